// Tangent-linear and adjoint kernels of solution order 3.
#include "adjoint_kernels.hpp"
GOCFD_INSTANTIATE_ADJOINT(3)
