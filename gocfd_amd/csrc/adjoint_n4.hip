// Tangent-linear and adjoint kernels of solution order 4.
#include "adjoint_kernels.hpp"
GOCFD_INSTANTIATE_ADJOINT(4)
