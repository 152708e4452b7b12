// Tangent-linear and adjoint kernels of solution order 2.
#include "adjoint_kernels.hpp"
GOCFD_INSTANTIATE_ADJOINT(2)
