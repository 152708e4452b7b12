// Tangent-linear and adjoint kernels of solution order 1.
#include "adjoint_kernels.hpp"
GOCFD_INSTANTIATE_ADJOINT(1)
