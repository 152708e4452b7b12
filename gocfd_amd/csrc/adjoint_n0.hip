// Tangent-linear and adjoint kernels of solution order 0.
#include "adjoint_kernels.hpp"
GOCFD_INSTANTIATE_ADJOINT(0)
