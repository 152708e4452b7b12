// Multi-order stage kernels of flux 3 (all non-shock order groups of an ensemble in one launch).
#define GOCFD_MULTI_FLUX 3
#include "stage_kernel.hpp"
