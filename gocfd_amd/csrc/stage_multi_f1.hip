// Multi-order stage kernels of flux 1 (all non-shock order groups of an ensemble in one launch).
#define GOCFD_MULTI_FLUX 1
#include "stage_kernel.hpp"
