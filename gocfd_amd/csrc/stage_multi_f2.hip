// Multi-order stage kernels of flux 2 (all non-shock order groups of an ensemble in one launch).
#define GOCFD_MULTI_FLUX 2
#include "stage_kernel.hpp"
