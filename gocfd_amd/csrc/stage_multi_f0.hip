// Multi-order stage kernels of flux 0 (all non-shock order groups of an ensemble in one launch).
#define GOCFD_MULTI_FLUX 0
#include "stage_kernel.hpp"
