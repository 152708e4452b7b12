// Stage, prologue and shock-capturing kernels of solution order 1.
#include "stage_kernel.hpp"
GOCFD_INSTANTIATE_ORDER(1)
