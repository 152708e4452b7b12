// Stage, prologue and shock-capturing kernels of solution order 0.
#include "stage_kernel.hpp"
GOCFD_INSTANTIATE_ORDER(0)
