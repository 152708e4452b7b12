// Order dispatch of the stage kernels, and the small control kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace gocfd {
namespace dev {

#define GOCFD_DISPATCH_ORDER(N, CALL) \
  switch (N) {                        \
    case 0: return CALL(0);           \
    case 1: return CALL(1);           \
    case 2: return CALL(2);           \
    case 3: return CALL(3);           \
    case 4: return CALL(4);           \
    default: return hipErrorInvalidValue; \
  }

hipError_t launch_stage(int N, int prec, int shock, const StageArgs &a, hipStream_t s) {
#define CALL(n) launch_stage_n<n>(prec, shock, a, s)
  GOCFD_DISPATCH_ORDER(N, CALL)
#undef CALL
}

hipError_t launch_prologue(int N, int prec, const StageArgs &a, hipStream_t s) {
#define CALL(n) launch_prologue_n<n>(prec, a, s)
  GOCFD_DISPATCH_ORDER(N, CALL)
#undef CALL
}

hipError_t launch_halo_export(int N, int prec, const StageArgs &a, hipStream_t s) {
#define CALL(n) launch_halo_export_n<n>(prec, a, s)
  GOCFD_DISPATCH_ORDER(N, CALL)
#undef CALL
}

hipError_t launch_shock(int N, int prec, int pass, const StageArgs &a, hipStream_t s) {
#define CALL(n) launch_shock_n<n>(prec, pass, a, s)
  GOCFD_DISPATCH_ORDER(N, CALL)
#undef CALL
}

hipError_t launch_stage_multi(int prec, int flux, const MultiArgs &m, hipStream_t s) {
  switch (flux) {
    case 0: return launch_stage_multi_f<0>(prec, m, s);
    case 1: return launch_stage_multi_f<1>(prec, m, s);
    case 2: return launch_stage_multi_f<2>(prec, m, s);
    case 3: return launch_stage_multi_f<3>(prec, m, s);
  }
  return hipErrorInvalidValue;
}

hipError_t upload_operators_multi(int flux, int N, const double *divT, const double *interp) {
  switch (flux) {
    case 0: return upload_operators_multi_f<0>(N, divT, interp);
    case 1: return upload_operators_multi_f<1>(N, divT, interp);
    case 2: return upload_operators_multi_f<2>(N, divT, interp);
    case 3: return upload_operators_multi_f<3>(N, divT, interp);
  }
  return hipErrorInvalidValue;
}

hipError_t upload_operators(int N, const double *divT, const double *interp) {
#define CALL(n) upload_operators_n<n>(divT, interp)
  GOCFD_DISPATCH_ORDER(N, CALL)
#undef CALL
}

// Dynamic LDS of the stage kernel, at most: the tile state [NV][kTile+1],
// the trace/flux buffer [4][3 NpEdge kTP + NpEdge Hmax] (the halo part
// only with shock capturing), the Riemann items' geometry and the metric.
size_t stage_lds_bytes(int N, int prec, int Hmax) {
  if (N < 0 || N > 4) return 0;
  const size_t cs = prec == 0 ? sizeof(double) : sizeof(float);
  const size_t NpInt = size_t(N + 1) * (N + 2) / 2, NpEdge = size_t(N) + 2;
  return (cs * (4 * NpInt * (kTile + 1) + 4 * (3 * NpEdge * kTP + NpEdge * size_t(Hmax))) + 31) / 32 * 32 +
         cs * (4 * 3 * kTile + 6 * kTile) +                                  // item geometry (Fmax <= 3 kTile), metric
         sizeof(double) * kTile + 2 * sizeof(int) * 6 * kTile +              // phase G lists (Wmax, Xmax <= 3 kTile)
         sizeof(double) * kTile + sizeof(unsigned long long) * 3 * kTile;  // with the LTS buffers
}

__global__ void k_ctl_init(Ctl *ctls, double time, int ncases) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncases) return;
  Ctl *ctl = ctls + c;
  ctl->time[0] = time;
  ctl->time[1] = time;
  for (int i = 0; i < 5; ++i) ctl->dt[i] = 0.0;
  for (int i = 0; i < 3; ++i) {
    ctl->wmax[i] = 0ull;
    ctl->wmaxv[i] = 0ull;
  }
  ctl->nan = 0;
  ctl->steps = 0;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

hipError_t launch_ctl_init(Ctl *ctl, double time, hipStream_t s, int ncases) {
  hipLaunchKernelGGL(k_ctl_init, dim3((ncases + 63) / 64), dim3(64), 0, s, ctl, time, ncases);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int blocks = int(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int blocks = int(std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace gocfd
