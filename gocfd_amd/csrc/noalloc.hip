// RHS of a device-resident state without any device allocation
// (gocfd_euler2d_rhs_noalloc, include/gocfd_euler2d.h), for callers that
// cannot allocate: code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc. The element-major staging of the input state and of
// the RHS lives in one file-scope device array, `scratch`, of fixed size;
// each device that loads the library holds its own copy until the process
// ends, and every call on that device shares it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "euler2d_handle.hpp"

namespace {

// 32 MiB per device: the input state and the RHS, Kp * 4 NpInt doubles each
// (K up to about 34,900 triangles at N=4, 87,300 at N=2)
constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr size_t kScratchVals = kScratchBytes / sizeof(double);

__device__ __align__(256) double scratch[kScratchVals];

// Device address of `scratch` on each device, resolved once per device with
// that device current. The array's host-side name is not a device address:
// host code uses these pointers only.
constexpr int kMaxDevices = 64;
double *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Reference layout [4][NpInt][K] -> element-major [Kp][4 NpInt]; padding
// elements (perm = -1) take the inert state {1, 0, 0, 2.5}, as the host-side
// re-layout gives them. Every value of the output is written.
__global__ void __launch_bounds__(256) k_gather_state(const double *__restrict__ Q, const int *__restrict__ perm,
                                                      int K, int Kp, int NpInt, double *__restrict__ out) {
  const int NV = 4 * NpInt;
  const size_t n = size_t(Kp) * NV;
  for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n; idx += size_t(gridDim.x) * blockDim.x) {
    const int kk = int(idx / NV), r = int(idx - size_t(kk) * NV);
    const int k = perm[kk];
    double v;
    if (k < 0) {
      const int var = r / NpInt;
      v = var == 0 ? 1.0 : var == 3 ? 2.5 : 0.0;
    } else {
      v = Q[size_t(r) * K + k];
    }
    out[idx] = v;
  }
}

// Element-major RHS of the K real elements -> reference layout
__global__ void __launch_bounds__(256) k_scatter_rhs(const double *__restrict__ in, const int *__restrict__ perm,
                                                     int K, int NpInt, double *__restrict__ RHS) {
  const int NV = 4 * NpInt;
  const size_t n = size_t(K) * NV;
  for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < n; idx += size_t(gridDim.x) * blockDim.x) {
    const int kk = int(idx / NV), r = int(idx - size_t(kk) * NV);
    RHS[size_t(r) * K + perm[kk]] = in[idx];
  }
}

int grid_for(size_t n) { return int(std::min<size_t>((n + 255) / 256, 8192)); }

}  // namespace

namespace gocfd {
namespace capi {

void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw SetupError(GOCFD_ERR_DEVICE, "device ordinal " + std::to_string(device) + " beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return;
  int cur = -1;
  HIPCHECK(hipGetDevice(&cur));
  if (cur != device) throw SetupError(GOCFD_ERR_DEVICE, "scratch set-up needs its device current");
  void *p = nullptr;
  HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw SetupError(GOCFD_ERR_DEVICE, "hipGetSymbolAddress gave no address for the scratch array");
  g_scratch[device] = static_cast<double *>(p);
}

}  // namespace capi
}  // namespace gocfd

extern "C" int gocfd_euler2d_rhs_noalloc(gocfd_euler2d *h, const double *Q, double t, double *RHS, void *stream) {
  GUARD_BEGIN
  if (!h || !Q || !RHS) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  const size_t vals = size_t(h->L.Kp) * h->NV;
  if (2 * vals > kScratchVals)
    return fail(GOCFD_ERR_ARG, "rhs_noalloc: this mesh and order need " + std::to_string(2 * vals * sizeof(double)) +
                                   " bytes of scratch; the maximum is " + std::to_string(kScratchBytes) +
                                   " (use gocfd_euler2d_rhs)");
  HIPCHECK(hipSetDevice(h->device));
  noalloc_setup(h->device);  // done at create; a no-op here
  double *in = g_scratch[h->device], *out = in + vals;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  const int NpInt = h->d.ops.NpInt;
  hipLaunchKernelGGL(k_gather_state, dim3(grid_for(vals)), dim3(256), 0, s, Q, h->perm.p, h->L.K, h->L.Kp, NpInt,
                     in);
  HIPCHECK(hipGetLastError());
  dev::StageArgs a = h->rhs_args(in, t, out, s);
  h->launch_one_stage(a, s);
  if (h->shock) {  // the vertex buffers this evaluation used must be clear for the next stage
    HIPCHECK(dev::launch_fill_u64(h->sigV.p, h->sigV.n, 0ull, s));
    HIPCHECK(dev::launch_fill_u64(h->epsV.p, h->epsV.n, 0ull, s));
  }
  const size_t nout = size_t(h->L.K) * h->NV;
  hipLaunchKernelGGL(k_scatter_rhs, dim3(grid_for(nout)), dim3(256), 0, s, out, h->perm.p, h->L.K, NpInt, RHS);
  HIPCHECK(hipGetLastError());
  if (!stream) HIPCHECK(hipStreamSynchronize(s));
  return GOCFD_OK;
  GUARD_END
}
