// Order-independent kernels of the tangent-linear and adjoint step: the
// transposed SSP54 combinations, the wave-speed arg-max and dt, the
// far-state cotangent sum, dual packing; and the order dispatch.
#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "dual.hpp"

namespace gocfd {
namespace dev {

using ad::Dual;

__device__ __forceinline__ double val_of(double a) { return a; }
__device__ __forceinline__ double val_of(const Dual &a) { return a.v; }

// Partial arg-max of the wave speeds: the largest value, ties to the
// smallest index (a deterministic choice of the point the dt follows)
template <class T>
__global__ void __launch_bounds__(256) k_adj_argmax_part(const T *SP, size_t n, double *pv, int *pi) {
  __shared__ double sv[256];
  __shared__ int si[256];
  double bv = -2.;
  int bi = -1;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double v = val_of(SP[i]);
    if (v > bv) { bv = v; bi = int(i); }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      const double ov = sv[threadIdx.x + w];
      const int oi = si[threadIdx.x + w];
      if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi >= 0 && (si[threadIdx.x] < 0 || oi < si[threadIdx.x]))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

// The arg-max over the partials in one block (ties to the smallest index)
__device__ int argmax_of_parts(const double *pv, const int *pi, int nparts, double *sv, int *si) {
  double bv = -2.;
  int bi = -1;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    if (pv[i] > bv || (pv[i] == bv && pi[i] >= 0 && (bi < 0 || pi[i] < bi))) { bv = pv[i]; bi = pi[i]; }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      const double ov = sv[threadIdx.x + w];
      const int oi = si[threadIdx.x + w];
      if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi >= 0 && (si[threadIdx.x] < 0 || oi < si[threadIdx.x]))) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const int r = si[0];
  __syncthreads();
  return r;
}

// Final arg-max over the partials; dt = CFL / W (in T), with an active
// shock sensor the smaller of that and Cdiff / Wvisc (the maximum viscous
// speed; calculateGlobalDT euler.go:945-971), the FinalTime check of the
// solver's clamp, and the time advance after the last stage
template <class T>
__global__ void __launch_bounds__(256) k_adj_argmax_final(const T *SP, const double *pv, const int *pi, int nparts,
                                                          int stage, double cfl, double final_time, T *dts,
                                                          int *argmax, double *Wv, double *time, int *flag,
                                                          const T *SPV, const double *pv2, const int *pi2,
                                                          double cdiff, int *argmaxv, double *Wvv, int *vbr) {
  __shared__ double sv[256];
  __shared__ int si[256];
  const int idx = argmax_of_parts(pv, pi, nparts, sv, si);
  const int idv = SPV ? argmax_of_parts(pv2, pi2, nparts, sv, si) : -1;
  if (threadIdx.x == 0) {
    argmax[stage] = idx;
    if (SPV) {
      argmaxv[stage] = idv;
      Wvv[stage] = idv >= 0 ? val_of(SPV[idv]) : 0.;
      vbr[stage] = 0;
    }
    if (idx < 0) {
      *flag |= 2;  // no wave face
      dts[stage] = T(0.);
      Wv[stage] = 0.;
      return;
    }
    const T W = SP[idx];
    T dt = cfl / W;
    if (idv >= 0 && val_of(SPV[idv]) > 0.) {
      const T dtv = cdiff / SPV[idv];
      if (val_of(dtv) < val_of(dt)) {
        dt = dtv;
        vbr[stage] = 1;
      }
    }
    dts[stage] = dt;
    Wv[stage] = val_of(W);
    if (time) {
      if (*time + val_of(dt) > final_time) *flag |= 1;
      if (stage == 4) *time += val_of(dt);
    }
  }
}

// Transpose of one SSP54 stage. lam is the cotangent of the stage's output;
// it is read before lamU0 is written, so that the last stage may take the
// step result's cotangent in the lamU0 buffer itself. Writes the cotangent
// of the stage's RHS (lamR) and adds to those of the stage inputs; block
// partials of the dt cotangent's dot product go to part.
__global__ void __launch_bounds__(256) k_adj_ssp(int stage, size_t n, const double *lam, double *lamU0,
                                                 double *lamQ1, double *lamQ2, double *lamQ3, double *lamQ4,
                                                 double *lamR, double *lamR3, const double *Rs, const double *R3,
                                                 const double *dts, double *part) {
  __shared__ double red[256];
  const double dt = dts[stage];
  double dot = 0.;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
    const double l = lam[i];
    switch (stage) {
      case 4:
        lamU0[i] = 0.;
        lamQ2[i] = kSspD2 * l;
        lamQ3[i] = kSspD3 * l;
        lamQ4[i] = kSspD4 * l;
        lamR3[i] = kSspE3 * dt * l;
        lamR[i] = kSspE4 * dt * l;
        dot += l * (kSspE3 * R3[i] + kSspE4 * Rs[i]);
        break;
      case 3:
        lamU0[i] += kSspB3 * l;
        lamQ3[i] += kSspC3 * l;
        lamR[i] = lamR3[i] + kSspA3 * dt * l;
        dot += kSspA3 * l * Rs[i];
        break;
      case 2:
        lamU0[i] += kSspB2 * l;
        lamQ2[i] += kSspC2 * l;
        lamR[i] = kSspA2 * dt * l;
        dot += kSspA2 * l * Rs[i];
        break;
      case 1:
        lamU0[i] += kSspB1 * l;
        lamQ1[i] = kSspC1 * l;
        lamR[i] = kSspA1 * dt * l;
        dot += kSspA1 * l * Rs[i];
        break;
      default:
        lamU0[i] += l;
        lamR[i] = kSspA0 * dt * l;
        dot += kSspA0 * l * Rs[i];
    }
  }
  red[threadIdx.x] = dot;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Sum of the block partials into out[slot] (one block, fixed order)
__global__ void __launch_bounds__(256) k_adj_sum(const double *part, int nparts, double *out, int slot) {
  __shared__ double red[256];
  double s = 0.;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[slot] = red[0];
}

// Far-state cotangent: the face-point records summed per component (block
// partials, then one block in a fixed order), plus the dt terms' part
__global__ void __launch_bounds__(256) k_adj_far_part(const double *pts, size_t npts, double *part) {
  __shared__ double red[4][256];
  double s[4] = {0., 0., 0., 0.};
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < npts; i += size_t(gridDim.x) * blockDim.x)
    for (int c = 0; c < 4; ++c) s[c] += pts[4 * i + c];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) part[4 * blockIdx.x + c] = red[c][0];
}

__global__ void __launch_bounds__(256) k_adj_far_final(const double *part, int nparts, const double *farW,
                                                       double *out) {
  __shared__ double red[4][256];
  double s[4] = {0., 0., 0., 0.};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    for (int c = 0; c < 4; ++c) s[c] += part[4 * i + c];
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = s[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) out[c] = red[c][0] + farW[c];
}

__global__ void __launch_bounds__(256) k_adj_pack(size_t n, const double *v, const double *d, Dual *out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = Dual(v[i], d ? d[i] : 0.);
}

__global__ void __launch_bounds__(256) k_adj_unpack(size_t n, const Dual *in, double *v, double *d) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  v[i] = in[i].v;
  d[i] = in[i].d;
}

static unsigned blocks_of(size_t n) { return unsigned((n + 255) / 256); }

hipError_t adj_launch_common(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s) {
  switch (op) {
    case ADJ_ARGMAX: {
      const size_t nfp = size_t(3) * a.Kp * (a.N + 2);
      if (b.dual) {
        hipLaunchKernelGGL((k_adj_argmax_part<Dual>), dim3(b.nparts), dim3(256), 0, s,
                           static_cast<const Dual *>(b.SP), nfp, b.part_v, b.part_i);
        if (a.act)
          hipLaunchKernelGGL((k_adj_argmax_part<Dual>), dim3(b.nparts), dim3(256), 0, s,
                             static_cast<const Dual *>(b.SPV), nfp, b.part_v2, b.part_i2);
        hipLaunchKernelGGL((k_adj_argmax_final<Dual>), dim3(1), dim3(256), 0, s, static_cast<const Dual *>(b.SP),
                           b.part_v, b.part_i, b.nparts, b.stage, a.ph.cfl, a.ph.final_time,
                           static_cast<Dual *>(b.dts), b.argmax, b.Wv, b.time, b.flag,
                           a.act ? static_cast<const Dual *>(b.SPV) : nullptr, b.part_v2, b.part_i2, a.cdiff,
                           b.argmaxv, b.Wvv, b.vbr);
      } else {
        hipLaunchKernelGGL((k_adj_argmax_part<double>), dim3(b.nparts), dim3(256), 0, s,
                           static_cast<const double *>(b.SP), nfp, b.part_v, b.part_i);
        if (a.act)
          hipLaunchKernelGGL((k_adj_argmax_part<double>), dim3(b.nparts), dim3(256), 0, s,
                             static_cast<const double *>(b.SPV), nfp, b.part_v2, b.part_i2);
        hipLaunchKernelGGL((k_adj_argmax_final<double>), dim3(1), dim3(256), 0, s,
                           static_cast<const double *>(b.SP), b.part_v, b.part_i, b.nparts, b.stage, a.ph.cfl,
                           a.ph.final_time, static_cast<double *>(b.dts), b.argmax, b.Wv, b.time, b.flag,
                           a.act ? static_cast<const double *>(b.SPV) : nullptr, b.part_v2, b.part_i2, a.cdiff,
                           b.argmaxv, b.Wvv, b.vbr);
      }
      break;
    }
    case ADJ_ADJ_SSP:
      // nparts blocks stride over the element array; one partial each
      hipLaunchKernelGGL(k_adj_ssp, dim3(b.nparts), dim3(256), 0, s, b.stage, b.n, b.lam, b.lamU0, b.lamQ[1],
                         b.lamQ[2], b.lamQ[3], b.lamQ[4], b.lamR, b.lamR3, static_cast<const double *>(b.Rs),
                         static_cast<const double *>(b.R3), static_cast<const double *>(b.dts), b.part_v);
      hipLaunchKernelGGL(k_adj_sum, dim3(1), dim3(256), 0, s, b.part_v, b.nparts, b.lam_dt, b.stage);
      break;
    case ADJ_FAR_SUM:
      // 4 partials per block; the result goes to b.target[0..3]
      hipLaunchKernelGGL(k_adj_far_part, dim3(b.nparts), dim3(256), 0, s, b.lamFarPts,
                         size_t(3) * a.Kp * (a.N + 2), b.part_v);
      hipLaunchKernelGGL(k_adj_far_final, dim3(1), dim3(256), 0, s, b.part_v, b.nparts, b.lamFarW, b.target);
      break;
    case ADJ_PACK:
      hipLaunchKernelGGL(k_adj_pack, dim3(blocks_of(b.n)), dim3(256), 0, s, b.n, b.pack_v, b.pack_d,
                         static_cast<Dual *>(b.out));
      break;
    case ADJ_UNPACK:
      hipLaunchKernelGGL(k_adj_unpack, dim3(blocks_of(b.n)), dim3(256), 0, s, b.n, static_cast<const Dual *>(b.Us),
                         b.unpack_v, b.unpack_d);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t adj_launch(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s) {
  if (op == ADJ_ARGMAX || op == ADJ_ADJ_SSP || op == ADJ_FAR_SUM || op == ADJ_PACK || op == ADJ_UNPACK)
    return adj_launch_common(op, a, b, s);
  switch (a.N) {
    case 0: return adj_launch_n<0>(op, a, b, s);
    case 1: return adj_launch_n<1>(op, a, b, s);
    case 2: return adj_launch_n<2>(op, a, b, s);
    case 3: return adj_launch_n<3>(op, a, b, s);
    case 4: return adj_launch_n<4>(op, a, b, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace dev
}  // namespace gocfd
