// Order-templated tangent-linear and adjoint kernels of the FP64 SSP54 step
// (interface and scope in adjoint.hpp). Each adjoint_n<N>.hip translation
// unit includes this header and instantiates one order.
//
// Forward kernels (arithmetic type T = double or ad::Dual), one thread per
// face point or per element, on the reference's formulation:
//   face:  the edge RT DOF value of a face point: the numerical flux against
//          the neighbour's trace at the reversed point (owner orientation),
//          or the boundary flux, times the side's ||n|| and sign
//          (CalculateEdgeEulerFlux edges.go:325-452, SetRTFluxOnEdges
//          edges.go:454-483)
//   rhs:   -(1/Jdet) DivInt [Jdet Jinv F(Q) ; edge values] (RHSInternalPoints
//          euler.go:665-699, SetRTFluxInternal euler.go:701-726)
//   wave:  (|U| + c) / hK on the owner side's post-BC trace (StoreEdgeAggregates
//          edges.go:246-289)
// Adjoint kernels: the transposes of the same maps at a recomputed stage
// input, with the pointwise Jacobians formed in dual arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include "adjoint.hpp"
#include "dfr_setup.hpp"
#include "dual.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace gocfd {
namespace dev {

using ad::Dual;

template <int N>
struct AdjShape {
  static constexpr int NpInt = (N + 1) * (N + 2) / 2;
  static constexpr int NpEdge = N + 2;
  static constexpr int NpFlux = (N + 2) * (N + 4);
  static constexpr int NF = 3 * NpEdge;
  static constexpr int NV = 4 * NpInt;
};

// A value whose primal is `v` and whose tangent is that of `t`
__device__ __forceinline__ double with_tangent(double v, double) { return v; }
__device__ __forceinline__ Dual with_tangent(double v, Dual t) { return Dual(v, t.d); }

template <class T>
__device__ __forceinline__ void far_state(const AdjArgs &a, T q[4]);
template <>
__device__ __forceinline__ void far_state<double>(const AdjArgs &a, double q[4]) {
  for (int c = 0; c < 4; ++c) q[c] = a.ph.qinf[c];
}
template <>
__device__ __forceinline__ void far_state<Dual>(const AdjArgs &a, Dual q[4]) {
  for (int c = 0; c < 4; ++c) q[c] = Dual(a.ph.qinf[c], a.dfar[c]);
}

// Riemann-invariant state of a far-field face point (edges.go:425-431):
// FSFar, or FSIn / FSOut. pinf and cinf follow the far state: their values
// are the solver's, their tangents those of the state's pressure and sound
// speed.
template <class T>
__device__ __forceinline__ void far_bc(const AdjArgs &a, int bidx, const T qfar[4], const T own[4], double nx,
                                       double ny, T qb[4]) {
  const Phys &ph = a.ph;
  if (bidx == 0 || a.io_tied) {
    const T g = T(ph.gamma);
    const T pinf = with_tangent(ph.pinf, pressure(g, qfar[0], qfar[1], qfar[2], qfar[3]));
    const T cinf = with_tangent(ph.cinf, sound_speed(g, qfar));
    riemann_bc(ph.gamma, ph.minf, pinf, cinf, own, qfar, nx, ny, qb);
  } else {
    const FarState &fs = ph.io[bidx - 1];
    const T qf[4] = {T(fs.q[0]), T(fs.q[1]), T(fs.q[2]), T(fs.q[3])};
    riemann_bc(fs.gamma, fs.minf, fs.pinf, fs.cinf, own, qf, nx, ny, qb);
  }
}

// Edge RT DOF values of face point (fi = f Kp + kk, pp) from its own trace,
// the neighbour's trace at the reversed point and the far state
template <class T>
__device__ __forceinline__ void face_value(const AdjArgs &a, int code, size_t fi, const T own[4], const T oth[4],
                                           const T qfar[4], T F[4]) {
  const int kind = code & 15;
  const double nxd = a.Nx[fi], nyd = a.Ny[fi], w = a.IInII[fi];
  const T g = T(a.ph.gamma), nx = T(nxd), ny = T(nyd);
  for (int c = 0; c < 4; ++c) F[c] = T(0.);
  if (kind == FK_Shared) {
    const bool owner = code & FC_FLUX_OWNER;
    if (owner) numflux(a.ph.flux, g, own, oth, nx, ny, F);
    else numflux(a.ph.flux, g, oth, own, nx, ny, F);
    for (int c = 0; c < 4; ++c) F[c] = owner ? F[c] * w : -F[c] * w;
    return;
  }
  if (kind == FK_Far) {
    T qb[4];
    far_bc(a, code >> 8, qfar, own, nxd, nyd, qb);
    T Fx[4], Fy[4];
    flux_xy(g, qb, Fx, Fy);
    for (int c = 0; c < 4; ++c) F[c] = nx * Fx[c] + ny * Fy[c];
  } else if (kind == FK_Wall) {
    const T pr = pressure(g, own[0], own[1], own[2], own[3]);
    F[1] = nx * pr;
    F[2] = ny * pr;
  } else if (kind == FK_Open) {
    T Fx[4], Fy[4];
    flux_xy(g, own, Fx, Fy);
    for (int c = 0; c < 4; ++c) F[c] = nx * Fx[c] + ny * Fy[c];
  }
  for (int c = 0; c < 4; ++c) F[c] = F[c] * w;
}

// Wave speed (|U| + c) / hK of a face point on the owner's post-BC trace
template <class T>
__device__ __forceinline__ T wave_value(const AdjArgs &a, int code, size_t fi, double oohK, const T own[4],
                                        const T qfar[4]) {
  T q[4] = {own[0], own[1], own[2], own[3]};
  if ((code & 15) == FK_Far) far_bc(a, code >> 8, qfar, own, a.Nx[fi], a.Ny[fi], q);
  const T oorho = T(1.) / q[0], u = q[1] * oorho, v = q[2] * oorho;
  const T U = sqrt(u * u + v * v);
  const T c = sound_speed(T(a.ph.gamma), q);
  return oohK * (U + c);
}

template <int N, class T>
__device__ __forceinline__ void trace_at(const AdjArgs &a, const T *U, int kk, int p, T q[4]) {
  using Sh = AdjShape<N>;
  const T *u = U + size_t(kk) * Sh::NV;
  const double *row = a.Interp + p * Sh::NpInt;
  for (int c = 0; c < 4; ++c) {
    T t = T(0.);
    for (int i = 0; i < Sh::NpInt; ++i) t += row[i] * u[c * Sh::NpInt + i];
    q[c] = t;
  }
}

// (element kk, face f, point pp) of face-point index t = (f Kp + kk) NpEdge + pp
template <int N>
__device__ __forceinline__ void face_point(const AdjArgs &a, size_t t, int &f, int &kk, int &pp, size_t &fi) {
  constexpr int NpEdge = AdjShape<N>::NpEdge;
  pp = int(t % NpEdge);
  fi = t / NpEdge;
  f = int(fi / a.Kp);
  kk = int(fi - size_t(f) * a.Kp);
}

template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_face(AdjArgs a, const T *U, T *FE) {
  using Sh = AdjShape<N>;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * Sh::NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  T F[4] = {T(0.), T(0.), T(0.), T(0.)};
  if (kk < a.K) {
    const int code = a.code[fi];
    T own[4], oth[4] = {T(0.), T(0.), T(0.), T(0.)}, qfar[4];
    trace_at<N>(a, U, kk, f * Sh::NpEdge + pp, own);
    if ((code & 15) == FK_Shared)
      trace_at<N>(a, U, a.nbr_elem[fi], a.nbr_face[fi] * Sh::NpEdge + (Sh::NpEdge - 1 - pp), oth);
    far_state(a, qfar);
    face_value(a, code, fi, own, oth, qfar, F);
  }
  for (int c = 0; c < 4; ++c) FE[4 * t + c] = F[c];
}

template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_rhs(AdjArgs a, const T *U, const T *FE, T *R) {
  using Sh = AdjShape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.Kp) return;
  T *r = R + size_t(kk) * Sh::NV;
  if (kk >= a.K) {
    for (int i = 0; i < Sh::NV; ++i) r[i] = T(0.);
    return;
  }
  const T *u = U + size_t(kk) * Sh::NV;
  const double Jd = a.Jdet[kk];
  const double J0 = a.Jinv[kk], J1 = a.Jinv[size_t(a.Kp) + kk];
  const double J2 = a.Jinv[2 * size_t(a.Kp) + kk], J3 = a.Jinv[3 * size_t(a.Kp) + kk];
  const double *D = a.DivInt;
  T acc[4][NpInt];
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < NpInt; ++i) acc[c][i] = T(0.);
  for (int j = 0; j < NpInt; ++j) {
    const T q[4] = {u[j], u[NpInt + j], u[2 * NpInt + j], u[3 * NpInt + j]};
    T Fx[4], Fy[4];
    flux_xy(T(a.ph.gamma), q, Fx, Fy);
    for (int c = 0; c < 4; ++c) {
      const T Fr = Jd * (J0 * Fx[c] + J1 * Fy[c]), Fs = Jd * (J2 * Fx[c] + J3 * Fy[c]);
      for (int i = 0; i < NpInt; ++i) acc[c][i] += D[i * NpFlux + j] * Fr + D[i * NpFlux + NpInt + j] * Fs;
    }
  }
  for (int p = 0; p < Sh::NF; ++p) {
    const int f = p / NpEdge, pp = p - f * NpEdge;
    const T *fe = FE + 4 * ((size_t(f) * a.Kp + kk) * NpEdge + pp);
    for (int c = 0; c < 4; ++c)
      for (int i = 0; i < NpInt; ++i) acc[c][i] += D[i * NpFlux + 2 * NpInt + p] * fe[c];
  }
  const double m = -(1. / Jd);
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < NpInt; ++i) r[c * NpInt + i] = acc[c][i] * m;
}

template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_wave(AdjArgs a, const T *U, T *SP) {
  using Sh = AdjShape<N>;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * Sh::NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  T s = T(-1.);
  if (kk < a.K) {
    const int code = a.code[fi];
    if (code & FC_WAVE) {
      T own[4], qfar[4];
      trace_at<N>(a, U, kk, f * Sh::NpEdge + pp, own);
      far_state(a, qfar);
      s = wave_value(a, code, fi, a.oohK[kk], own, qfar);
    }
  }
  SP[t] = s;
}

// ---- adjoint ----

// Cotangent of the edge values of face point (f, pp) of element kk:
// -(1/Jdet) sum_i DivInt[i][2 NpInt + p] lamR[c][i]
template <int N>
__device__ __forceinline__ void edge_cotangent(const AdjArgs &a, const double *lamR, int kk, int p, double lf[4]) {
  using Sh = AdjShape<N>;
  const double *lr = lamR + size_t(kk) * Sh::NV;
  const double m = -(1. / a.Jdet[kk]);
  for (int c = 0; c < 4; ++c) {
    double t = 0.;
    for (int i = 0; i < Sh::NpInt; ++i) t += a.DivInt[i * Sh::NpFlux + 2 * Sh::NpInt + p] * lr[c * Sh::NpInt + i];
    lf[c] = t * m;
  }
}

// Per face point: the transposed Jacobians of its edge value with respect
// to the own trace, the neighbour's trace and the far state, applied to the
// edge value's cotangent.
template <int N>
__global__ void __launch_bounds__(256) k_adj_face(AdjArgs a, const double *U, const double *lamR, double *lamOwn,
                                                  double *lamOth, double *lamFarPts) {
  using Sh = AdjShape<N>;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * Sh::NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  double *lo = lamOwn + 4 * t, *lt = lamOth + 4 * t, *lq = lamFarPts + 4 * t;
  for (int c = 0; c < 4; ++c) {
    lo[c] = 0.;
    lt[c] = 0.;
  }
  if (kk >= a.K) return;
  const int code = a.code[fi], kind = code & 15;
  double lf[4];
  edge_cotangent<N>(a, lamR, kk, f * Sh::NpEdge + pp, lf);
  double own[4], oth[4] = {0., 0., 0., 0.}, qfar[4];
  trace_at<N>(a, U, kk, f * Sh::NpEdge + pp, own);
  if (kind == FK_Shared)
    trace_at<N>(a, U, a.nbr_elem[fi], a.nbr_face[fi] * Sh::NpEdge + (Sh::NpEdge - 1 - pp), oth);
  far_state(a, qfar);
  const bool far = kind == FK_Far && ((code >> 8) == 0 || a.io_tied);
  // input directions j = 4 set + component; sets: 0 the own trace, 1 the
  // neighbour's trace, 2 the far state (its cotangent accumulates)
  for (int j = 0; j < 12; ++j) {
    const int which = j >> 2, d = j & 3;
    if ((which == 1 && kind != FK_Shared) || (which == 2 && !far)) continue;
    Dual o[4], n[4], q[4], F[4];
    for (int c = 0; c < 4; ++c) {
      o[c] = Dual(own[c], which == 0 && c == d ? 1. : 0.);
      n[c] = Dual(oth[c], which == 1 && c == d ? 1. : 0.);
      q[c] = Dual(qfar[c], which == 2 && c == d ? 1. : 0.);
    }
    face_value(a, code, fi, o, n, q, F);
    double sum = 0.;
    for (int c = 0; c < 4; ++c) sum += lf[c] * F[c].d;
    if (which == 0) lo[d] = sum;
    else if (which == 1) lt[d] = sum;
    else lq[d] += sum;
  }
}

// Per element: the interior-flux transpose (DivInt columns of the r and s
// components, the contravariant transform and the physical-flux Jacobian
// at each solution point) and the FluxEdgeInterp transpose of the face-point
// trace cotangents (own side, plus the neighbour's cotangent of this
// element's trace at the reversed point), added to target.
template <int N>
__global__ void __launch_bounds__(256) k_adj_elem(AdjArgs a, const double *U, const double *lamR,
                                                  const double *lamOwn, const double *lamOth, double *target) {
  using Sh = AdjShape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.K) return;
  const double *u = U + size_t(kk) * Sh::NV;
  const double *lr = lamR + size_t(kk) * Sh::NV;
  const double Jd = a.Jdet[kk], m = -(1. / Jd);
  const double J0 = a.Jinv[kk], J1 = a.Jinv[size_t(a.Kp) + kk];
  const double J2 = a.Jinv[2 * size_t(a.Kp) + kk], J3 = a.Jinv[3 * size_t(a.Kp) + kk];
  const double *D = a.DivInt;
  double *tg = target + size_t(kk) * Sh::NV;
  // face points: FluxEdgeInterp^T of the trace cotangents
  double lu[4][NpInt];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < NpInt; ++i) lu[c][i] = 0.;
  for (int p = 0; p < Sh::NF; ++p) {
    const int f = p / NpEdge, pp = p - f * NpEdge;
    const size_t fi = size_t(f) * a.Kp + kk;
    const size_t t = fi * NpEdge + pp;
    double lq[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) lq[c] = lamOwn[4 * t + c];
    if ((a.code[fi] & 15) == FK_Shared) {
      const size_t t2 = (size_t(a.nbr_face[fi]) * a.Kp + a.nbr_elem[fi]) * NpEdge + (NpEdge - 1 - pp);
#pragma unroll
      for (int c = 0; c < 4; ++c) lq[c] += lamOth[4 * t2 + c];
    }
    const double *row = a.Interp + p * NpInt;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int i = 0; i < NpInt; ++i) lu[c][i] += row[i] * lq[c];
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int i = 0; i < NpInt; ++i) tg[c * NpInt + i] += lu[c][i];
  // solution points: DivInt^T (r and s columns), the contravariant
  // transform and the physical-flux Jacobian
  for (int j = 0; j < NpInt; ++j) {
    double lFx[4], lFy[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double lr_ = 0., ls_ = 0.;
      for (int i = 0; i < NpInt; ++i) {
        lr_ += D[i * NpFlux + j] * lr[c * NpInt + i];
        ls_ += D[i * NpFlux + NpInt + j] * lr[c * NpInt + i];
      }
      lr_ *= m;
      ls_ *= m;
      lFx[c] = Jd * (J0 * lr_ + J2 * ls_);
      lFy[c] = Jd * (J1 * lr_ + J3 * ls_);
    }
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      Dual q[4], Fx[4], Fy[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) q[c] = Dual(u[c * NpInt + j], c == d ? 1. : 0.);
      flux_xy(Dual(a.ph.gamma), q, Fx, Fy);
      double sum = 0.;
#pragma unroll
      for (int c = 0; c < 4; ++c) sum += lFx[c] * Fx[c].d + lFy[c] * Fy[c].d;
      tg[d * NpInt + j] += sum;
    }
  }
}

// The dt term of a stage: dt = CFL / W with W the wave speed at the stage's
// arg-max face point; its cotangent goes to that point's trace (and the far
// state for a far-field face). One thread.
template <int N>
__global__ void k_adj_wave(AdjArgs a, int stage, const double *U, const int *argmax, const double *Wv,
                           const double *lam_dt, double *target, double *lamFarW) {
  using Sh = AdjShape<N>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int idx = argmax[stage];
  if (idx < 0) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, size_t(idx), f, kk, pp, fi);
  const double W = Wv[stage];
  const double lW = -(a.ph.cfl / (W * W)) * lam_dt[stage];
  const int code = a.code[fi];
  double own[4], qfar[4];
  trace_at<N>(a, U, kk, f * Sh::NpEdge + pp, own);
  far_state(a, qfar);
  const bool far = (code & 15) == FK_Far && ((code >> 8) == 0 || a.io_tied);
  double lo[4];
  for (int d = 0; d < 4; ++d) {
    Dual o[4], q[4];
    for (int c = 0; c < 4; ++c) {
      o[c] = Dual(own[c], c == d ? 1. : 0.);
      q[c] = Dual(qfar[c]);
    }
    lo[d] = lW * wave_value(a, code, fi, a.oohK[kk], o, q).d;
    if (far) {
      for (int c = 0; c < 4; ++c) {
        o[c] = Dual(own[c]);
        q[c] = Dual(qfar[c], c == d ? 1. : 0.);
      }
      lamFarW[d] += lW * wave_value(a, code, fi, a.oohK[kk], o, q).d;
    }
  }
  const double *row = a.Interp + (f * Sh::NpEdge + pp) * Sh::NpInt;
  double *tg = target + size_t(kk) * Sh::NV;
  for (int c = 0; c < 4; ++c)
    for (int i = 0; i < Sh::NpInt; ++i) tg[c * Sh::NpInt + i] += row[i] * lo[c];
}

inline unsigned adj_blocks(size_t n) { return unsigned((n + 255) / 256); }

template <int N>
hipError_t adj_launch_n(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s) {
  using Sh = AdjShape<N>;
  const size_t nfp = size_t(3) * a.Kp * Sh::NpEdge;
  switch (op) {
    case ADJ_FWD_FACE:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_fwd_face<N, Dual>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                           static_cast<const Dual *>(b.Us), static_cast<Dual *>(b.FE));
      else
        hipLaunchKernelGGL((k_adj_fwd_face<N, double>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                           static_cast<const double *>(b.Us), static_cast<double *>(b.FE));
      break;
    case ADJ_FWD_RHS:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_fwd_rhs<N, Dual>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a,
                           static_cast<const Dual *>(b.Us), static_cast<const Dual *>(b.FE),
                           static_cast<Dual *>(b.R));
      else
        hipLaunchKernelGGL((k_adj_fwd_rhs<N, double>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a,
                           static_cast<const double *>(b.Us), static_cast<const double *>(b.FE),
                           static_cast<double *>(b.R));
      break;
    case ADJ_FWD_WAVE:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_fwd_wave<N, Dual>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                           static_cast<const Dual *>(b.Us), static_cast<Dual *>(b.SP));
      else
        hipLaunchKernelGGL((k_adj_fwd_wave<N, double>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                           static_cast<const double *>(b.Us), static_cast<double *>(b.SP));
      break;
    case ADJ_ADJ_FACE:
      hipLaunchKernelGGL((k_adj_face<N>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamR, b.lamOwn, b.lamOth, b.lamFarPts);
      break;
    case ADJ_ADJ_ELEM:
      hipLaunchKernelGGL((k_adj_elem<N>), dim3(adj_blocks(a.K)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamR, b.lamOwn, b.lamOth, b.target);
      break;
    case ADJ_ADJ_WAVE:
      hipLaunchKernelGGL((k_adj_wave<N>), dim3(1), dim3(64), 0, s, a, b.stage, static_cast<const double *>(b.Us),
                         b.argmax, b.Wv, b.lam_dt, b.target, b.lamFarW);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace dev
}  // namespace gocfd

#define GOCFD_INSTANTIATE_ADJOINT(NN)                                                                \
  namespace gocfd {                                                                                  \
  namespace dev {                                                                                    \
  template hipError_t adj_launch_n<NN>(int, const AdjArgs &, const AdjBufs &, hipStream_t);          \
  }                                                                                                  \
  }
