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

// Edge RT DOF values of an unshared face point (fi = f Kp + kk) from its own
// trace and the far state: the boundary flux times ||n|| (zero for an
// uncoupled periodic face)
template <class T>
__device__ __forceinline__ void boundary_value(const AdjArgs &a, int code, size_t fi, const T own[4],
                                               const T qfar[4], T F[4]) {
  const int kind = code & 15;
  const double nxd = a.Nx[fi], nyd = a.Ny[fi], w = a.IInII[fi];
  const T g = T(a.ph.gamma), nx = T(nxd), ny = T(nyd);
  for (int c = 0; c < 4; ++c) F[c] = T(0.);
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

// A shared edge point's flux is evaluated once, by the owner's face point,
// which writes both sides' edge values: w_own F* and -w_nbr F* at the
// partner's reversed point (the value the reference stores once per edge
// and reads from both sides). The same pass stores each face point's wave
// speed (|U| + c) / hK on its post-BC trace where the face bounds the time
// step (the edge's owner side), -1 elsewhere.
template <int N, class T, int FLUX>
__global__ void __launch_bounds__(256) k_adj_fwd_face(AdjArgs a, const T *U, T *FE, T *SP) {
  using Sh = AdjShape<N>;
  constexpr int NpEdge = Sh::NpEdge;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  if (kk >= a.K) {
    for (int c = 0; c < 4; ++c) FE[4 * t + c] = T(0.);
    SP[t] = T(-1.);
    return;
  }
  const int code = a.code[fi], kind = code & 15;
  T own[4], qfar[4], F[4];
  trace_at<N>(a, U, kk, f * NpEdge + pp, own);
  far_state(a, qfar);
  SP[t] = (code & FC_WAVE) ? wave_value(a, code, fi, a.oohK[kk], own, qfar) : T(-1.);
  if (kind == FK_Shared) {
    if (!(code & FC_FLUX_OWNER)) return;  // written by the owner
    const int nb = a.nbr_elem[fi], nf = a.nbr_face[fi], rp = NpEdge - 1 - pp;
    T oth[4];
    trace_at<N>(a, U, nb, nf * NpEdge + rp, oth);
    numflux(FLUX, T(a.ph.gamma), own, oth, T(a.Nx[fi]), T(a.Ny[fi]), F);
    const size_t fn = size_t(nf) * a.Kp + nb, t2 = fn * NpEdge + rp;
    const double wo = a.IInII[fi], wn = a.IInII[fn];
    for (int c = 0; c < 4; ++c) {
      FE[4 * t + c] = F[c] * wo;
      FE[4 * t2 + c] = -F[c] * wn;
    }
    return;
  }
  boundary_value(a, code, fi, own, qfar, F);
  for (int c = 0; c < 4; ++c) FE[4 * t + c] = F[c];
}

// Per (element, solution point i): the RHS row i of the four variables
// (the interior flux at every solution point of the element, its
// contravariant transform, the DivInt row and the edge values), stored in R,
// then the stage's SSP54 combination at those four values in the order of
// the solver's stage kernel (stage s reads the step start U0, its input,
// its RHS and dt_s; the last stage also U2, U3 and the stage-3 RHS).
// A workgroup holds RhsElems<N> whole elements; thread (e, i) first forms
// the transformed flux at its own point i into LDS for its element's rows.
template <int N>
struct RhsElems {
  static constexpr int value = 256 / AdjShape<N>::NpInt;
};

template <int N>
struct AdjShock {
  using Sh = AdjShape<N>;
  static constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux;
  static constexpr int NR = NpInt + 3 * NpEdge;  // the used RT points: interior, then the edges
  static constexpr int n2 = NpInt * NpInt;
  static constexpr int oP = 0, oM = n2, oD = 2 * n2, oV = 3 * n2, oVi = 4 * n2, omf = 5 * n2, oDR = 5 * n2 + NpInt,
                       oBa = oDR + NR * NpFlux;
  // eps grad Q of variable n, direction dir, RT row r of element kk in [4][2][NR][Kp]
  __device__ static size_t dis_at(int Kp, int n, int dir, int r, int kk) {
    return (size_t((n * 2 + dir) * NR + r)) * Kp + kk;
  }
};

template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_rhs(AdjArgs a, int stage, const T *U, const T *FE, T *R,
                                                     const T *U0, const T *U2, const T *U3, const T *R3,
                                                     const T *dts, const T *dtk, T *out, const T *dis,
                                                     const T *sigb, double *Rpre) {
  using Sh = AdjShape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux, EB = RhsElems<N>::value;
  // Fr of the 4 variables, then Fs, per element: [EB][8][NpInt] (raw
  // storage: the dual type has a constructor)
  __shared__ double sraw[EB * 8 * NpInt * (sizeof(T) / sizeof(double))];
  T(*sF)[8][NpInt] = reinterpret_cast<T(*)[8][NpInt]>(sraw);
  const int el = threadIdx.x / NpInt, i = threadIdx.x - el * NpInt;
  const int kk = blockIdx.x * EB + el;
  const bool live = el < EB && kk < a.Kp;
  const bool real = live && kk < a.K;
  double Jd = 1., J0 = 0., J1 = 0., J2 = 0., J3 = 0.;
  const T *u = U + size_t(live ? kk : 0) * Sh::NV;
  if (real) {
    Jd = a.Jdet[kk];
    J0 = a.Jinv[kk];
    J1 = a.Jinv[size_t(a.Kp) + kk];
    J2 = a.Jinv[2 * size_t(a.Kp) + kk];
    J3 = a.Jinv[3 * size_t(a.Kp) + kk];
    const T q[4] = {u[i], u[NpInt + i], u[2 * NpInt + i], u[3 * NpInt + i]};
    T Fx[4], Fy[4];
    flux_xy(T(a.ph.gamma), q, Fx, Fy);
    if (a.act) {  // the physical flux minus eps grad Q (AddDissipation)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        Fx[c] -= dis[AdjShock<N>::dis_at(a.Kp, c, 0, i, kk)];
        Fy[c] -= dis[AdjShock<N>::dis_at(a.Kp, c, 1, i, kk)];
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      sF[el][c][i] = Jd * (J0 * Fx[c] + J1 * Fy[c]);
      sF[el][4 + c][i] = Jd * (J2 * Fx[c] + J3 * Fy[c]);
    }
  }
  __syncthreads();
  T acc[4] = {T(0.), T(0.), T(0.), T(0.)};
  if (live && real) {
    const double *D = a.DivInt + i * NpFlux;
    for (int j = 0; j < NpInt; ++j) {
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += D[j] * sF[el][c][j] + D[NpInt + j] * sF[el][4 + c][j];
    }
    for (int p = 0; p < Sh::NF; ++p) {
      const int f = p / NpEdge, pp = p - f * NpEdge;
      const T *fe = FE + 4 * ((size_t(f) * a.Kp + kk) * NpEdge + pp);
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += D[2 * NpInt + p] * fe[c];
    }
    const double m = -(1. / Jd);
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = acc[c] * m;
  }
  if (a.shock) {
    // the modal filter of the RHS: row i of F times the element's rows
    __syncthreads();  // every thread has read sF
    if (live) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sF[el][c][i] = acc[c];
      if (Rpre) {
#pragma unroll
        for (int c = 0; c < 4; ++c) Rpre[size_t(kk) * Sh::NV + c * NpInt + i] = ad::val(acc[c]);
      }
    }
    __syncthreads();
    if (a.act) {
      // F(sigma) = V diag(1, mf (1 - sin(pi sigma / 2))) Vinv, sigma the
      // vertex mean: modal coefficients, scaled, back to the points
      using Sk = AdjShock<N>;
      const double *V = a.shops + Sk::oV, *Vi = a.shops + Sk::oVi, *mf = a.shops + Sk::omf;
      T mm[4];
      if (live && real) {
        const T fac = 1. - sin((0.5 * M_PI) * sigb[kk]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          T t = T(0.);
          for (int j = 0; j < NpInt; ++j) t += Vi[i * NpInt + j] * sF[el][c][j];
          mm[c] = i == 0 ? t : t * (mf[i] * fac);
        }
      }
      __syncthreads();
      if (live && real) {
#pragma unroll
        for (int c = 0; c < 4; ++c) sF[el][c][i] = mm[c];
      }
      __syncthreads();
      if (live && real) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          T t = T(0.);
          for (int m = 0; m < NpInt; ++m) t += V[i * NpInt + m] * sF[el][c][m];
          acc[c] = t;
        }
      }
    } else if (live && real) {
      const double *Fi = a.filt + i * NpInt;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T t = T(0.);
        for (int k = 0; k < NpInt; ++k) t += Fi[k] * sF[el][c][k];
        acc[c] = t;
      }
    }
  }
  if (!live) return;
  const T dt = a.lts ? dtk[kk] : dts[stage];  // the element's dt with local time stepping
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const size_t x = size_t(kk) * Sh::NV + c * NpInt + i;
    R[x] = acc[c];
    const T dtR = dt * acc[c];
    switch (stage) {
      case 0: out[x] = U[x] + kSspA0 * dtR; break;
      case 1: out[x] = kSspB1 * U0[x] + kSspC1 * U[x] + kSspA1 * dtR; break;
      case 2: out[x] = kSspB2 * U0[x] + kSspC2 * U[x] + kSspA2 * dtR; break;
      case 3: out[x] = kSspB3 * U0[x] + kSspC3 * U[x] + kSspA3 * dtR; break;
      default: {
        const T dtR3 = dt * R3[x];
        const T Rn = -U0[x] + kSspD2 * U2[x] + kSspD3 * U3[x] + kSspD4 * U[x] + kSspE3 * dtR3 + kSspE4 * dtR;
        out[x] = U0[x] + Rn;
      }
    }
  }
}

// Local time stepping (CalculateElementMaxWaveSpeed edges.go:291-323,
// CalculateLocalDT euler.go:973-1002): per element, w = the previous
// stage's element dt (-100 at stage 0, the reference's unit-mixing max),
// then the max with the speeds of the points of its three faces (read from
// the edge owner's face); dt = CFL / w. Records the point that sets w (-1:
// the previous dt) for the adjoint. Ties keep the earlier candidate.
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_lts(AdjArgs a, int stage, const T *SP, const T *dtk_prev, T *dtk,
                                                     int *srck, double *wk) {
  constexpr int NpEdge = AdjShape<N>::NpEdge;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.Kp) return;
  if (kk >= a.K) {
    dtk[kk] = T(0.);
    srck[kk] = -1;
    wk[kk] = 0.;
    return;
  }
  T w = stage == 0 ? T(-100.) : dtk_prev[kk];
  int src = -1;
  for (int f = 0; f < 3; ++f) {
    const size_t fs = size_t(a.wsrc[size_t(f) * a.Kp + kk]);
    for (int pp = 0; pp < NpEdge; ++pp) {
      const size_t t = fs * NpEdge + pp;
      const T sp = SP[t];
      if (ad::val(sp) > ad::val(w)) {
        w = sp;
        src = int(t);
      }
    }
  }
  dtk[kk] = a.ph.cfl / w;
  srck[kk] = src;
  wk[kk] = ad::val(w);
}

// PerssonC0 sensor of each element's density (UpdateSeMoment
// DG2D/dfr_shock_capturing.go:142-166): Se = log10(sum DQ.(P r) / sum r.(M r))
// with DQ = D r; flags the stage (flag |= 4) where Se reaches S0 - Kappa,
// the start of the sigma ramp, beyond which the quiet-sensor derivative
// does not hold.
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_sensor(AdjArgs a, const T *U, int *flag) {
  constexpr int NpInt = AdjShape<N>::NpInt;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.K) return;
  const T *rho = U + size_t(kk) * AdjShape<N>::NV;
  const double *Pm = a.sens, *Mm = Pm + NpInt * NpInt, *Dm = Mm + NpInt * NpInt;
  double num = 0., den = 0.;
  for (int i = 0; i < NpInt; ++i) {
    double X = 0., Y = 0., DQ = 0.;
    for (int j = 0; j < NpInt; ++j) {
      const double r = ad::val(rho[j]);
      X += Pm[i * NpInt + j] * r;
      Y += Mm[i * NpInt + j] * r;
      DQ += Dm[i * NpInt + j] * r;
    }
    num += DQ * X;
    den += ad::val(rho[i]) * Y;
  }
  const double se = log10(num / den);
  if (se >= a.sens_left) atomicOr(flag, 4);
}

// X <- F X (or F^T X) for each element and variable, in place
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_filter(AdjArgs a, T *X, int transpose) {
  constexpr int NpInt = AdjShape<N>::NpInt;
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(a.K) * 4) return;
  T *x = X + gid * NpInt;  // element gid / 4, variable gid % 4
  T u[NpInt];
  for (int i = 0; i < NpInt; ++i) u[i] = x[i];
  for (int i = 0; i < NpInt; ++i) {
    T t = T(0.);
    for (int k = 0; k < NpInt; ++k) t += (transpose ? a.filt[k * NpInt + i] : a.filt[i * NpInt + k]) * u[k];
    x[i] = t;
  }
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

// Reverse-mode derivative of roe_flux (RoeFlux, fluxes.go:284-413), by
// hand: given the cotangent bF of F*(L, R, n), adds (dF*/dL)^T bF to bL and
// (dF*/dR)^T bF to bR. Statement by statement the transpose of the forward
// evaluation in physics.hpp; |x| takes the derivative sign(x) with +1 at 0.
__device__ __forceinline__ void roe_flux_vjp(double g, const double L[4], const double R[4], double nx, double ny,
                                             const double bF[4], double bL[4], double bR[4]) {
  const double GM1 = g - 1.;
  // forward
  const double rULr = L[1] * nx + L[2] * ny, rVLr = L[1] * (-ny) + L[2] * nx;
  const double rURr = R[1] * nx + R[2] * ny, rVRr = R[1] * (-ny) + R[2] * nx;
  const double rhoL = L[0], uL = rULr / L[0], vL = rVLr / L[0];
  const double rhoR = R[0], uR = rURr / R[0], vR = rVRr / R[0];
  const double cuL = L[1] / L[0], cvL = L[2] / L[0], cuR = R[1] / R[0], cvR = R[2] / R[0];
  const double pL = GM1 * (L[3] - 0.5 * rhoL * (cuL * cuL + cvL * cvL));
  const double pR = GM1 * (R[3] - 0.5 * rhoR * (cuR * cuR + cvR * cvR));
  const double hL = (L[3] + pL) / rhoL, hR = (R[3] + pR) / rhoR;
  const double sL = sqrt(rhoL), sR = sqrt(rhoR), sLR = sL + sR;
  const double rho = sL * sR;
  const double u = (sL * uL + sR * uR) / sLR;
  const double v = (sL * vL + sR * vR) / sLR;
  const double h = (sL * hL + sR * hR) / sLR;
  const double c2 = GM1 * (h - 0.5 * (u * u + v * v));
  const double cc = sqrt(c2);
  const double du = uR - uL, dv = vR - vL, dp = pR - pL;
  const double X = rho * du / cc, Y = dp / c2;
  const double dW1 = -0.5 * X + 0.5 * Y, dW2 = (rhoR - rhoL) - Y, dW3 = rho * dv, dW4 = 0.5 * X + 0.5 * Y;
  const double s1 = (u - cc) < 0. ? -1. : 1., s2 = u < 0. ? -1. : 1., s4 = (u + cc) < 0. ? -1. : 1.;
  const double a1 = s1 * (u - cc), a2 = s2 * u, a4 = s4 * (u + cc);
  const double W1 = a1 * dW1, W2 = a2 * dW2, W3 = a2 * dW3, W4 = a4 * dW4;
  // reverse: the rotation back to x/y
  const double bf0 = bF[0], bf1 = nx * bF[1] + ny * bF[2], bf2 = -ny * bF[1] + nx * bF[2], bf3 = bF[3];
  double b_rULr = 0.5 * bf0 + 0.5 * bf1 * uL, b_rURr = 0.5 * bf0 + 0.5 * bf1 * uR;
  double b_rVLr = 0.5 * bf2 * uL, b_rVRr = 0.5 * bf2 * uR;
  double b_uL = 0.5 * bf1 * rULr + 0.5 * bf2 * rVLr + 0.5 * bf3 * (pL + L[3]);
  double b_uR = 0.5 * bf1 * rURr + 0.5 * bf2 * rVRr + 0.5 * bf3 * (pR + R[3]);
  double b_pL = 0.5 * bf1 + 0.5 * bf3 * uL, b_pR = 0.5 * bf1 + 0.5 * bf3 * uR;
  double bL3 = 0.5 * bf3 * uL, bR3 = 0.5 * bf3 * uR;
  const double bW1 = -0.5 * (bf0 + bf1 * (u - cc) + bf2 * v + bf3 * (h - u * cc));
  const double bW2 = -0.5 * (bf0 + bf1 * u + bf2 * v + bf3 * 0.5 * (u * u + v * v));
  const double bW3 = -0.5 * (bf2 + bf3 * v);
  const double bW4 = -0.5 * (bf0 + bf1 * (u + cc) + bf2 * v + bf3 * (h + u * cc));
  double b_u = -0.5 * bf1 * (W1 + W2 + W4) - 0.5 * bf3 * (cc * (W4 - W1) + W2 * u);
  double b_v = -0.5 * bf2 * (W1 + W2 + W4) - 0.5 * bf3 * (W2 * v + W3);
  double b_h = -0.5 * bf3 * (W1 + W4);
  double b_cc = -0.5 * bf1 * (W4 - W1) - 0.5 * bf3 * u * (W4 - W1);
  // the wave strengths and their speeds
  const double b_a1 = bW1 * dW1, b_a2 = bW2 * dW2 + bW3 * dW3, b_a4 = bW4 * dW4;
  const double b_dW1 = bW1 * a1, b_dW2 = bW2 * a2, b_dW3 = bW3 * a2, b_dW4 = bW4 * a4;
  b_u += b_a1 * s1 + b_a2 * s2 + b_a4 * s4;
  b_cc += -b_a1 * s1 + b_a4 * s4;
  const double b_X = 0.5 * (b_dW4 - b_dW1), b_Y = 0.5 * (b_dW1 + b_dW4) - b_dW2;
  double b_rhoL = -b_dW2, b_rhoR = b_dW2;
  double b_rho = b_dW3 * dv + b_X * du / cc;
  double b_vL = -b_dW3 * rho, b_vR = b_dW3 * rho;
  const double b_du = b_X * rho / cc;
  b_uR += b_du;
  b_uL -= b_du;
  b_cc += -b_X * X / cc;
  const double b_dp = b_Y / c2;
  b_pR += b_dp;
  b_pL -= b_dp;
  double b_c2 = -b_Y * Y / c2 + b_cc / (2. * cc);
  b_h += GM1 * b_c2;
  b_u += -GM1 * u * b_c2;
  b_v += -GM1 * v * b_c2;
  // Roe averages
  double b_sL = b_rho * sR + (b_u * (uL - u) + b_v * (vL - v) + b_h * (hL - h)) / sLR;
  double b_sR = b_rho * sL + (b_u * (uR - u) + b_v * (vR - v) + b_h * (hR - h)) / sLR;
  b_uL += b_u * sL / sLR;
  b_uR += b_u * sR / sLR;
  b_vL += b_v * sL / sLR;
  b_vR += b_v * sR / sLR;
  const double b_hL = b_h * sL / sLR, b_hR = b_h * sR / sLR;
  b_rhoL += b_sL / (2. * sL);
  b_rhoR += b_sR / (2. * sR);
  // enthalpies
  bL3 += b_hL / rhoL;
  b_pL += b_hL / rhoL;
  b_rhoL += -b_hL * hL / rhoL;
  bR3 += b_hR / rhoR;
  b_pR += b_hR / rhoR;
  b_rhoR += -b_hR * hR / rhoR;
  // normal-frame velocities
  b_rULr += b_uL / L[0];
  b_rVLr += b_vL / L[0];
  b_rhoL += -(b_uL * uL + b_vL * vL) / L[0];
  b_rURr += b_uR / R[0];
  b_rVRr += b_vR / R[0];
  b_rhoR += -(b_uR * uR + b_vR * vR) / R[0];
  // pressures
  bL[0] += b_rhoL + b_pL * GM1 * 0.5 * (cuL * cuL + cvL * cvL);
  bL[1] += -b_pL * GM1 * cuL + b_rULr * nx - b_rVLr * ny;
  bL[2] += -b_pL * GM1 * cvL + b_rULr * ny + b_rVLr * nx;
  bL[3] += b_pL * GM1 + bL3;
  bR[0] += b_rhoR + b_pR * GM1 * 0.5 * (cuR * cuR + cvR * cvR);
  bR[1] += -b_pR * GM1 * cuR + b_rURr * nx - b_rVRr * ny;
  bR[2] += -b_pR * GM1 * cvR + b_rURr * ny + b_rVRr * nx;
  bR[3] += b_pR * GM1 + bR3;
}

// Reverse-mode derivative of flux_xy (FluxCalcBase, fluxes.go:76-87), by
// hand: adds (dFx/dq)^T bFx + (dFy/dq)^T bFy to bq, with u = rhoU/rho,
// v = rhoV/rho, p = (g-1)(E - (rhoU^2 + rhoV^2)/(2 rho)):
// Fx = (rhoU, rhoU u + p, rhoU v, u (E + p)), Fy = (rhoV, rhoV u, rhoV v + p, v (E + p)).
__device__ __forceinline__ void flux_xy_vjp(double g, const double q[4], const double bFx[4], const double bFy[4],
                                            double bq[4]) {
  const double GM1 = g - 1.;
  const double r = q[0], u = q[1] / r, v = q[2] / r, E = q[3];
  const double p = GM1 * (E - 0.5 * r * (u * u + v * v)), H = E + p;
  // cotangents of u, v, p, and the direct rhoU, rhoV, E terms
  double b_u = bFx[1] * q[1] + bFx[3] * H + bFy[1] * q[2];
  double b_v = bFx[2] * q[1] + bFy[2] * q[2] + bFy[3] * H;
  const double b_p = bFx[1] + bFy[2] + bFx[3] * u + bFy[3] * v;
  double b_ru = bFx[0] + bFx[1] * u + bFx[2] * v;
  double b_rv = bFy[0] + bFy[1] * u + bFy[2] * v;
  double b_E = bFx[3] * u + bFy[3] * v;
  // p
  b_E += GM1 * b_p;
  b_ru += -GM1 * u * b_p;
  b_rv += -GM1 * v * b_p;
  double b_r = GM1 * 0.5 * (u * u + v * v) * b_p;
  // u = rhoU / rho, v = rhoV / rho
  b_ru += b_u / r;
  b_rv += b_v / r;
  b_r += -(b_u * u + b_v * v) / r;
  bq[0] += b_r;
  bq[1] += b_ru;
  bq[2] += b_rv;
  bq[3] += b_E;
}

// Reverse-mode derivative of lax_flux (LaxFlux, fluxes.go:161-190), by
// hand: adds (dF*/dL)^T bF to bL and (dF*/dR)^T bF to bR. The maximum of
// the two sides' |V| + c takes the left side's branch on a tie, as the
// dual-number fmax does.
__device__ __forceinline__ void lax_flux_vjp(double g, const double L[4], const double R[4], double nx, double ny,
                                             const double bF[4], double bL[4], double bR[4]) {
  const double GM1 = g - 1.;
  const double *Q[2] = {L, R};
  double *B[2] = {bL, bR};
  double uu[2], vv[2], pp[2], cc[2], zz[2], ss[2], VV[2];
  for (int k = 0; k < 2; ++k) {
    const double *q = Q[k];
    uu[k] = q[1] / q[0];
    vv[k] = q[2] / q[0];
    pp[k] = GM1 * (q[3] - 0.5 * q[0] * (uu[k] * uu[k] + vv[k] * vv[k]));
    zz[k] = g * pp[k] / q[0];
    cc[k] = sqrt(fabs(zz[k]));
    ss[k] = sqrt(uu[k] * uu[k] + vv[k] * vv[k]);
    VV[k] = ss[k] + cc[k];
  }
  const bool left = VV[0] >= VV[1];
  const double maxV = left ? VV[0] : VV[1];
  double b_maxV = 0.;
  for (int c = 0; c < 4; ++c) {
    bL[c] += 0.5 * maxV * bF[c];
    bR[c] -= 0.5 * maxV * bF[c];
    b_maxV += 0.5 * bF[c] * (L[c] - R[c]);
  }
  for (int k = 0; k < 2; ++k) {
    const double *q = Q[k];
    double *b = B[k];
    const double u = uu[k], v = vv[k], p = pp[k], H = p + q[3];
    // the central part: each side's physical flux along n, halved
    b[1] += 0.5 * nx * bF[0] + 0.5 * bF[1] * (nx * u + ny * v);
    b[2] += 0.5 * ny * bF[0] + 0.5 * bF[2] * (nx * u + ny * v);
    double b_u = 0.5 * (bF[1] * nx * q[1] + bF[2] * nx * q[2] + bF[3] * nx * H);
    double b_v = 0.5 * (bF[1] * ny * q[1] + bF[2] * ny * q[2] + bF[3] * ny * H);
    const double b_H = 0.5 * bF[3] * (nx * u + ny * v);
    double b_p = 0.5 * (bF[1] * nx + bF[2] * ny) + b_H;
    b[3] += b_H;
    // |V| + c of the side that attains the maximum
    if ((k == 0) == left) {
      if (ss[k] > 0.) {
        b_u += b_maxV * u / ss[k];
        b_v += b_maxV * v / ss[k];
      }
      const double b_z = b_maxV * (zz[k] < 0. ? -1. : 1.) / (2. * cc[k]);
      b_p += b_z * g / q[0];
      b[0] += -b_z * zz[k] / q[0];
    }
    // velocities and pressure
    b[1] += b_u / q[0] - GM1 * u * b_p;
    b[2] += b_v / q[0] - GM1 * v * b_p;
    b[0] += -(b_u * u + b_v * v) / q[0] + GM1 * 0.5 * (u * u + v * v) * b_p;
    b[3] += GM1 * b_p;
  }
}

// Per face point: the transposed Jacobians of its edge value with respect
// to its inputs, applied to the edge value's cotangent, into the trace
// cotangent of each face point (lamOwn, written exactly once per launch)
// and the far state's (lamFarPts, accumulated). A shared edge point is
// handled by the owner's face point: the cotangent of F* is
// w_own lambda_own - w_nbr lambda_nbr, and the transposed Jacobians with
// respect to the two traces go to the owner's and the partner's records.
// The Jacobians are formed one input direction at a time in dual
// arithmetic.
// (Capping its registers for occupancy spills and measured slower.)
template <int N, int FLUX>
__global__ void __launch_bounds__(256) k_adj_face(AdjArgs a, const double *U, const double *lamR, double *lamOwn) {
  using Sh = AdjShape<N>;
  constexpr int NpEdge = Sh::NpEdge;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  double *lo = lamOwn + 4 * t;
  if (kk >= a.K) {
    for (int c = 0; c < 4; ++c) lo[c] = 0.;
    return;
  }
  const int code = a.code[fi], kind = code & 15;
  // unshared points: k_adj_face_bnd; a non-owner's point: written by its owner
  if (kind != FK_Shared || !(code & FC_FLUX_OWNER)) return;
  double own[4];
  trace_at<N>(a, U, kk, f * NpEdge + pp, own);
  const int nb = a.nbr_elem[fi], nf = a.nbr_face[fi], rp = NpEdge - 1 - pp;
  const size_t fn = size_t(nf) * a.Kp + nb;
  double oth[4], lfo[4], lfn[4], lF[4];
  trace_at<N>(a, U, nb, nf * NpEdge + rp, oth);
  edge_cotangent<N>(a, lamR, kk, f * NpEdge + pp, lfo);
  edge_cotangent<N>(a, lamR, nb, nf * NpEdge + rp, lfn);
  const double wo = a.IInII[fi], wn = a.IInII[fn];
  for (int c = 0; c < 4; ++c) lF[c] = wo * lfo[c] - wn * lfn[c];
  double *ln = lamOwn + 4 * (fn * NpEdge + rp);
  if constexpr (FLUX <= 2) {
    double bL[4] = {0., 0., 0., 0.}, bR[4] = {0., 0., 0., 0.};
    if constexpr (FLUX == 2) {
      roe_flux_vjp(a.ph.gamma, own, oth, a.Nx[fi], a.Ny[fi], lF, bL, bR);
    } else if constexpr (FLUX == 1) {
      lax_flux_vjp(a.ph.gamma, own, oth, a.Nx[fi], a.Ny[fi], lF, bL, bR);
    } else {
      // AvgFlux: 0.5 (n . F(L) + n . F(R)) (fluxes.go:135-159)
      double bx[4], by[4];
      for (int c = 0; c < 4; ++c) {
        bx[c] = 0.5 * a.Nx[fi] * lF[c];
        by[c] = 0.5 * a.Ny[fi] * lF[c];
      }
      flux_xy_vjp(a.ph.gamma, own, bx, by, bL);
      flux_xy_vjp(a.ph.gamma, oth, bx, by, bR);
    }
    for (int c = 0; c < 4; ++c) {
      lo[c] = bL[c];
      ln[c] = bR[c];
    }
  } else {
    const Dual g(a.ph.gamma), nx(a.Nx[fi]), ny(a.Ny[fi]);
#pragma unroll 1
    for (int j = 0; j < 8; ++j) {
      const int side = j >> 2, d = j & 3;
      Dual L[4], R[4], F[4];
      for (int c = 0; c < 4; ++c) {
        L[c] = Dual(own[c], side == 0 && c == d ? 1. : 0.);
        R[c] = Dual(oth[c], side == 1 && c == d ? 1. : 0.);
      }
      numflux(FLUX, g, L, R, nx, ny, F);
      double sum = 0.;
      for (int c = 0; c < 4; ++c) sum += lF[c] * F[c].d;
      if (side == 0) lo[d] = sum;
      else ln[d] = sum;
    }
  }
}

// The unshared face points (boundary kinds) of real elements, from a list:
// the boundary flux's transposed Jacobians with respect to the own trace
// and the far state, in dual arithmetic, one thread per (point, input
// direction): 0..3 the own trace, 4..7 the far state
template <int N>
__global__ void __launch_bounds__(256) k_adj_face_bnd(AdjArgs a, const double *U, const double *lamR,
                                                      double *lamOwn, double *lamFarPts) {
  using Sh = AdjShape<N>;
  constexpr int NpEdge = Sh::NpEdge;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= 8 * a.nbnd) return;
  const int ib = gid >> 3, j = gid & 7, side = j >> 2, d = j & 3;
  const size_t t = size_t(a.bnd[ib]);
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  const int code = a.code[fi], kind = code & 15;
  const bool far = kind == FK_Far && ((code >> 8) == 0 || a.io_tied);
  if (side == 1 && !far) return;
  double own[4], lf[4], qfar[4];
  trace_at<N>(a, U, kk, f * NpEdge + pp, own);
  edge_cotangent<N>(a, lamR, kk, f * NpEdge + pp, lf);
  far_state(a, qfar);
  Dual o[4], q[4], F[4];
  for (int c = 0; c < 4; ++c) {
    o[c] = Dual(own[c], side == 0 && c == d ? 1. : 0.);
    q[c] = Dual(qfar[c], side == 1 && c == d ? 1. : 0.);
  }
  boundary_value(a, code, fi, o, q, F);
  double sum = 0.;
  for (int c = 0; c < 4; ++c) sum += lf[c] * F[c].d;
  if (side == 0) lamOwn[4 * t + d] = sum;
  else lamFarPts[4 * t + d] += sum;
}

// Per (element, solution point j): column j of the FluxEdgeInterp transpose
// of the element's face-point trace cotangents, and the interior-flux
// transpose at point j (the DivInt columns of its r and s components, the
// contravariant transform and the physical-flux Jacobian), added to target.
template <int N>
__global__ void __launch_bounds__(256) k_adj_elem(AdjArgs a, const double *U, const double *lamR,
                                                  const double *lamOwn, double *target) {
  using Sh = AdjShape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux;
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(a.K) * NpInt) return;
  const int kk = int(gid / NpInt), j = int(gid - size_t(kk) * NpInt);
  const double *u = U + size_t(kk) * Sh::NV;
  const double *lr = lamR + size_t(kk) * Sh::NV;
  const double Jd = a.Jdet[kk], m = -(1. / Jd);
  const double J0 = a.Jinv[kk], J1 = a.Jinv[size_t(a.Kp) + kk];
  const double J2 = a.Jinv[2 * size_t(a.Kp) + kk], J3 = a.Jinv[3 * size_t(a.Kp) + kk];
  const double *D = a.DivInt;
  double acc[4] = {0., 0., 0., 0.};
  for (int p = 0; p < Sh::NF; ++p) {
    const int f = p / NpEdge, pp = p - f * NpEdge;
    const double Ij = a.Interp[p * NpInt + j];
    const double *lq = lamOwn + 4 * ((size_t(f) * a.Kp + kk) * NpEdge + pp);
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += Ij * lq[c];
  }
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
    acc[d] += sum;
  }
  double *tg = target + size_t(kk) * Sh::NV;
#pragma unroll
  for (int c = 0; c < 4; ++c) tg[c * NpInt + j] += acc[c];
}

// Transpose of one SSP54 stage with element dts (one thread per element):
// the updates of k_adj_ssp, and the element's dt cotangent (its dot product
// plus what the next stage's dt passed back through the previous-dt term),
// routed through dt = CFL / w to the point that set w (a face-point speed
// cotangent; at most the two elements of an edge add to one point, so the
// sum does not depend on their order) or, if w was the previous stage's dt,
// carried to that stage.
template <int N>
__global__ void __launch_bounds__(256) k_adj_ssp_lts(AdjArgs a, int stage, const double *lam, double *lamU0,
                                                     double *lamQ1, double *lamQ2, double *lamQ3, double *lamQ4,
                                                     double *lamR, double *lamR3, const double *Rs, const double *R3,
                                                     const double *dtk, const int *srck, const double *wk,
                                                     double *carry, double *lamSpeed) {
  constexpr int NV = AdjShape<N>::NV;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.Kp) return;
  const double dt = dtk[kk];
  double dot = 0.;
  for (size_t i = size_t(kk) * NV; i < size_t(kk + 1) * NV; ++i) {
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
  if (kk >= a.K) {
    carry[kk] = 0.;
    return;
  }
  const double lam_d = dot + carry[kk];
  const double W = wk[kk];
  const double lw = -(a.ph.cfl / (W * W)) * lam_d;
  const int src = srck[kk];
  if (src < 0) {
    carry[kk] = lw;
  } else {
    atomicAdd(&lamSpeed[src], lw);
    carry[kk] = 0.;
  }
}

// The face-point speed cotangents of local time stepping into the trace
// (and far-state) cotangents of the point; clears them for the next stage
template <int N>
__global__ void __launch_bounds__(256) k_adj_lts_wave(AdjArgs a, const double *U, double *lamSpeed, double *lamOwn,
                                                      double *lamFarPts) {
  using Sh = AdjShape<N>;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * Sh::NpEdge) return;
  const double ls = lamSpeed[t];
  if (ls == 0.) return;
  lamSpeed[t] = 0.;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  const int code = a.code[fi];
  double own[4], qfar[4];
  trace_at<N>(a, U, kk, f * Sh::NpEdge + pp, own);
  far_state(a, qfar);
  const bool far = (code & 15) == FK_Far && ((code >> 8) == 0 || a.io_tied);
  for (int d = 0; d < 4; ++d) {
    Dual o[4], q[4];
    for (int c = 0; c < 4; ++c) {
      o[c] = Dual(own[c], c == d ? 1. : 0.);
      q[c] = Dual(qfar[c]);
    }
    lamOwn[4 * t + d] += ls * wave_value(a, code, fi, a.oohK[kk], o, q).d;
    if (far) {
      for (int c = 0; c < 4; ++c) {
        o[c] = Dual(own[c]);
        q[c] = Dual(qfar[c], c == d ? 1. : 0.);
      }
      lamFarPts[4 * t + d] += ls * wave_value(a, code, fi, a.oohK[kk], o, q).d;
    }
  }
}

// The dt term of a stage: dt = CFL / W with W the wave speed at the stage's
// arg-max face point; its cotangent goes to that point's trace (and the far
// state for a far-field face). One thread.
template <int N>
__global__ void k_adj_wave(AdjArgs a, int stage, const double *U, const int *argmax, const double *Wv,
                           const double *lam_dt, double *target, double *lamFarW, const int *vbr) {
  using Sh = AdjShape<N>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (vbr && vbr[stage]) return;  // the viscous dt bound this stage
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

// ---- PerssonC0 with an active sensor ----
//
// Forward (T = double or ad::Dual), per stage: the sensor fields of the
// stage input (k_adj_fwd_sens, k_adj_fwd_vmax, k_adj_fwd_efield), the
// stage-2 state filter with the vertex-mean sigma (k_adj_filter_sig), the
// RT gradient times eps (k_adj_fwd_grad) and the viscous edge values
// (k_adj_fwd_visc). Adjoint: their transposes (k_adj_filter_sig_t,
// k_adj_visc_t, k_adj_grad_t, k_adj_eq_gather, k_adj_vdt, k_adj_vert_t,
// k_adj_sens_t), deterministic (each cotangent is written by one thread or
// gathered in a fixed order).

// sigma(Se) of UpdateShockFinderSigma (dissipation.go:491-518) with the
// stepping path's S0 = 4/(N+1)^4, and its slope
template <class T>
__device__ __forceinline__ T sigma_ramp(const AdjArgs &a, T se) {
  const double left = a.S0 - a.kappa, right = a.S0 + a.kappa, ook = 0.5 / a.kappa;
  if (se < left) return T(0.);
  if (se <= right) return 0.5 * (1. + sin((M_PI * ook) * (se - a.S0)));
  if (se > right) return T(1.);
  return T(0.);
}
__device__ __forceinline__ double sigma_slope(const AdjArgs &a, double se) {
  const double left = a.S0 - a.kappa, right = a.S0 + a.kappa, ook = 0.5 / a.kappa;
  if (!(se >= left && se <= right)) return 0.;
  return 0.5 * (M_PI * ook) * cos((M_PI * ook) * (se - a.S0));
}

// Per element: Se of the stage input's density (UpdateSeMoment
// DG2D/dfr_shock_capturing.go:142-166), sigma and eps = eps0 hK sigma
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_sens(AdjArgs a, const T *U, T *sig, T *eps) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.Kp) return;
  if (kk >= a.K) {
    sig[kk] = T(0.);
    eps[kk] = T(0.);
    return;
  }
  const T *rho = U + size_t(kk) * AdjShape<N>::NV;
  const double *Pm = a.shops + Sk::oP, *Mm = a.shops + Sk::oM, *Dm = a.shops + Sk::oD;
  T num = T(0.), den = T(0.);
  for (int i = 0; i < NpInt; ++i) {
    T X = T(0.), Y = T(0.), DQ = T(0.);
    for (int j = 0; j < NpInt; ++j) {
      const T r = rho[j];
      X += Pm[i * NpInt + j] * r;
      Y += Mm[i * NpInt + j] * r;
      DQ += Dm[i * NpInt + j] * r;
    }
    num += DQ * X;
    den += rho[i] * Y;
  }
  const T sg = sigma_ramp(a, log10(num / den));
  sig[kk] = sg;
  eps[kk] = (a.eps0 * a.hK[kk]) * sg;
}

// Vertex maxima of sigma and eps over the real elements at the vertex
// (MergeElementScalarToVertices euler.go:1048-1065), with the element that
// attains each (the first in incidence order among equals)
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_vmax(AdjArgs a, const T *sig, const T *eps, T *sigv, T *epsv,
                                                      int *args, int *arge) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.Nv) return;
  T bs = T(0.), be = T(0.);
  int is = -1, ie = -1;
  for (int q = a.vptr[v]; q < a.vptr[v + 1]; ++q) {
    const int kk = a.vinc[q] >> 2;
    const T s = sig[kk], e = eps[kk];
    if (is < 0 || s > bs) {
      bs = s;
      is = kk;
    }
    if (ie < 0 || e > be) {
      be = e;
      ie = kk;
    }
  }
  sigv[v] = bs;
  epsv[v] = be;
  args[v] = is;
  arge[v] = ie;
}

// Per element: sigma as the mean of its vertices' maxima, eps at the used
// RT points by barycentric interpolation of its vertices' maxima
// (InterpolateEpsilonSigma dissipation.go:219-242), and eps / hK^2 at its
// owner-side face points, the viscous speed of StoreEdgeAggregates
// (edges.go:246-289); -1 at the other face points
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_efield(AdjArgs a, const T *sigv, const T *epsv, T *sigb, T *epsf,
                                                        T *SPV) {
  using Sk = AdjShock<N>;
  constexpr int NR = Sk::NR, NpInt = Sk::NpInt, NpEdge = Sk::NpEdge;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.Kp) return;
  const size_t Kp = size_t(a.Kp);
  if (kk >= a.K) {
    sigb[kk] = T(0.);
    for (int r = 0; r < NR; ++r) epsf[r * Kp + kk] = T(0.);
    for (int f = 0; f < 3; ++f)
      for (int pp = 0; pp < NpEdge; ++pp) SPV[(f * Kp + kk) * NpEdge + pp] = T(-1.);
    return;
  }
  const int v0 = a.evert[kk], v1 = a.evert[Kp + kk], v2 = a.evert[2 * Kp + kk];
  sigb[kk] = ((sigv[v0] + sigv[v1]) + sigv[v2]) / 3.;
  const T e0 = epsv[v0], e1 = epsv[v1], e2 = epsv[v2];
  const double *B = a.shops + Sk::oBa;
  for (int r = 0; r < NR; ++r) epsf[r * Kp + kk] = B[3 * r] * e0 + B[3 * r + 1] * e1 + B[3 * r + 2] * e2;
  const double oohK = a.oohK[kk];
  for (int f = 0; f < 3; ++f) {
    const bool owner = a.code[f * Kp + kk] & FC_WAVE;
    for (int pp = 0; pp < NpEdge; ++pp)
      SPV[(f * Kp + kk) * NpEdge + pp] = owner ? (oohK * oohK) * epsf[(NpInt + f * NpEdge + pp) * Kp + kk] : T(-1.);
  }
}

// X <- F(sigma) X per element and variable, in place (limitAndFilterSolution
// dissipation.go:606-622 with the vertex-mean sigma)
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_filter_sig(AdjArgs a, T *X, const T *sigb) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt;
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(a.K) * 4) return;
  const int kk = int(gid >> 2);
  T *x = X + gid * NpInt;
  const double *V = a.shops + Sk::oV, *Vi = a.shops + Sk::oVi, *mf = a.shops + Sk::omf;
  const T fac = 1. - sin((0.5 * M_PI) * sigb[kk]);
  T u[NpInt], m[NpInt];
  for (int i = 0; i < NpInt; ++i) u[i] = x[i];
  for (int i = 0; i < NpInt; ++i) {
    T t = T(0.);
    for (int j = 0; j < NpInt; ++j) t += Vi[i * NpInt + j] * u[j];
    m[i] = i == 0 ? t : t * (mf[i] * fac);
  }
  for (int i = 0; i < NpInt; ++i) {
    T t = T(0.);
    for (int j = 0; j < NpInt; ++j) t += V[i * NpInt + j] * m[j];
    x[i] = t;
  }
}

// Variable n of the edge-owner's stored trace that face point (f, i) of
// element kk reads: its own trace where it owns the edge (or the face is
// unshared), else the owner's at the reversed point (edges.go:225-237,
// euler.go:889-903)
template <int N, class T>
__device__ __forceinline__ T owner_trace(const AdjArgs &a, const T *U, int kk, int f, int i, int n) {
  using Sh = AdjShape<N>;
  const size_t fi = size_t(f) * a.Kp + kk;
  const int code = a.code[fi];
  int ke = kk, p = f * Sh::NpEdge + i;
  if ((code & 15) == FK_Shared && !(code & FC_FLUX_OWNER)) {
    ke = a.nbr_elem[fi];
    p = a.nbr_face[fi] * Sh::NpEdge + (Sh::NpEdge - 1 - i);
  }
  const T *u = U + size_t(ke) * Sh::NV + n * Sh::NpInt;
  const double *row = a.Interp + p * Sh::NpInt;
  T t = T(0.);
  for (int j = 0; j < Sh::NpInt; ++j) t += row[j] * u[j];
  return t;
}

// The RT-element gradient DOFs of variable n of element kk
// (GetSolutionGradientUsingRTElement euler.go:864-918): interior Jinv U,
// edges n IInII / Jdet times the owner's stored trace
template <int N, class T>
__device__ __forceinline__ void grad_dofs(const AdjArgs &a, const T *U, int kk, int n, T *DX, T *DY) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt, NpEdge = Sk::NpEdge;
  const size_t Kp = size_t(a.Kp);
  const double ooJd = 1. / a.Jdet[kk];
  const double J0 = a.Jinv[kk], J1 = a.Jinv[Kp + kk], J2 = a.Jinv[2 * Kp + kk], J3 = a.Jinv[3 * Kp + kk];
  const T *u = U + size_t(kk) * AdjShape<N>::NV + n * NpInt;
  for (int i = 0; i < NpInt; ++i) {
    const T un = u[i];
    DX[i] = J0 * un;
    DY[i] = J1 * un;
    DX[i + NpInt] = J2 * un;
    DY[i + NpInt] = J3 * un;
  }
  for (int f = 0; f < 3; ++f) {
    const size_t fi = f * Kp + kk;
    const double nx = a.Nxo[fi], ny = a.Nyo[fi], w = a.IInII[fi];
    for (int i = 0; i < NpEdge; ++i) {
      const T q = owner_trace<N>(a, U, kk, f, i, n);
      DX[2 * NpInt + f * NpEdge + i] = ooJd * nx * w * q;
      DY[2 * NpInt + f * NpEdge + i] = ooJd * ny * w * q;
    }
  }
}

__device__ __forceinline__ bool any_nonzero(double x) { return x != 0.; }
__device__ __forceinline__ bool any_nonzero(const Dual &x) { return x.v != 0. || x.d != 0.; }

// eps grad Q at the used RT points, one thread per (element, variable)
// (CalculateEpsilonGradient dissipation.go:244-272); zero where eps is
// zero at every point of the element
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_grad(AdjArgs a, const T *U, const T *epsf, T *dis) {
  using Sk = AdjShock<N>;
  constexpr int NR = Sk::NR, NpFlux = Sk::NpFlux;
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(a.Kp) * 4) return;
  const int kk = int(gid >> 2), n = int(gid & 3);
  const size_t Kp = size_t(a.Kp);
  bool act = false;
  if (kk < a.K)
    for (int r = 0; r < NR; ++r) act = act || any_nonzero(epsf[r * Kp + kk]);
  if (!act) {
    for (int r = 0; r < NR; ++r) {
      dis[Sk::dis_at(a.Kp, n, 0, r, kk)] = T(0.);
      dis[Sk::dis_at(a.Kp, n, 1, r, kk)] = T(0.);
    }
    return;
  }
  T DX[NpFlux], DY[NpFlux];
  grad_dofs<N>(a, U, kk, n, DX, DY);
  const double *DR = a.shops + Sk::oDR;
  for (int r = 0; r < NR; ++r) {
    T ax = T(0.), ay = T(0.);
    for (int j = 0; j < NpFlux; ++j) {
      ax += DR[r * NpFlux + j] * DX[j];
      ay += DR[r * NpFlux + j] * DY[j];
    }
    const T e = epsf[r * Kp + kk];
    dis[Sk::dis_at(a.Kp, n, 0, r, kk)] = ax * e;
    dis[Sk::dis_at(a.Kp, n, 1, r, kk)] = ay * e;
  }
}

// The viscous edge flux (StoreEdgeViscousFlux edges.go:151-244) taken off
// the edge values: at a shared edge's owner point i, the mean of both
// sides' eps grad Q along the owner normal minus the penalty
// N^2 mean(eps) / L (Q_i - Q_rev i) on the owner's stored trace, written to
// both sides like the numerical flux; at an unshared point the own side's
// eps grad Q along its normal
template <int N, class T>
__global__ void __launch_bounds__(256) k_adj_fwd_visc(AdjArgs a, const T *U, const T *epsf, const T *dis, T *FE) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt, NpEdge = Sk::NpEdge;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  if (kk >= a.K) return;
  const size_t Kp = size_t(a.Kp);
  const int code = a.code[fi];
  const double nx = a.Nx[fi], ny = a.Ny[fi];
  const int rowL = NpInt + f * NpEdge + pp;
  if ((code & 15) == FK_Shared) {
    if (!(code & FC_FLUX_OWNER)) return;  // written by the owner
    const int nb = a.nbr_elem[fi], nf = a.nbr_face[fi], rp = NpEdge - 1 - pp;
    const int rowR = NpInt + nf * NpEdge + rp;
    const size_t fn = size_t(nf) * Kp + nb, t2 = fn * NpEdge + rp;
    const T lam = 0.5 * (epsf[rowL * Kp + kk] + epsf[rowR * Kp + nb]);
    const T coef = (a.omega * lam) * a.vooL[fi];
    const double wo = a.IInII[fi], wn = a.IInII[fn];
    for (int n = 0; n < 4; ++n) {
      const T vFL = nx * dis[Sk::dis_at(a.Kp, n, 0, rowL, kk)] + ny * dis[Sk::dis_at(a.Kp, n, 1, rowL, kk)];
      const T vFR = nx * dis[Sk::dis_at(a.Kp, n, 0, rowR, nb)] + ny * dis[Sk::dis_at(a.Kp, n, 1, rowR, nb)];
      T VF = 0.5 * (vFL + vFR);
      VF -= coef * (owner_trace<N>(a, U, kk, f, pp, n) - owner_trace<N>(a, U, kk, f, rp, n));
      FE[4 * t + n] -= wo * VF;
      FE[4 * t2 + n] += wn * VF;
    }
    return;
  }
  const double w = a.IInII[fi];
  for (int n = 0; n < 4; ++n) {
    const T VF = nx * dis[Sk::dis_at(a.Kp, n, 0, rowL, kk)] + ny * dis[Sk::dis_at(a.Kp, n, 1, rowL, kk)];
    FE[4 * t + n] -= w * VF;
  }
}

// ---- adjoint of the active sensor path ----

// lam <- F(sigma)^T lam for the element's four variables (in place) and
// lamSigb[kk] (+)= sum_n lam_n . dF/dsigma X_n at the pre-filter values X
template <int N>
__global__ void __launch_bounds__(256) k_adj_filter_sig_t(AdjArgs a, double *lam, const double *X, const double *sigb,
                                                          double *lamSigb, int accumulate) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.K) return;
  const double *V = a.shops + Sk::oV, *Vi = a.shops + Sk::oVi, *mf = a.shops + Sk::omf;
  const double sb = sigb[kk];
  const double fac = 1. - sin((0.5 * M_PI) * sb), dfac = -(0.5 * M_PI) * cos((0.5 * M_PI) * sb);
  double dsum = 0.;
  for (int n = 0; n < 4; ++n) {
    double *l = lam + size_t(kk) * AdjShape<N>::NV + n * NpInt;
    const double *x = X + size_t(kk) * AdjShape<N>::NV + n * NpInt;
    double w[NpInt];
    for (int m = 0; m < NpInt; ++m) {
      double t = 0., mx = 0.;
      for (int i = 0; i < NpInt; ++i) t += V[i * NpInt + m] * l[i];
      for (int j = 0; j < NpInt; ++j) mx += Vi[m * NpInt + j] * x[j];
      if (m > 0) {
        dsum += t * mf[m] * dfac * mx;
        t *= mf[m] * fac;
      }
      w[m] = t;
    }
    for (int j = 0; j < NpInt; ++j) {
      double t = 0.;
      for (int m = 0; m < NpInt; ++m) t += Vi[m * NpInt + j] * w[m];
      l[j] = t;
    }
  }
  lamSigb[kk] = accumulate ? lamSigb[kk] + dsum : dsum;
}

// Cotangents of the viscous edge flux, one thread per face point; the
// owner point of a shared edge writes both sides' rows: eps grad Q at the
// edge rows (lamDis), eps at the edge rows from the penalty (lamEpsf), and
// its own stored trace's cotangent from the penalty jumps of its point and
// of the reversed one (lamEQv; zero at every other point)
template <int N>
__global__ void __launch_bounds__(256) k_adj_visc_t(AdjArgs a, const double *U, const double *lamR, const double *epsf,
                                                    double *lamDis, double *lamEpsf, double *lamEQv) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt, NpEdge = Sk::NpEdge;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  for (int c = 0; c < 4; ++c) lamEQv[4 * t + c] = 0.;
  if (kk >= a.K) return;
  const size_t Kp = size_t(a.Kp);
  const int code = a.code[fi];
  const double nx = a.Nx[fi], ny = a.Ny[fi];
  const int rowL = NpInt + f * NpEdge + pp;
  if ((code & 15) == FK_Shared) {
    if (!(code & FC_FLUX_OWNER)) return;
    const int nb = a.nbr_elem[fi], nf = a.nbr_face[fi], rp = NpEdge - 1 - pp;
    const int rowR = NpInt + nf * NpEdge + rp;
    const size_t fn = size_t(nf) * Kp + nb;
    const double wo = a.IInII[fi], wn = a.IInII[fn];
    // the edge values' cotangents at this point and at the reversed one
    double lfo[4], lfn[4], lfor[4], lfnr[4], lV[4], lVr[4];
    edge_cotangent<N>(a, lamR, kk, f * NpEdge + pp, lfo);
    edge_cotangent<N>(a, lamR, nb, nf * NpEdge + rp, lfn);
    edge_cotangent<N>(a, lamR, kk, f * NpEdge + rp, lfor);
    edge_cotangent<N>(a, lamR, nb, nf * NpEdge + pp, lfnr);
    for (int c = 0; c < 4; ++c) {
      lV[c] = -(wo * lfo[c] - wn * lfn[c]);
      lVr[c] = -(wo * lfor[c] - wn * lfnr[c]);
    }
    const double pc = a.omega * 0.5 * a.vooL[fi];
    const double coef = pc * (epsf[rowL * Kp + kk] + epsf[rowR * Kp + nb]);
    const double coefr = pc * (epsf[(NpInt + f * NpEdge + rp) * Kp + kk] + epsf[(NpInt + nf * NpEdge + pp) * Kp + nb]);
    double jump = 0.;
    for (int c = 0; c < 4; ++c) {
      const double qi = owner_trace<N>(a, U, kk, f, pp, c), qr = owner_trace<N>(a, U, kk, f, rp, c);
      jump += (qi - qr) * lV[c];
      lamDis[Sk::dis_at(a.Kp, c, 0, rowL, kk)] = 0.5 * nx * lV[c];
      lamDis[Sk::dis_at(a.Kp, c, 1, rowL, kk)] = 0.5 * ny * lV[c];
      lamDis[Sk::dis_at(a.Kp, c, 0, rowR, nb)] = 0.5 * nx * lV[c];
      lamDis[Sk::dis_at(a.Kp, c, 1, rowR, nb)] = 0.5 * ny * lV[c];
      lamEQv[4 * t + c] = -coef * lV[c] + coefr * lVr[c];
    }
    lamEpsf[rowL * Kp + kk] = -pc * jump;
    lamEpsf[rowR * Kp + nb] = -pc * jump;
    return;
  }
  const double w = a.IInII[fi];
  double lf[4];
  edge_cotangent<N>(a, lamR, kk, f * NpEdge + pp, lf);
  for (int c = 0; c < 4; ++c) {
    const double lv = -w * lf[c];
    lamDis[Sk::dis_at(a.Kp, c, 0, rowL, kk)] = nx * lv;
    lamDis[Sk::dis_at(a.Kp, c, 1, rowL, kk)] = ny * lv;
  }
  lamEpsf[rowL * Kp + kk] = 0.;
}

// Cotangents through eps grad Q, one thread per (element, variable), four
// consecutive lanes per element: the interior rows' cotangent from the
// interior flux (the flux is F - eps grad Q), the edge rows' from
// k_adj_visc_t; then eps's (summed over the variables: interior rows set,
// edge rows added), the interior state's (added to target) and the
// owner-stored traces' at the element's face points (lamEQg). Elements with
// eps zero everywhere take none (their eps cotangent is set to zero).
template <int N>
__global__ void __launch_bounds__(256) k_adj_grad_t(AdjArgs a, const double *U, const double *lamR, const double *epsf,
                                                    const double *lamDis, double *lamEpsf, double *lamEQg,
                                                    double *target) {
  using Sk = AdjShock<N>;
  constexpr int NR = Sk::NR, NpInt = Sk::NpInt, NpEdge = Sk::NpEdge, NpFlux = Sk::NpFlux;
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= size_t(a.Kp) * 4) return;  // Kp * 4 is a multiple of 64: whole element groups
  const int kk = int(gid >> 2), n = int(gid & 3);
  const size_t Kp = size_t(a.Kp);
  bool act = false;
  if (kk < a.K)
    for (int r = 0; r < NR; ++r) act = act || epsf[r * Kp + kk] != 0.;
  if (!act) {
    for (int f = 0; f < 3; ++f)
      for (int i = 0; i < NpEdge; ++i) lamEQg[4 * ((f * Kp + kk) * NpEdge + i) + n] = 0.;
    if (n == 0)
      for (int r = 0; r < NR; ++r) lamEpsf[r * Kp + kk] = 0.;
    return;
  }
  const double Jd = a.Jdet[kk], m = -(1. / Jd), ooJd = 1. / Jd;
  const double J0 = a.Jinv[kk], J1 = a.Jinv[Kp + kk], J2 = a.Jinv[2 * Kp + kk], J3 = a.Jinv[3 * Kp + kk];
  const double *lr = lamR + size_t(kk) * AdjShape<N>::NV + n * NpInt;
  const double *D = a.DivInt, *DR = a.shops + Sk::oDR;
  double DX[NpFlux], DY[NpFlux];
  grad_dofs<N>(a, U, kk, n, DX, DY);
  double gX[NpFlux], gY[NpFlux];  // cotangents of the gradient DOFs
  for (int j = 0; j < NpFlux; ++j) {
    gX[j] = 0.;
    gY[j] = 0.;
  }
  for (int r = 0; r < NR; ++r) {
    double lDX, lDY;
    if (r < NpInt) {
      // the interior flux's cotangent at point r (k_adj_elem), negated
      double lr_ = 0., ls_ = 0.;
      for (int i = 0; i < NpInt; ++i) {
        lr_ += D[i * NpFlux + r] * lr[i];
        ls_ += D[i * NpFlux + NpInt + r] * lr[i];
      }
      lr_ *= m;
      ls_ *= m;
      lDX = -(Jd * (J0 * lr_ + J2 * ls_));
      lDY = -(Jd * (J1 * lr_ + J3 * ls_));
    } else {
      lDX = lamDis[Sk::dis_at(a.Kp, n, 0, r, kk)];
      lDY = lamDis[Sk::dis_at(a.Kp, n, 1, r, kk)];
    }
    double ax = 0., ay = 0.;
    for (int j = 0; j < NpFlux; ++j) {
      ax += DR[r * NpFlux + j] * DX[j];
      ay += DR[r * NpFlux + j] * DY[j];
    }
    double le = lDX * ax + lDY * ay;  // eps's cotangent from this variable
    le += __shfl_xor(le, 1, 64);
    le += __shfl_xor(le, 2, 64);
    if (n == 0) lamEpsf[r * Kp + kk] = (r < NpInt ? 0. : lamEpsf[r * Kp + kk]) + le;
    const double e = epsf[r * Kp + kk], bx = e * lDX, by = e * lDY;
    for (int j = 0; j < NpFlux; ++j) {
      gX[j] += DR[r * NpFlux + j] * bx;
      gY[j] += DR[r * NpFlux + j] * by;
    }
  }
  double *tg = target + size_t(kk) * AdjShape<N>::NV + n * NpInt;
  for (int i = 0; i < NpInt; ++i) tg[i] += J0 * gX[i] + J2 * gX[i + NpInt] + J1 * gY[i] + J3 * gY[i + NpInt];
  for (int f = 0; f < 3; ++f) {
    const size_t fi = f * Kp + kk;
    const double nx = a.Nxo[fi], ny = a.Nyo[fi], w = a.IInII[fi];
    for (int i = 0; i < NpEdge; ++i) {
      const int j = 2 * NpInt + f * NpEdge + i;
      lamEQg[4 * ((f * Kp + kk) * NpEdge + i) + n] = ooJd * w * (nx * gX[j] + ny * gY[j]);
    }
  }
}

// The owner-stored traces' cotangents into the face-point trace cotangents
// of the edge owner: its viscous-flux term, its own gradient term and, for a
// shared edge, the partner's gradient term at the reversed point
template <int N>
__global__ void __launch_bounds__(256) k_adj_eq_gather(AdjArgs a, const double *lamEQv, const double *lamEQg,
                                                       double *lamOwn) {
  constexpr int NpEdge = AdjShape<N>::NpEdge;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= size_t(3) * a.Kp * NpEdge) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, t, f, kk, pp, fi);
  if (kk >= a.K) return;
  const int code = a.code[fi];
  const bool shared = (code & 15) == FK_Shared;
  if (shared && !(code & FC_FLUX_OWNER)) return;
  size_t t2 = 0;
  if (shared) t2 = (size_t(a.nbr_face[fi]) * a.Kp + a.nbr_elem[fi]) * NpEdge + (NpEdge - 1 - pp);
  for (int c = 0; c < 4; ++c) {
    double s = lamEQv[4 * t + c] + lamEQg[4 * t + c];
    if (shared) s += lamEQg[4 * t2 + c];
    lamOwn[4 * t + c] += s;
  }
}

// The viscous dt term where it set the stage's dt (dt = Cdiff / Wv at the
// arg-max owner face point): into eps at that point's RT row. One thread.
template <int N>
__global__ void k_adj_vdt(AdjArgs a, int stage, const int *vbr, const int *argmaxv, const double *Wvv,
                          const double *lam_dt, double *lamEpsf) {
  using Sh = AdjShape<N>;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (!vbr[stage]) return;
  int f, kk, pp;
  size_t fi;
  face_point<N>(a, size_t(argmaxv[stage]), f, kk, pp, fi);
  const double W = Wvv[stage];
  const double lW = -(a.cdiff / (W * W)) * lam_dt[stage];
  const double oohK = a.oohK[kk];
  lamEpsf[size_t(Sh::NpInt + f * Sh::NpEdge + pp) * a.Kp + kk] += (oohK * oohK) * lW;
}

// Per vertex: the cotangents of its eps and sigma maxima, summed over its
// elements in incidence order (the transposes of the barycentric
// interpolation and of the vertex mean)
template <int N>
__global__ void __launch_bounds__(256) k_adj_vert_t(AdjArgs a, const double *lamEpsf, const double *lamSigb,
                                                    double *lamEpsv, double *lamSigv) {
  using Sk = AdjShock<N>;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= a.Nv) return;
  const double *B = a.shops + Sk::oBa;
  double le = 0., ls = 0.;
  for (int q = a.vptr[v]; q < a.vptr[v + 1]; ++q) {
    const int kk = a.vinc[q] >> 2, lv = a.vinc[q] & 3;
    for (int r = 0; r < Sk::NR; ++r) le += B[3 * r + lv] * lamEpsf[size_t(r) * a.Kp + kk];
    ls += lamSigb[kk] / 3.;
  }
  lamEpsv[v] = le;
  lamSigv[v] = ls;
}

// Per element: the cotangents of the vertex maxima it attains, through
// eps = eps0 hK sigma and the ramp to Se and the density of the sensor's
// input (the stage input before the stage-2 filter), added to target
template <int N>
__global__ void __launch_bounds__(256) k_adj_sens_t(AdjArgs a, const double *Us, const int *args, const int *arge,
                                                    const double *lamEpsv, const double *lamSigv, double *target) {
  using Sk = AdjShock<N>;
  constexpr int NpInt = Sk::NpInt;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  if (kk >= a.K) return;
  const size_t Kp = size_t(a.Kp);
  double le = 0., ls = 0.;
  for (int c = 0; c < 3; ++c) {
    const int v = a.evert[c * Kp + kk];
    if (arge[v] == kk) le += lamEpsv[v];
    if (args[v] == kk) ls += lamSigv[v];
  }
  ls += (a.eps0 * a.hK[kk]) * le;
  if (ls == 0.) return;
  const double *rho = Us + size_t(kk) * AdjShape<N>::NV;
  const double *Pm = a.shops + Sk::oP, *Mm = a.shops + Sk::oM, *Dm = a.shops + Sk::oD;
  double X[NpInt], Y[NpInt], DQ[NpInt], num = 0., den = 0.;
  for (int i = 0; i < NpInt; ++i) {
    double x = 0., y = 0., dq = 0.;
    for (int j = 0; j < NpInt; ++j) {
      x += Pm[i * NpInt + j] * rho[j];
      y += Mm[i * NpInt + j] * rho[j];
      dq += Dm[i * NpInt + j] * rho[j];
    }
    X[i] = x;
    Y[i] = y;
    DQ[i] = dq;
    num += dq * x;
    den += rho[i] * y;
  }
  const double slope = sigma_slope(a, log10(num / den));
  if (slope == 0.) return;
  const double lse = slope * ls / 2.302585092994045684;
  double *tg = target + size_t(kk) * AdjShape<N>::NV;  // variable 0: the density
  for (int j = 0; j < NpInt; ++j) {
    double gn = 0., gd = Y[j];
    for (int i = 0; i < NpInt; ++i) {
      gn += Dm[i * NpInt + j] * X[i] + DQ[i] * Pm[i * NpInt + j];
      gd += rho[i] * Mm[i * NpInt + j];
    }
    tg[j] += lse * (gn / num - gd / den);
  }
}

inline unsigned adj_blocks(size_t n) { return unsigned((n + 255) / 256); }

template <int N>
hipError_t adj_launch_n(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s) {
  using Sh = AdjShape<N>;
  const size_t nfp = size_t(3) * a.Kp * Sh::NpEdge;
  switch (op) {
    case ADJ_FWD_FACE:
#define GOCFD_FWD_FACE(FL)                                                                                 \
  if (b.dual)                                                                                              \
    hipLaunchKernelGGL((k_adj_fwd_face<N, Dual, FL>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,           \
                       static_cast<const Dual *>(b.Us), static_cast<Dual *>(b.FE), static_cast<Dual *>(b.SP)); \
  else                                                                                                     \
    hipLaunchKernelGGL((k_adj_fwd_face<N, double, FL>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,         \
                       static_cast<const double *>(b.Us), static_cast<double *>(b.FE),                     \
                       static_cast<double *>(b.SP));
      switch (a.ph.flux) {
        case 0: GOCFD_FWD_FACE(0) break;
        case 1: GOCFD_FWD_FACE(1) break;
        case 2: GOCFD_FWD_FACE(2) break;
        default: GOCFD_FWD_FACE(3) break;
      }
#undef GOCFD_FWD_FACE
      break;
    case ADJ_FWD_RHS: {
      const unsigned nb = unsigned((a.Kp + RhsElems<N>::value - 1) / RhsElems<N>::value);
      if (b.dual)
        hipLaunchKernelGGL((k_adj_fwd_rhs<N, Dual>), dim3(nb), dim3(256), 0, s, a, b.stage,
                           static_cast<const Dual *>(b.Us), static_cast<const Dual *>(b.FE),
                           static_cast<Dual *>(b.R), static_cast<const Dual *>(b.U0),
                           static_cast<const Dual *>(b.U2), static_cast<const Dual *>(b.U3),
                           static_cast<const Dual *>(b.R3), static_cast<const Dual *>(b.dts),
                           static_cast<const Dual *>(b.dtk), static_cast<Dual *>(b.out),
                           static_cast<const Dual *>(b.dis), static_cast<const Dual *>(b.sigb), nullptr);
      else
        hipLaunchKernelGGL((k_adj_fwd_rhs<N, double>), dim3(nb), dim3(256), 0, s, a, b.stage,
                           static_cast<const double *>(b.Us), static_cast<const double *>(b.FE),
                           static_cast<double *>(b.R), static_cast<const double *>(b.U0),
                           static_cast<const double *>(b.U2), static_cast<const double *>(b.U3),
                           static_cast<const double *>(b.R3), static_cast<const double *>(b.dts),
                           static_cast<const double *>(b.dtk), static_cast<double *>(b.out),
                           static_cast<const double *>(b.dis), static_cast<const double *>(b.sigb), b.Rpre);
      break;
    }
    case ADJ_FWD_LTS:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_fwd_lts<N, Dual>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a, b.stage,
                           static_cast<const Dual *>(b.SP), static_cast<const Dual *>(b.dtk_prev),
                           static_cast<Dual *>(b.dtk), b.srck, b.wk);
      else
        hipLaunchKernelGGL((k_adj_fwd_lts<N, double>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a, b.stage,
                           static_cast<const double *>(b.SP), static_cast<const double *>(b.dtk_prev),
                           static_cast<double *>(b.dtk), b.srck, b.wk);
      break;
    case ADJ_ADJ_SSP_LTS:
      hipLaunchKernelGGL((k_adj_ssp_lts<N>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a, b.stage, b.lam, b.lamU0,
                         b.lamQ[1], b.lamQ[2], b.lamQ[3], b.lamQ[4], b.lamR, b.lamR3,
                         static_cast<const double *>(b.Rs), static_cast<const double *>(b.R3),
                         static_cast<const double *>(b.dtk), b.srck, b.wk, b.carry, b.lamSpeed);
      break;
    case ADJ_ADJ_LTS_WAVE:
      hipLaunchKernelGGL((k_adj_lts_wave<N>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamSpeed, b.lamOwn, b.lamFarPts);
      break;
    case ADJ_SENSOR:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_sensor<N, Dual>), dim3(adj_blocks(a.K)), dim3(256), 0, s, a,
                           static_cast<const Dual *>(b.Us), b.flag);
      else
        hipLaunchKernelGGL((k_adj_sensor<N, double>), dim3(adj_blocks(a.K)), dim3(256), 0, s, a,
                           static_cast<const double *>(b.Us), b.flag);
      break;
    case ADJ_FILTER:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_filter<N, Dual>), dim3(adj_blocks(size_t(a.K) * 4)), dim3(256), 0, s, a,
                           static_cast<Dual *>(b.out), 0);
      else
        hipLaunchKernelGGL((k_adj_filter<N, double>), dim3(adj_blocks(size_t(a.K) * 4)), dim3(256), 0, s, a,
                           static_cast<double *>(b.out), 0);
      break;
    case ADJ_FILTER_T:
      hipLaunchKernelGGL((k_adj_filter<N, double>), dim3(adj_blocks(size_t(a.K) * 4)), dim3(256), 0, s, a, b.lam, 1);
      break;
    case ADJ_ADJ_FACE:
#define GOCFD_ADJ_FACE(FL)                                                                                 \
  hipLaunchKernelGGL((k_adj_face<N, FL>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,                        \
                     static_cast<const double *>(b.Us), b.lamR, b.lamOwn);
      switch (a.ph.flux) {
        case 0: GOCFD_ADJ_FACE(0) break;
        case 1: GOCFD_ADJ_FACE(1) break;
        case 2: GOCFD_ADJ_FACE(2) break;
        default: GOCFD_ADJ_FACE(3) break;
      }
#undef GOCFD_ADJ_FACE
      if (a.nbnd > 0)
        hipLaunchKernelGGL((k_adj_face_bnd<N>), dim3(adj_blocks(8 * size_t(a.nbnd))), dim3(256), 0, s, a,
                           static_cast<const double *>(b.Us), b.lamR, b.lamOwn, b.lamFarPts);
      break;
    case ADJ_ADJ_ELEM:
      hipLaunchKernelGGL((k_adj_elem<N>), dim3(adj_blocks(size_t(a.K) * Sh::NpInt)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamR, b.lamOwn, b.target);
      break;
    case ADJ_FWD_SHOCK:
#define GOCFD_FWD_SHOCK(TT)                                                                                       \
  hipLaunchKernelGGL((k_adj_fwd_sens<N, TT>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a,                          \
                     static_cast<const TT *>(b.Us), static_cast<TT *>(b.sig), static_cast<TT *>(b.eps));            \
  hipLaunchKernelGGL((k_adj_fwd_vmax<N, TT>), dim3(adj_blocks(a.Nv)), dim3(256), 0, s, a,                          \
                     static_cast<const TT *>(b.sig), static_cast<const TT *>(b.eps), static_cast<TT *>(b.sigv),     \
                     static_cast<TT *>(b.epsv), b.args, b.arge);                                                    \
  hipLaunchKernelGGL((k_adj_fwd_efield<N, TT>), dim3(adj_blocks(a.Kp)), dim3(256), 0, s, a,                        \
                     static_cast<const TT *>(b.sigv), static_cast<const TT *>(b.epsv), static_cast<TT *>(b.sigb),   \
                     static_cast<TT *>(b.epsf), static_cast<TT *>(b.SPV));
      if (b.dual) {
        GOCFD_FWD_SHOCK(Dual)
      } else {
        GOCFD_FWD_SHOCK(double)
      }
#undef GOCFD_FWD_SHOCK
      break;
    case ADJ_FILTER_SIG:
      if (b.dual)
        hipLaunchKernelGGL((k_adj_filter_sig<N, Dual>), dim3(adj_blocks(size_t(a.K) * 4)), dim3(256), 0, s, a,
                           static_cast<Dual *>(b.out), static_cast<const Dual *>(b.sigb));
      else
        hipLaunchKernelGGL((k_adj_filter_sig<N, double>), dim3(adj_blocks(size_t(a.K) * 4)), dim3(256), 0, s, a,
                           static_cast<double *>(b.out), static_cast<const double *>(b.sigb));
      break;
    case ADJ_FWD_VISC:
#define GOCFD_FWD_VISC(TT)                                                                                        \
  hipLaunchKernelGGL((k_adj_fwd_grad<N, TT>), dim3(adj_blocks(size_t(a.Kp) * 4)), dim3(256), 0, s, a,              \
                     static_cast<const TT *>(b.Us), static_cast<const TT *>(b.epsf), static_cast<TT *>(b.dis));     \
  hipLaunchKernelGGL((k_adj_fwd_visc<N, TT>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,                           \
                     static_cast<const TT *>(b.Us), static_cast<const TT *>(b.epsf), static_cast<const TT *>(b.dis),\
                     static_cast<TT *>(b.FE));
      if (b.dual) {
        GOCFD_FWD_VISC(Dual)
      } else {
        GOCFD_FWD_VISC(double)
      }
#undef GOCFD_FWD_VISC
      break;
    case ADJ_ADJ_FILTER_SIG_T:
      hipLaunchKernelGGL((k_adj_filter_sig_t<N>), dim3(adj_blocks(a.K)), dim3(256), 0, s, a, b.lam, b.Xpre,
                         static_cast<const double *>(b.sigb), b.lamSigb, b.accumulate);
      break;
    case ADJ_ADJ_VISC:
      hipLaunchKernelGGL((k_adj_visc_t<N>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamR, static_cast<const double *>(b.epsf), b.lamDis,
                         b.lamEpsf, b.lamEQv);
      hipLaunchKernelGGL((k_adj_grad_t<N>), dim3(adj_blocks(size_t(a.Kp) * 4)), dim3(256), 0, s, a,
                         static_cast<const double *>(b.Us), b.lamR, static_cast<const double *>(b.epsf), b.lamDis,
                         b.lamEpsf, b.lamEQg, b.target);
      hipLaunchKernelGGL((k_adj_eq_gather<N>), dim3(adj_blocks(nfp)), dim3(256), 0, s, a, b.lamEQv, b.lamEQg,
                         b.lamOwn);
      break;
    case ADJ_ADJ_VDT:
      hipLaunchKernelGGL((k_adj_vdt<N>), dim3(1), dim3(64), 0, s, a, b.stage, b.vbr, b.argmaxv, b.Wvv, b.lam_dt,
                         b.lamEpsf);
      break;
    case ADJ_ADJ_SENSOR:
      hipLaunchKernelGGL((k_adj_vert_t<N>), dim3(adj_blocks(a.Nv)), dim3(256), 0, s, a, b.lamEpsf, b.lamSigb,
                         b.lamEpsv, b.lamSigv);
      hipLaunchKernelGGL((k_adj_sens_t<N>), dim3(adj_blocks(a.K)), dim3(256), 0, s, a, b.Usens, b.args, b.arge,
                         b.lamEpsv, b.lamSigv, b.target);
      break;
    case ADJ_ADJ_WAVE:
      hipLaunchKernelGGL((k_adj_wave<N>), dim3(1), dim3(64), 0, s, a, b.stage, static_cast<const double *>(b.Us),
                         b.argmax, b.Wv, b.lam_dt, b.target, b.lamFarW, a.act ? b.vbr : nullptr);
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
