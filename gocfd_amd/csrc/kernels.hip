// gfx950 kernels for the 2D Euler DFR right-hand side and SSP54 stages.
//
// One launch per Runge-Kutta stage. A workgroup owns a tile of kTile
// consecutive elements (Hilbert order), whose element-major state is one
// contiguous chunk of HBM. Per stage and tile:
//   A. load the stage state into LDS; face traces by the FluxEdgeInterp
//      contraction (model_problems/Euler2D/edges.go:485-491) for the tile's
//      faces and for the out-of-tile neighbour faces listed in the halo table
//   B. Riemann / boundary flux at every face point (edges.go:325-452,
//      bcs.go:11-133), scaled by ||n|| (SetRTFluxOnEdges edges.go:454-483)
//   C. interior physical flux, contravariant transform and the DivInt
//      contraction (SetRTFluxInternal euler.go:701-726, RHSInternalPoints
//      euler.go:665-699)
//   E. SSP54 combination (euler.go:502-565) with the global dt of this stage
//   F/G. traces of the new state, boundary states at the next stage's time
//      and the owner-side CFL wave-speed maximum (edges.go:246-323), folded
//      into one atomic per workgroup for the next stage's dt
//      (calculateGlobalDT euler.go:945-971)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dfr_setup.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace gocfd {
namespace dev {

// Reference-element operators of every order 0..4 in constant memory, so
// that every coefficient is a uniform scalar load: per order, DivInt^T
// [NpFlux][NpInt] followed by FluxEdgeInterp [3 NpEdge][NpInt]. They depend
// on the order only, so all solvers on a device share them.
constexpr int ops_size(int N) {
  return ((N + 2) * (N + 4) + 3 * (N + 2)) * ((N + 1) * (N + 2) / 2);
}
constexpr int ops_offset(int N) { return N == 0 ? 0 : ops_offset(N - 1) + ops_size(N - 1); }
constexpr int kOpsTotal = ops_offset(5);
__constant__ double cOpsD[kOpsTotal];
__constant__ float cOpsF[kOpsTotal];
template <class C>
__device__ __forceinline__ const C *ops_base();
template <>
__device__ __forceinline__ const double *ops_base<double>() { return cOpsD; }
template <>
__device__ __forceinline__ const float *ops_base<float>() { return cOpsF; }

template <int N>
struct Shape {
  static constexpr int NpInt = (N + 1) * (N + 2) / 2;
  static constexpr int NpEdge = N + 2;
  static constexpr int NpFlux = (N + 2) * (N + 4);
  static constexpr int NF = 3 * NpEdge;  // face points per element
  static constexpr int NV = 4 * NpInt;   // state values per element
};

// Tile index of a workgroup: with workgroups dealt round-robin to the 8 XCDs,
// XCD x processes the contiguous tile range [x*q, (x+1)*q), so neighbouring
// tiles share an L2.
__device__ __forceinline__ int tile_of_block() {
  const int q = gridDim.x >> 3;
  return (blockIdx.x & 7) * q + (blockIdx.x >> 3);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <class C>
__device__ __forceinline__ void boundary_state(int kind, const Phys &ph, const double *bxy, int bidx,
                                               int NpEdge, int pp, double t, const C own[4],
                                               double nx, double ny, double qb[4]) {
  double qi[4] = {double(own[0]), double(own[1]), double(own[2]), double(own[3])};
  double qf[4];
  if (kind == FK_Far) {
    qf[0] = ph.qinf[0]; qf[1] = ph.qinf[1]; qf[2] = ph.qinf[2]; qf[3] = ph.qinf[3];
  } else {
    const double *xy = bxy + 2 * (size_t(bidx) * NpEdge + pp);
    ivortex_state(t, xy[0], xy[1], qf);
  }
  riemann_bc(ph.gamma, ph.minf, ph.pinf, ph.cinf, qi, qf, nx, ny, qb);
}

// Wave speed (|U| + c) / hK of a face point (StoreEdgeAggregates, edges.go:264-272)
__device__ __forceinline__ double point_speed(double g, const double q[4], double oohK) {
  const double oorho = 1. / q[0], u = q[1] * oorho, v = q[2] * oorho;
  const double U = sqrt(u * u + v * v);
  const double c = sound_speed(g, q);
  return oohK * (U + c);
}

template <int N, class S, class C, int FLUX, bool LAST>
__global__ void __launch_bounds__(kThreads) k_stage(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NF = Sh::NF, NV = Sh::NV;
  constexpr int E = kTile, P = kTile + 1;
  static_assert(kThreads == 4 * kTile, "one thread per (element, variable)");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C *sQ = reinterpret_cast<C *>(smem_raw);  // [NV][P]
  const int TS = NF * E + NpEdge * a.Hmax;
  C *sT = sQ + NV * P;                       // [4][TS]: traces, then fluxes
  C *sR = sT;                                // [NV][P]: RHS (aliases sT)
  __shared__ double red[kThreads / 64];

  const int tile = tile_of_block();
  if (tile >= a.ntiles) return;
  const int tid = threadIdx.x;
  const bool rhs_only = a.stage < 0;
  const int par = a.step_index & 1;
  Ctl *ctl = a.ctl;
  const double time = rhs_only ? a.t_rhs : ctl->time[par];
  // CheckIfFinished (euler.go:796-801): no further steps once Time >= FinalTime
  if (!rhs_only && a.step_index > 0 && time >= a.ph.final_time) {
    if (a.stage == 4 && blockIdx.x == 0 && threadIdx.x == 0) ctl->time[par ^ 1] = time;
    return;
  }
  const int k0 = tile * E;
  const int Kp = a.Kp;
  const S *qs = static_cast<const S *>(a.qs);
  const C *DvT = ops_base<C>() + ops_offset(N);         // DivInt^T [NpFlux][NpInt]
  const C *Itp = DvT + Sh::NpFlux * NpInt;              // FluxEdgeInterp [NF][NpInt]
  const C g = C(a.ph.gamma);

  // ---- A: tile state -> LDS (coalesced, contiguous chunk) ----
  // Every load is issued before the first LDS store, so a thread's LPT
  // loads are in flight together.
  constexpr int LPT = (E * NV + kThreads - 1) / kThreads;
  {
    const S *qt = qs + size_t(k0) * NV;
    S v[LPT];
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      v[j] = (idx < E * NV) ? qt[idx] : S(0);
    }
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      if (idx < E * NV) {
        const int e = idx / NV, r = idx - e * NV;
        sQ[r * P + e] = C(v[j]);
      }
    }
  }
  __syncthreads();
  const int te = tid % E, tn = tid / E;  // (element, variable) of this thread
  {
    C q[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) q[i] = sQ[(tn * NpInt + i) * P + te];
#pragma unroll 2
    for (int p = 0; p < NF; ++p) {
      C t = C(0);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * q[i];
      sT[tn * TS + p * E + te] = t;
    }
  }
  for (int it = tid; it < a.Hmax * 4; it += kThreads) {
    const int h = it % a.Hmax, n = it / a.Hmax;
    const int kh = a.g.halo_elem[size_t(tile) * a.Hmax + h];
    if (kh < 0) continue;
    const int f = a.g.halo_face[size_t(tile) * a.Hmax + h];
    const S *qn = qs + size_t(kh) * NV + n * NpInt;
    C q[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) q[i] = C(qn[i]);
#pragma unroll
    for (int pp = 0; pp < NpEdge; ++pp) {
      const C *row = Itp + (f * NpEdge + pp) * NpInt;
      C t = C(0);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) t += row[i] * q[i];
      sT[n * TS + NF * E + h * NpEdge + pp] = t;
    }
  }
  __syncthreads();

  // ---- B: numerical flux at every face point ----
  // Shared faces first (unrolled, results held in registers until every
  // neighbour trace has been read), then boundary faces in a rolled loop: a
  // boundary face point reads and writes only its own LDS slot.
  constexpr int ITB = (E * NF + kThreads - 1) / kThreads;
  C fl[ITB][4];
#pragma unroll
  for (int j = 0; j < ITB; ++j) {
    const int it = tid + j * kThreads;
    if (it < E * NF) {
      const int e = it % E, p = it / E, f = p / NpEdge, pp = p - f * NpEdge;
      const size_t fi = size_t(f) * Kp + k0 + e;
      const int code = a.g.code[fi];
      if ((code & 15) == FK_Shared) {
        C own[4], oth[4], F[4];
        const unsigned slot = unsigned(a.g.nbr_slot[fi]);
        const int base = int(slot & 0x7fffffffu);
        const int stride = (slot >> 31) ? 1 : E;
        const int idx = base + (NpEdge - 1 - pp) * stride;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          own[c] = sT[c * TS + p * E + e];
          oth[c] = sT[c * TS + idx];
        }
        const C nx = C(a.g.Nx[fi]), ny = C(a.g.Ny[fi]);
        const C w = C(a.g.IInII[fi]);
        if (code & FC_FLUX_OWNER) {
          numflux(FLUX, g, own, oth, nx, ny, F);
#pragma unroll
          for (int c = 0; c < 4; ++c) fl[j][c] = F[c] * w;
        } else {
          numflux(FLUX, g, oth, own, nx, ny, F);
#pragma unroll
          for (int c = 0; c < 4; ++c) fl[j][c] = -F[c] * w;
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ITB; ++j) {
    const int it = tid + j * kThreads;
    if (it < E * NF) {
      const int e = it % E, p = it / E, f = p / NpEdge;
      if ((a.g.code[size_t(f) * Kp + k0 + e] & 15) == FK_Shared) {
#pragma unroll
        for (int c = 0; c < 4; ++c) sT[c * TS + p * E + e] = fl[j][c];
      }
    }
  }
#pragma unroll 1
  for (int it = tid; it < E * NF; it += kThreads) {
    const int e = it % E, p = it / E, f = p / NpEdge, pp = p - f * NpEdge;
    const size_t fi = size_t(f) * Kp + k0 + e;
    const int code = a.g.code[fi];
    const int kind = code & 15;
    if (kind == FK_Shared) continue;
    C own[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) own[c] = sT[c * TS + p * E + e];
    const double nxd = a.g.Nx[fi], nyd = a.g.Ny[fi];
    const C nx = C(nxd), ny = C(nyd);
    C F[4] = {C(0), C(0), C(0), C(0)};
    if (kind == FK_Far || kind == FK_IVortex) {
      double qb[4];
      boundary_state(kind, a.ph, a.g.bxy, code >> 8, NpEdge, pp, time, own, nxd, nyd, qb);
      const C q[4] = {C(qb[0]), C(qb[1]), C(qb[2]), C(qb[3])};
      C Fx[4], Fy[4];
      flux_xy(g, q, Fx, Fy);
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = nx * Fx[c] + ny * Fy[c];
    } else if (kind == FK_Wall) {
      const C pr = pressure(g, own[0], own[1], own[2], own[3]);
      F[1] = nx * pr;
      F[2] = ny * pr;
    } else if (kind == FK_Open) {
      C Fx[4], Fy[4];
      flux_xy(g, own, Fx, Fy);
#pragma unroll
      for (int c = 0; c < 4; ++c) F[c] = nx * Fx[c] + ny * Fy[c];
    }
    const C w = C(a.g.IInII[fi]);
#pragma unroll
    for (int c = 0; c < 4; ++c) sT[c * TS + p * E + e] = F[c] * w;
  }
  __syncthreads();

  // ---- C: interior flux + divergence, thread (element te, variable tn) ----
  {
    const int kk = k0 + te;
    const double Jd = a.g.Jdet[kk];
    const C Jdc = C(Jd);
    const C J0 = C(a.g.Jinv[kk]), J1 = C(a.g.Jinv[size_t(Kp) + kk]);
    const C J2 = C(a.g.Jinv[2 * size_t(Kp) + kk]), J3 = C(a.g.Jinv[3 * size_t(Kp) + kk]);
    C acc[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) acc[i] = C(0);
#pragma unroll 1
    for (int j = 0; j < NpInt; ++j) {
      const C r0 = sQ[(0 * NpInt + j) * P + te], r1 = sQ[(1 * NpInt + j) * P + te];
      const C r2 = sQ[(2 * NpInt + j) * P + te], r3 = sQ[(3 * NpInt + j) * P + te];
      C Fx, Fy;
      flux_component(tn, g, r0, r1, r2, r3, Fx, Fy);
      const C Fr = Jdc * (J0 * Fx + J1 * Fy), Fs = Jdc * (J2 * Fx + J3 * Fy);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) acc[i] += DvT[j * NpInt + i] * Fr + DvT[(NpInt + j) * NpInt + i] * Fs;
    }
#pragma unroll 2
    for (int p = 0; p < NF; ++p) {
      const C Fe = sT[tn * TS + p * E + te];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) acc[i] += DvT[(2 * NpInt + p) * NpInt + i] * Fe;
    }
    const C moojd = -C(1. / Jd);
    __syncthreads();  // every thread has read its fluxes; sR may now overwrite sT
#pragma unroll
    for (int i = 0; i < NpInt; ++i) sR[(tn * NpInt + i) * P + te] = acc[i] * moojd;
  }
  __syncthreads();

  if (rhs_only) {
    double *out = a.rhs_out + size_t(k0) * NV;
    for (int idx = tid; idx < E * NV; idx += kThreads) {
      const int e = idx / NV, r = idx - e * NV;
      out[idx] = double(sR[r * P + e]);
    }
    return;
  }

  // ---- E: SSP54 stage combination (euler.go:511-565), FP64 arithmetic ----
  // The operands of this stage are loaded for all LPT values of the thread
  // first, then combined and stored.
  const int stage = a.stage;
  double dt = a.ph.cfl / __longlong_as_double((long long)ctl->wmax[a.ring]);
  if (time + dt > a.ph.final_time) dt = a.ph.final_time - time;
  int bad = 0;
  {
    const size_t base = size_t(k0) * NV;
    S *q0 = static_cast<S *>(a.q0) + base;
    S *q1 = static_cast<S *>(a.q1) + base;
    S *q2 = static_cast<S *>(a.q2) + base;
    S *q3 = static_cast<S *>(a.q3) + base;
    S *q4 = static_cast<S *>(a.q4) + base;
    S *Rr = static_cast<S *>(a.R) + base;
    const int nreal = min(E, a.K - k0);
    if constexpr (LAST) {
      // stage 4: R = -q0 + a q2 + b q3 + c q4 + d dt R3 + e dt RHS, q0 += R
      S v0[LPT], v2[LPT], v3[LPT], vR[LPT];
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int idx = tid + j * kThreads;
        const bool ok = idx < E * NV;
        v0[j] = ok ? q0[idx] : S(0);
        v2[j] = ok ? q2[idx] : S(0);
        v3[j] = ok ? q3[idx] : S(0);
        vR[j] = ok ? Rr[idx] : S(0);
      }
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int idx = tid + j * kThreads;
        if (idx >= E * NV) continue;
        const int e = idx / NV, r = idx - e * NV;
        const double dtR = dt * double(sR[r * P + e]);
        const double qsv = double(sQ[r * P + e]);
        const double q0v = double(v0[j]);
        const double dtR3 = dt * double(vR[j]);
        const double Rn = -q0v + 0.517231671970585 * double(v2[j]) + 0.096059710526146 * double(v3[j]) +
                          0.386708617503269 * qsv + 0.063692468666290 * dtR3 + 0.226007483236906 * dtR;
        Rr[idx] = S(Rn);
        const S nv = S(q0v + Rn);
        q0[idx] = nv;
        sQ[r * P + e] = C(nv);
        if (e < nreal && !(nv == nv)) bad = 1;
      }
    } else {
      S v0[LPT];
      if (stage >= 1) {
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
          const int idx = tid + j * kThreads;
          v0[j] = (idx < E * NV) ? q0[idx] : S(0);
        }
      }
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int idx = tid + j * kThreads;
        if (idx >= E * NV) continue;
        const int e = idx / NV, r = idx - e * NV;
        const double rhs = double(sR[r * P + e]);
        const double qsv = double(sQ[r * P + e]);
        const double dtR = dt * rhs;
        S nv;
        switch (stage) {
          case 0:
            nv = S(qsv + 0.391752226571890 * dtR);
            q1[idx] = nv;
            break;
          case 1:
            nv = S(0.444370493651235 * double(v0[j]) + 0.555629506348765 * qsv + 0.368410593050371 * dtR);
            q2[idx] = nv;
            break;
          case 2:
            nv = S(0.620101851488403 * double(v0[j]) + 0.379898148511597 * qsv + 0.251891774271694 * dtR);
            q3[idx] = nv;
            break;
          default:
            Rr[idx] = S(rhs);
            nv = S(0.178079954393132 * double(v0[j]) + 0.821920045606868 * qsv + 0.544974750228521 * dtR);
            q4[idx] = nv;
            break;
        }
        sQ[r * P + e] = C(nv);
        if (e < nreal && !(nv == nv)) bad = 1;
      }
    }
  }
  __syncthreads();

  // ---- F: traces of the new state (own faces) ----
  {
    C q[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) q[i] = sQ[(tn * NpInt + i) * P + te];
#pragma unroll 2
    for (int p = 0; p < NF; ++p) {
      C t = C(0);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * q[i];
      sT[tn * TS + p * E + te] = t;
    }
  }
  __syncthreads();

  // ---- G: owner-side wave speed of the next stage, max over the tile ----
  const double tnext = LAST ? time + dt : time;
  double wl = 0.0;
  {
    const int nw = a.g.wave_cnt[tile] * NpEdge;
    const int *wlist = a.g.wave_list + size_t(tile) * a.Wmax;
#pragma unroll 1
    for (int it = tid; it < nw; it += kThreads) {
      const int w = it / NpEdge, pp = it - w * NpEdge;
      const int packed = wlist[w];
      const int e = packed & 255, f = packed >> 8, p = f * NpEdge + pp;
      const int kk = k0 + e;
      const size_t fi = size_t(f) * Kp + kk;
      const int code = a.g.code[fi];
      const int kind = code & 15;
      C own[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) own[c] = sT[c * TS + p * E + e];
      double q[4] = {double(own[0]), double(own[1]), double(own[2]), double(own[3])};
      if (kind == FK_Far || kind == FK_IVortex)
        boundary_state(kind, a.ph, a.g.bxy, code >> 8, NpEdge, pp, tnext, own, a.g.Nx[fi], a.g.Ny[fi], q);
      const double wsp = point_speed(a.ph.gamma, q, a.g.oohK[kk]);
      if (!(wsp == wsp)) bad = 1;
      wl = fmax(wl, wsp);
    }
  }
  wl = wave_max(wl);
  if ((tid & 63) == 0) red[tid >> 6] = wl;
  bad = __syncthreads_or(bad);
  if (tid == 0) {
    double m = red[0];
    for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, red[w]);
    atomicMax(&ctl->wmax[(a.ring + 1) % 3], (unsigned long long)__double_as_longlong(m));
    if (bad) atomicOr(&ctl->nan, 1);
    if (blockIdx.x == 0) {
      ctl->wmax[(a.ring + 2) % 3] = 0ull;
      ctl->dt[stage] = dt;
      if (LAST) {
        ctl->time[par ^ 1] = tnext;
        ctl->steps += 1;
      }
    }
  }
}

// Wave speed of the current state (start of a step sequence): one thread per
// (element, face point), traces straight from HBM.
template <int N, class S>
__global__ void __launch_bounds__(256) k_prologue(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NF = Sh::NF, NV = Sh::NV;
  __shared__ double red[4];
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = it / NF, p = it % NF;
  double wl = 0.0;
  int bad = 0;
  const double *Itp = a.Interp64;
  if (kk < a.K) {
    const int f = p / NpEdge, pp = p - f * NpEdge;
    const size_t fi = size_t(f) * a.Kp + kk;
    const int code = a.g.code[fi];
    if (code & FC_WAVE) {
      const S *q = static_cast<const S *>(a.qs) + size_t(kk) * NV;
      double tr[4];
      for (int n = 0; n < 4; ++n) {
        double t = 0;
        for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * double(q[n * NpInt + i]);
        tr[n] = t;
      }
      const int kind = code & 15;
      if (kind == FK_Far || kind == FK_IVortex) {
        const double own[4] = {tr[0], tr[1], tr[2], tr[3]};
        boundary_state(kind, a.ph, a.g.bxy, code >> 8, NpEdge, pp, a.ctl->time[a.step_index & 1], own,
                       a.g.Nx[fi], a.g.Ny[fi], tr);
      }
      wl = point_speed(a.ph.gamma, tr, a.g.oohK[kk]);
      if (!(wl == wl)) bad = 1;
    }
  }
  wl = wave_max(wl);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = wl;
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) {
    double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(&a.ctl->wmax[a.ring], (unsigned long long)__double_as_longlong(m));
    if (bad) atomicOr(&a.ctl->nan, 1);
  }
}

__global__ void k_ctl_init(Ctl *ctl, double time) {
  ctl->time[0] = time;
  ctl->time[1] = time;
  for (int i = 0; i < 5; ++i) ctl->dt[i] = 0.0;
  for (int i = 0; i < 3; ++i) ctl->wmax[i] = 0ull;
  ctl->nan = 0;
  ctl->steps = 0;
}

template <int N>
size_t lds_bytes_n(size_t csize, int Hmax) {
  using Sh = Shape<N>;
  return csize * (size_t(Sh::NV) * (kTile + 1) + 4 * (size_t(Sh::NF) * kTile + size_t(Sh::NpEdge) * Hmax));
}

size_t stage_lds_bytes(int N, int prec, int Hmax) {
  const size_t cs = prec == 0 ? sizeof(double) : sizeof(float);
  switch (N) {
    case 0: return lds_bytes_n<0>(cs, Hmax);
    case 1: return lds_bytes_n<1>(cs, Hmax);
    case 2: return lds_bytes_n<2>(cs, Hmax);
    case 3: return lds_bytes_n<3>(cs, Hmax);
    case 4: return lds_bytes_n<4>(cs, Hmax);
    default: return 0;
  }
}

template <int N, class S, class C, int FLUX, bool LAST>
hipError_t launch_stage_t(const StageArgs &a, hipStream_t s) {
  const size_t lds = lds_bytes_n<N>(sizeof(C), a.Hmax);
  static size_t attr_bytes = 0;  // dynamic LDS the kernel is currently allowed
  if (lds > 64 * 1024 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage<N, S, C, FLUX, LAST>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    attr_bytes = lds;
  }
  const int q = (a.ntiles + 7) / 8;
  hipLaunchKernelGGL((k_stage<N, S, C, FLUX, LAST>), dim3(8 * q), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

template <int N, class S, class C>
hipError_t launch_stage_f(const StageArgs &a, hipStream_t s) {
  switch (a.ph.flux) {
    case 0: return a.stage == 4 ? launch_stage_t<N, S, C, 0, true>(a, s) : launch_stage_t<N, S, C, 0, false>(a, s);
    case 1: return a.stage == 4 ? launch_stage_t<N, S, C, 1, true>(a, s) : launch_stage_t<N, S, C, 1, false>(a, s);
    case 2: return a.stage == 4 ? launch_stage_t<N, S, C, 2, true>(a, s) : launch_stage_t<N, S, C, 2, false>(a, s);
    case 3: return a.stage == 4 ? launch_stage_t<N, S, C, 3, true>(a, s) : launch_stage_t<N, S, C, 3, false>(a, s);
  }
  return hipErrorInvalidValue;
}

template <int N, class S>
hipError_t launch_prologue_t(const StageArgs &a, hipStream_t s) {
  const long items = long(a.K) * Shape<N>::NF;
  const int blocks = int((items + 255) / 256);
  hipLaunchKernelGGL((k_prologue<N, S>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_stage(int N, int prec, const StageArgs &a, hipStream_t s) {
  if (prec == 0) {
    switch (N) {
      case 0: return launch_stage_f<0, double, double>(a, s);
      case 1: return launch_stage_f<1, double, double>(a, s);
      case 2: return launch_stage_f<2, double, double>(a, s);
      case 3: return launch_stage_f<3, double, double>(a, s);
      case 4: return launch_stage_f<4, double, double>(a, s);
    }
  } else {
    switch (N) {
      case 0: return launch_stage_f<0, float, float>(a, s);
      case 1: return launch_stage_f<1, float, float>(a, s);
      case 2: return launch_stage_f<2, float, float>(a, s);
      case 3: return launch_stage_f<3, float, float>(a, s);
      case 4: return launch_stage_f<4, float, float>(a, s);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_prologue(int N, int prec, const StageArgs &a, hipStream_t s) {
  if (prec == 0) {
    switch (N) {
      case 0: return launch_prologue_t<0, double>(a, s);
      case 1: return launch_prologue_t<1, double>(a, s);
      case 2: return launch_prologue_t<2, double>(a, s);
      case 3: return launch_prologue_t<3, double>(a, s);
      case 4: return launch_prologue_t<4, double>(a, s);
    }
  } else {
    switch (N) {
      case 0: return launch_prologue_t<0, float>(a, s);
      case 1: return launch_prologue_t<1, float>(a, s);
      case 2: return launch_prologue_t<2, float>(a, s);
      case 3: return launch_prologue_t<3, float>(a, s);
      case 4: return launch_prologue_t<4, float>(a, s);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t upload_operators(int N, const double *divT, const double *interp) {
  if (N < 0 || N > 4) return hipErrorInvalidValue;
  const int NpInt = (N + 1) * (N + 2) / 2, NpFlux = (N + 2) * (N + 4), NF = 3 * (N + 2);
  std::vector<double> d(size_t(ops_size(N)));
  std::copy(divT, divT + size_t(NpFlux) * NpInt, d.begin());
  std::copy(interp, interp + size_t(NF) * NpInt, d.begin() + size_t(NpFlux) * NpInt);
  std::vector<float> f(d.begin(), d.end());
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cOpsD), d.data(), d.size() * sizeof(double),
                                   size_t(ops_offset(N)) * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(cOpsF), f.data(), f.size() * sizeof(float),
                           size_t(ops_offset(N)) * sizeof(float), hipMemcpyHostToDevice);
}

hipError_t launch_ctl_init(Ctl *ctl, double time, hipStream_t s) {
  hipLaunchKernelGGL(k_ctl_init, dim3(1), dim3(1), 0, s, ctl, time);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace gocfd
