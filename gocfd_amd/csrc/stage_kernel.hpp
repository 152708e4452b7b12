// gfx950 kernels for the 2D Euler DFR right-hand side and SSP54 stages,
// templated on the solution order N. Each stage_n<N>.hip translation unit
// includes this header once and instantiates one order, so that the orders
// compile in parallel and each holds only its own operators in constant
// memory.
//
// One stage launch per Runge-Kutta stage. A workgroup owns a tile of kTile
// consecutive elements (Hilbert order), whose element-major state is one
// contiguous chunk of HBM. Per stage and tile:
//   A. load the stage state (into LDS, or on the matrix-core path into the
//      registers of the lanes that use it); face traces by the
//      FluxEdgeInterp contraction (model_problems/Euler2D/edges.go:485-491)
//      for the tile's faces; the cross-tile neighbour traces from the
//      halo-trace buffer, or with shock capturing interpolated from the
//      neighbours' state for the faces listed in the halo table
//   B. Riemann / boundary flux at every face point (edges.go:325-452,
//      bcs.go:11-133), scaled by ||n|| (SetRTFluxOnEdges edges.go:454-483);
//      with shock capturing, minus the viscous edge flux
//      (StoreEdgeViscousFlux edges.go:151-244)
//   C. interior physical flux (minus eps*grad Q with shock capturing),
//      contravariant transform and the DivInt contraction (SetRTFluxInternal
//      euler.go:701-726, RHSInternalPoints euler.go:665-699, AddDissipation
//      dissipation.go:274-346), then the modal filter of the RHS with shock
//      capturing (LimitFilterSolution dissipation.go:520-542)
//   E. SSP54 combination (euler.go:502-565) with the global dt of this stage
//      or the element's local dt (CalculateLocalDT euler.go:973-1002)
//   F/G. traces of the new state, boundary states at the next stage's time
//      and the owner-side CFL wave-speed maximum (edges.go:246-323), folded
//      into one atomic per workgroup for the next stage's dt
//      (calculateGlobalDT euler.go:945-971)
// From N=3 on without shock capturing, the FluxEdgeInterp (A, F) and DivInt
// (C) contractions run on the matrix cores as 16x16x4 FP32 or FP64 MFMA
// tiles, with the element state held in the MFMA result-row layout.
// With shock capturing three passes precede each stage launch: the sensor
// with the vertex max-merge, the element viscosity fields (and the stage-2
// state filter), and the RT gradient with the stage's wave speeds.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "dfr_setup.hpp"
#include "kernels.hpp"
#include "physics.hpp"

namespace gocfd {
namespace dev {

template <int N>
struct Shape {
  static constexpr int NpInt = (N + 1) * (N + 2) / 2;
  static constexpr int NpEdge = N + 2;
  static constexpr int NpFlux = (N + 2) * (N + 4);
  static constexpr int NF = 3 * NpEdge;  // face points per element
  static constexpr int NV = 4 * NpInt;   // state values per element
  static constexpr int NR = NpInt + NF;  // RT points whose gradient is used
  // shock-operator block offsets (see ShockArgs)
  static constexpr int oP = 0, oM = NpInt * NpInt, oD = 2 * NpInt * NpInt, oV = 3 * NpInt * NpInt,
                       oVi = 4 * NpInt * NpInt, omf = 5 * NpInt * NpInt, oDiv = omf + NpInt,
                       oBary = oDiv + NR * NpFlux, shock_ops = oBary + 3 * NR;
  static constexpr int ops = (NpFlux + NF) * NpInt;
  // MFMA forms (16x16x4 tiles): the face DOFs of the DivInt contraction in
  // KE K-steps of 4; the FluxEdgeInterp contraction in RB row blocks of 16
  // face points over KI K-steps of 4 solution points (MfShape sets the
  // K-steps the matrix-core stage path uses)
  static constexpr int KE = (NF + 3) / 4;
  static constexpr int RB = (NF + 15) / 16, KI = (NpInt + 3) / 4;
};

// The DivInt and FluxEdgeInterp contractions run on the matrix cores from
// N=3 on (NpInt = 10, 15 rows of a 16-row tile) without shock capturing
// (whose RHS filter needs a whole element column in one thread).
template <int N, class C, bool SHOCK>
struct UseMfma {
  static constexpr bool value = N >= 3 && !SHOCK;
};
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
// the four variables of one trace point, one vector access
template <class C>
struct Vec4;
template <>
struct Vec4<float> { using type = floatx4; };
template <>
struct Vec4<double> { using type = doublex4; };

// The FP64 shock operators (V, Vi, filter weights, sensor matrices,
// barycentric weights) through the constant address space (they are
// read-only during every kernel): uniform scalar loads where they are used.
// Through a generic pointer, which an FP32 state cannot alias, the compiler
// hoisted whole matrices into vector registers.
using ConstD = const __attribute__((address_space(4))) double;
__device__ __forceinline__ ConstD *shock_ops_c(const ShockArgs &s) { return (ConstD *)s.ops; }

// 16x16x4 MFMA tiles in FP32 and FP64. Operands: lane l holds A[l%16][l/16]
// and B[l/16][l%16]; result element v of lane l is D[row(l/16, v)][l%16],
// row = 4 krow + v in FP32 and krow + 4 v in FP64 (both checked on gfx950).
template <class C>
struct Mma;
template <>
struct Mma<float> {
  using vec = floatx4;
  static __device__ __forceinline__ vec zero() { return vec{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ vec mma(float a, float b, vec c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int krow, int v) { return 4 * krow + v; }
};
template <>
struct Mma<double> {
  using vec = doublex4;
  static __device__ __forceinline__ vec zero() { return vec{0., 0., 0., 0.}; }
  static __device__ __forceinline__ vec mma(double a, double b, vec c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int krow, int v) { return krow + 4 * v; }
};

// Matrix-core path of the stage kernel: wave w owns the 16 elements 16 w +
// col of the tile and lane (col, krow) holds the four variables of element
// 16 w + col at the solution points row(krow, j), j < KI, in registers; these
// are the rows the lane receives of a 16x16 MFMA result, so the same values
// feed the B operand of K-step j of the FluxEdgeInterp and DivInt
// contractions (whose A operands take their columns in that point order)
// and the SSP combination with the divergence. FP32 results hold rows
// 4 krow + j, which needs all four K-steps also for NpInt = 10; FP64 results
// hold rows krow + 4 j, the natural order.
template <int N, class C>
struct MfShape {
  using Sh = Shape<N>;
  static constexpr int KI = sizeof(C) == 4 ? 4 : Sh::KI;  // K-steps over solution points
  static constexpr int KS = 2 * KI + Sh::KE;               // DivInt K-steps: r, s components, faces
  static __host__ __device__ constexpr int point(int krow, int j) {
    return sizeof(C) == 4 ? 4 * krow + j : krow + 4 * j;
  }
};

// The lane's KI values of one variable of its element (matrix-core path);
// row points at the variable's NpInt values in HBM. An FP32 lane holds four
// consecutive points 4 krow + j: one vector access of the run starting at
// min(4 krow, NpInt - 4) (points past NpInt repeat the last one on loads and
// are not stored). FP64 lanes hold points krow + 4 j, one access each.
template <class T>
struct Run4 {
  T v[4];
};
template <int N, class C, class T>
__device__ __forceinline__ void lane_load(const T *row, int krow, T (&out)[MfShape<N, C>::KI]) {
  constexpr int NpInt = Shape<N>::NpInt, KI = MfShape<N, C>::KI;
  if constexpr (sizeof(C) == 4) {
    static_assert(KI == 4 && NpInt >= 4, "FP32 lanes hold four consecutive points");
    const int o = min(4 * krow, NpInt - 4);
    const Run4<T> r = *reinterpret_cast<const Run4<T> *>(row + o);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int d = min(4 * krow + j, NpInt - 1) - o;
      out[j] = d == 0 ? r.v[0] : d == 1 ? r.v[1] : d == 2 ? r.v[2] : r.v[3];
    }
  } else {
#pragma unroll
    for (int j = 0; j < KI; ++j) out[j] = row[min(MfShape<N, C>::point(krow, j), NpInt - 1)];
  }
}
template <int N, class C, class T>
__device__ __forceinline__ void lane_store(T *row, int krow, const T (&val)[MfShape<N, C>::KI]) {
  constexpr int NpInt = Shape<N>::NpInt, KI = MfShape<N, C>::KI;
  if constexpr (sizeof(C) == 4) {
    if (4 * krow + 3 < NpInt) {
      *reinterpret_cast<Run4<T> *>(row + 4 * krow) = Run4<T>{{val[0], val[1], val[2], val[3]}};
      return;
    }
  }
#pragma unroll
  for (int j = 0; j < KI; ++j) {
    const int i = MfShape<N, C>::point(krow, j);
    if (i < NpInt) row[i] = val[j];
  }
}

// The operators of order N in constant memory (each translation unit holds
// the orders it instantiates), so that every coefficient is a uniform scalar
// load: DivInt^T [NpFlux][NpInt] followed by FluxEdgeInterp [3 NpEdge][NpInt],
// in FP64 and FP32.
constexpr int kOpsMax = ((4 + 2) * (4 + 4) + 3 * (4 + 2)) * 15;
template <int N>
static __constant__ double cOpsD[kOpsMax];
template <int N>
static __constant__ float cOpsF[kOpsMax];
// DivInt as the A operand of the MFMA contraction: K-step s, lane l holds
// DivInt[row l%16][DOF of K-row 4s + l/16] (zero outside the matrix)
constexpr int kMfMax = 13 * 64;
template <int N>
static __constant__ float cMfA[kMfMax];
template <int N>
static __constant__ double cMfAD[kMfMax];
// FluxEdgeInterp as the A operand: row block rb, K-step s, lane l holds
// Interp[16 rb + l%16][4 s + l/16]
constexpr int kMfIMax = 8 * 64;
template <int N>
static __constant__ float cMfI[kMfIMax];
template <int N>
static __constant__ double cMfID[kMfIMax];
// The order-N tables in arithmetic type C
template <int N, class C>
struct Tabs;
template <int N>
struct Tabs<N, float> {
  static __device__ __forceinline__ const float *ops() { return cOpsF<N>; }
  static __device__ __forceinline__ const float *div() { return cMfA<N>; }
  static __device__ __forceinline__ const float *interp() { return cMfI<N>; }
};
template <int N>
struct Tabs<N, double> {
  static __device__ __forceinline__ const double *ops() { return cOpsD<N>; }
  static __device__ __forceinline__ const double *div() { return cMfAD<N>; }
  static __device__ __forceinline__ const double *interp() { return cMfID<N>; }
};

// Face traces of the lane's element (matrix-core path) from the state in
// registers: for each variable, RB row blocks of 16 face points over KI
// K-steps; lane (col, krow) stores face points 16 rb + row(krow, j).
template <int N, class C>
__device__ __forceinline__ void traces_mfma_regs(const C (&qv)[4][MfShape<N, C>::KI], C *sT, int lane, int krow,
                                                 int me, int TS) {
  using Sh = Shape<N>;
  using M = Mma<C>;
  constexpr int NF = Sh::NF, RB = Sh::RB, KI = MfShape<N, C>::KI, E = kTile;
  const C *tab = Tabs<N, C>::interp();
  C ai[RB][KI];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int s = 0; s < KI; ++s) ai[rb][s] = tab[(rb * KI + s) * 64 + lane];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    typename M::vec acc[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) acc[rb] = M::zero();
#pragma unroll
    for (int s = 0; s < KI; ++s)
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) acc[rb] = M::mma(ai[rb][s], qv[v][s], acc[rb]);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int p = 16 * rb + M::row(krow, j);
        if (p < NF) sT[v * TS + p * kTP + me] = acc[rb][j];
      }
  }
}

// Tile index of a workgroup: with workgroups dealt round-robin to the 8 XCDs,
// XCD x processes the contiguous tile range [x*q, (x+1)*q), so neighbouring
// tiles share an L2 (xcd_tiles = 1); otherwise tile = workgroup id.
__device__ __forceinline__ int tile_of_block(int xcd_tiles) {
  if (!xcd_tiles) return blockIdx.x;
  const int q = gridDim.x >> 3;
  return (blockIdx.x & 7) * q + (blockIdx.x >> 3);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ unsigned long long bits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double from_bits(unsigned long long b) { return __longlong_as_double((long long)b); }

// Device-wide maximum of non-negative doubles kept as IEEE bits (which order
// like the values). The atomic is skipped when the stored value is already at
// least as large: maxima only grow during a launch, so a stale read can only
// be smaller, and then the atomic runs. This keeps thousands of workgroups
// from serialising on one address.
__device__ __forceinline__ void atomic_max_bits(unsigned long long *p, double v) {
  const unsigned long long b = bits(v);
  if (b > *reinterpret_cast<volatile unsigned long long *>(p)) atomicMax(p, b);
}

template <class C>
__device__ __forceinline__ void boundary_state(int kind, const Phys &ph, const double *bxy, int bidx,
                                               int NpEdge, int pp, double t, const C own[4],
                                               double nx, double ny, double qb[4]) {
  double qi[4] = {double(own[0]), double(own[1]), double(own[2]), double(own[3])};
  double qf[4];
  double gam = ph.gamma, minf = ph.minf, pinf = ph.pinf, cinf = ph.cinf;
  if (kind == FK_Far) {
    // bidx selects FSFar (0), FSIn (1) or FSOut (2) (edges.go:425-431)
    const double *qs = ph.qinf;
    if (bidx > 0) {
      const FarState &fs = ph.io[bidx - 1];
      qs = fs.q; gam = fs.gamma; minf = fs.minf; pinf = fs.pinf; cinf = fs.cinf;
    }
    qf[0] = qs[0]; qf[1] = qs[1]; qf[2] = qs[2]; qf[3] = qs[3];
  } else {
    const double *xy = bxy + 2 * (size_t(bidx) * NpEdge + pp);
    ivortex_state(t, xy[0], xy[1], qf);
  }
  riemann_bc(gam, minf, pinf, cinf, qi, qf, nx, ny, qb);
}

// Wave speed (|U| + c) / hK of a face point (StoreEdgeAggregates, edges.go:264-272)
__device__ __forceinline__ double point_speed(double g, const double q[4], double oohK) {
  const double oorho = 1. / q[0], u = q[1] * oorho, v = q[2] * oorho;
  const double U = sqrt(u * u + v * v);
  const double c = sound_speed(g, q);
  return oohK * (U + c);
}

// Dynamic LDS of the stage kernel: the tile state [NV][kTile+1] (not on the
// matrix-core path, whose state stays in registers) and the trace/flux
// buffer [4][3 NpEdge kTP + NpEdge Hmax] (rows of kTP elements; rounded to
// 32 bytes), the Riemann
// items' geometry [Fmax][4] and the elements' metric terms [6][kTile];
// without shock capturing, phase G's lists: the elements' 1/hK [kTile]
// (FP64), the owned faces with their codes [2][Wmax] and the export items
// with their destinations [2][Xmax] (ints); local time stepping adds the
// element dt and face-speed buffers.
template <int N>
__host__ __device__ constexpr size_t lds_trace_bytes(size_t csize, int Hmax, bool mf) {
  return (csize * ((mf ? 0 : size_t(Shape<N>::NV) * (kTile + 1)) +
                   4 * (size_t(Shape<N>::NF) * kTP + size_t(Shape<N>::NpEdge) * Hmax)) + 31) / 32 * 32;
}
template <int N>
__host__ __device__ constexpr size_t lds_tile_bytes(size_t csize, int Hmax, int Fmax, int WXmax, bool mf) {
  return lds_trace_bytes<N>(csize, Hmax, mf) + csize * (4 * size_t(Fmax) + 6 * kTile) +
         (WXmax > 0 ? (sizeof(double) * kTile + 2 * sizeof(int) * size_t(WXmax) + 15) / 16 * 16 : 0);
}
constexpr size_t lds_lts_bytes = sizeof(double) * kTile + sizeof(unsigned long long) * 3 * kTile;

// CheckIfFinished (euler.go:796-801): no further steps once Time >= FinalTime
__device__ __forceinline__ bool step_skipped(const StageArgs &a, const Ctl *ctl, double final_time) {
  return a.stage >= 0 && a.step_index > 0 && ctl->time[a.step_index & 1] >= final_time;
}

// The values of a stage launch that change from launch to launch (the rest
// of StageArgs describes the solver or the ensemble group)
struct StageLaunch {
  int stage, step_index, ring, par2, hpar;
  const void *qs;
  double t_rhs;
  double *rhs_out;
};
__device__ __forceinline__ StageLaunch launch_of(const StageArgs &a) {
  return StageLaunch{a.stage, a.step_index, a.ring, a.par2, a.hpar, a.qs, a.t_rhs, a.rhs_out};
}
__device__ __forceinline__ bool step_skipped(const StageLaunch &l, const Ctl *ctl, double final_time) {
  return l.stage >= 0 && l.step_index > 0 && ctl->time[l.step_index & 1] >= final_time;
}

// Waves per SIMD the register allocator targets. On the matrix-core path an
// FP64 tile needs about 45 KB of LDS at N=4 (three workgroups per CU) and
// fits 168 VGPRs; the FP32 stages fit 128 VGPRs without spilling.
#ifndef GOCFD_WAVES_F32
#define GOCFD_WAVES_F32 4
#endif
#ifndef GOCFD_WAVES_F32_LAST
#define GOCFD_WAVES_F32_LAST 3
#endif
#ifndef GOCFD_WAVES_F64
#define GOCFD_WAVES_F64 3
#endif
template <class S, bool LAST, bool SHOCK>
struct StageWaves {
  static constexpr int value =
      SHOCK ? 1 : sizeof(S) == 4 ? (LAST ? GOCFD_WAVES_F32_LAST : GOCFD_WAVES_F32) : GOCFD_WAVES_F64;
};

// S: storage of the stage states q1..q4; SIN: storage of this stage's
// input (q0 at stage 0 and for the RHS evaluation, else S). q0 and the
// stage-3 RHS R, which accumulate the step, are always FP64.
template <int N, class S, class SIN, class C, int FLUX, bool LAST, bool SHOCK>
__device__ __forceinline__ void stage_tile(const StageArgs &a, const StageLaunch &l, const int tile) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NF = Sh::NF, NV = Sh::NV, NR = Sh::NR;
  constexpr int E = kTile, P = kTile + 1;
  static_assert(kThreads == 4 * kTile, "one thread per (element, variable)");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // Without shock capturing the traces of cross-tile neighbours come from
  // the halo-trace buffer, which the neighbours' tiles filled after the
  // previous stage; with it (the stage state may be filtered after the
  // previous stage) each tile interpolates them from the neighbours' state
  // into a halo part of the trace buffer.
  constexpr bool HX = !SHOCK;
  // Matrix-core path: the state stays in registers (no sQ, no sR)
  constexpr bool MF = UseMfma<N, C, SHOCK>::value;
  using V4 = typename Vec4<C>::type;
  C *sQ = reinterpret_cast<C *>(smem_raw);  // [NV][P] (VALU path)
  const int TS = NF * kTP + (HX ? 0 : NpEdge * a.Hmax);
  C *sT = MF ? sQ : sQ + NV * P;             // [4][TS]: traces, then fluxes
  C *sR = sT;                                // [NV][P]: RHS (aliases sT; VALU path)
  // the Riemann items' geometry and the metric terms (Jdet, Jinv[4], -1/Jdet)
  V4 *sFG = reinterpret_cast<V4 *>(smem_raw + lds_trace_bytes<N>(sizeof(C), HX ? 0 : a.Hmax, MF));  // [Fmax]
  C *sJ = reinterpret_cast<C *>(sFG + a.Fmax);                                                 // [6][E]
  // phase G's lists (without shock capturing)
  double *sOK = reinterpret_cast<double *>(sJ + 6 * E);  // [E] 1/hK
  int *sWP = reinterpret_cast<int *>(sOK + E);           // [Wmax] owned faces e | f << 8
  int *sWC = sWP + a.Wmax;                               // [Wmax] their codes
  int *sXP = sWC + a.Wmax, *sXD = sXP + a.Xmax;          // [Xmax] export items, destinations
  // local time stepping only: element dt and per-face speed maxima, after
  // the tile buffers (the launch adds their bytes only when LTS is on)
  double *sDT = reinterpret_cast<double *>(
      smem_raw + lds_tile_bytes<N>(sizeof(C), HX ? 0 : a.Hmax, a.Fmax, HX ? a.Wmax + a.Xmax : 0, MF));
  unsigned long long *sWf = reinterpret_cast<unsigned long long *>(sDT + kTile);
  __shared__ double red[kThreads / 64];

  const int tid = threadIdx.x;
  const int k0 = tile * E;
  const int Kp = a.Kp;
  const SIN *qs = static_cast<const SIN *>(l.qs);

  // The tile's work lists and state do not depend on the control block:
  // their loads go out first, so that their latency overlaps the chain
  // tile -> case -> control block and each other. The lists are read to
  // their padded lengths (zero past each tile's count), so that no load
  // waits for a count; the counts bound the loops that use the items.
  constexpr int ITB = (E * NF + kThreads - 1) / kThreads;
  const int nfl = a.g.flux_cnt[tile] * NpEdge;
  const int *flist = a.g.flux_list + size_t(tile) * a.Fmax;
  int fpk[ITB];
#pragma unroll
  for (int j = 0; j < ITB; ++j) {
    const int it = tid + j * kThreads;
    fpk[j] = it < a.Fmax * NpEdge ? flist[it / NpEdge] : 0;
  }
  // the items' geometry and the elements' metric terms, to LDS below
  V4 fgv = V4{C(0), C(0), C(0), C(0)};
  if (tid < a.Fmax) fgv = static_cast<const V4 *>(a.g.flux_geo)[size_t(tile) * a.Fmax + tid];
  // (wave 0: Jdet and Jinv[3]; wave c = 1..3: Jinv[c-1] of element te)
  double jm0, jm1 = 0.;
  {
    const int c = tid / E, e = tid - c * E;
    jm0 = c == 0 ? a.g.Jdet[k0 + e] : a.g.Jinv[size_t(c - 1) * Kp + k0 + e];
    if (c == 0) jm1 = a.g.Jinv[3 * size_t(Kp) + k0 + e];
  }
  // phase G's lists (owned faces, their codes, 1/hK; export items)
  const int nwc = HX ? a.g.wave_cnt[tile] : 0, nxc = HX ? a.g.xp_cnt[tile] : 0;
  int gw = 0, gc = 0, gx = 0, gd = 0;
  double gok = 0.;
  if constexpr (HX) {
    if (tid < a.Wmax) {
      gw = a.g.wave_list[size_t(tile) * a.Wmax + tid];
      gc = a.g.wave_code[size_t(tile) * a.Wmax + tid];
    }
    if (tid < a.Xmax) {
      gx = a.g.xp_list[size_t(tile) * a.Xmax + tid];
      gd = a.g.xp_dst[size_t(tile) * a.Xmax + tid];
    }
    if (tid < E) gok = a.g.oohK[k0 + tid];
  }
  const int nbnd = a.g.bnd_cnt[tile];  // non-shared faces: the boundary loop runs only if any
  int hk0 = -1, hf0 = 0;
  if (!HX && tid < a.Hmax * 4) {
    const int h = tid % a.Hmax;
    hk0 = a.g.halo_elem[size_t(tile) * a.Hmax + h];
    hf0 = a.g.halo_face[size_t(tile) * a.Hmax + h];
  }

  // ---- A: tile state -> LDS (coalesced, contiguous chunk), or on the
  // matrix-core path to the registers of the lanes that use it ----
  // Every load is issued before the first LDS store, so a thread's LPT
  // loads are in flight together.
  constexpr int LPT = (E * NV + kThreads - 1) / kThreads;
  using MS = MfShape<N, C>;
  constexpr int KI = MF ? MS::KI : 1;
  const int lane = tid & 63, krow = lane >> 4;
  const int me = (tid >> 6) * 16 + (lane & 15);  // element of the lane (matrix-core path)
  SIN v[MF ? 1 : LPT];
  C qv[4][KI];  // matrix-core path: the variables at points MS::point(krow, j)
  if constexpr (MF) {
    // a point past NpInt repeats the last one (a finite operand for the zero
    // columns of the A fragments)
    const SIN *qe = qs + size_t(k0 + me) * NV;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      SIN r[KI];
      lane_load<N, C>(qe + c * NpInt, krow, r);
#pragma unroll
      for (int j = 0; j < KI; ++j) qv[c][j] = C(r[j]);
    }
  } else {
    const SIN *qt = qs + size_t(k0) * NV;
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      v[j] = (idx < E * NV) ? qt[idx] : S(0);
    }
  }
  // the neighbour traces of the cross-tile faces, under the state loads:
  // the cross-tile faces lead the flux list and their traces sit at the
  // tile's item slots, so these loads need nothing else loaded first (items
  // below Xmax that are not cross-tile read slots they never use)
  V4 hx[HX ? ITB : 1];
  if constexpr (HX) {
    const V4 *hin = static_cast<const V4 *>(a.ht) + size_t(l.hpar) * a.ntiles * a.Fmax * NpEdge +
                    size_t(tile) * a.Fmax * NpEdge;
#pragma unroll
    for (int j = 0; j < ITB; ++j) {
      const int it = tid + j * kThreads;
      hx[j] = V4{C(0), C(0), C(0), C(0)};
      if (it < a.Xmax * NpEdge) hx[j] = hin[it];
    }
  }

  const bool rhs_only = l.stage < 0;
  const int par = l.step_index & 1;
  const int cs = a.tile_case[tile];
  const CaseInfo &ci = a.cases[cs];
  const Phys &ph = ci.ph;
  const bool first = tile == ci.tile0;  // does the case's control bookkeeping
  Ctl *ctl = a.ctl + cs;
  const double time = rhs_only ? l.t_rhs : ctl->time[par];
  // this stage's wave-speed maxima, written by the previous launch: read now
  // so that phase E does not wait for them
  const unsigned long long wmax_bits = rhs_only ? 0ull : ctl->wmax[l.ring];
  const unsigned long long wmaxv_bits = (rhs_only || !SHOCK) ? 0ull : ctl->wmaxv[l.ring];
  const C *DvT = Tabs<N, C>::ops();                     // DivInt^T [NpFlux][NpInt]
  const C *Itp = DvT + Sh::NpFlux * NpInt;              // FluxEdgeInterp [NF][NpInt]
  const C g = C(ph.gamma);
  if constexpr (!MF) {
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      if (idx < E * NV) {
        const int e = idx / NV, r = idx - e * NV;
        sQ[r * P + e] = C(v[j]);
      }
    }
  }
  if (tid < a.Fmax) sFG[tid] = fgv;
  if constexpr (HX) {
    if (tid < a.Wmax) {
      sWP[tid] = gw;
      sWC[tid] = gc;
    }
    if (tid < a.Xmax) {
      sXP[tid] = gx;
      sXD[tid] = gd;
    }
    if (tid < E) sOK[tid] = gok;
  }
  {
    const int c = tid / E, e = tid - c * E;
    sJ[c * E + e] = C(jm0);
    if (c == 0) {
      sJ[4 * E + e] = C(jm1);
      sJ[5 * E + e] = -C(1. / jm0);
    }
  }
  // matrix-core path: the traces need only the lane's registers, so they
  // are formed before the one barrier of phase A
  if constexpr (MF) traces_mfma_regs<N, C>(qv, sT, lane, krow, me, TS);
  __syncthreads();
  // the control-block chain (tile -> case -> time) has had the tile loads and
  // the LDS stores to arrive under
  if (step_skipped(l, ctl, ph.final_time)) {
    if (l.stage == 4 && first && threadIdx.x == 0) ctl->time[par ^ 1] = time;
    return;
  }
  const int te = tid % E, tn = tid / E;  // (element, variable) of this thread (VALU path)
  if constexpr (!MF) {
    C q[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) q[i] = sQ[(tn * NpInt + i) * P + te];
#pragma unroll 2
    for (int p = 0; p < NF; ++p) {
      C t = C(0);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * q[i];
      sT[tn * TS + p * kTP + te] = t;
    }
  }
  for (int it = tid; it < (HX ? 0 : a.Hmax * 4); it += kThreads) {
    const int h = it % a.Hmax, n = it / a.Hmax;
    const int kh = it == tid ? hk0 : a.g.halo_elem[size_t(tile) * a.Hmax + h];
    if (kh < 0) continue;
    const int f = it == tid ? hf0 : a.g.halo_face[size_t(tile) * a.Hmax + h];
    const SIN *qn = qs + size_t(kh) * NV + n * NpInt;
    C q[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) q[i] = C(qn[i]);
#pragma unroll
    for (int pp = 0; pp < NpEdge; ++pp) {
      const C *row = Itp + (f * NpEdge + pp) * NpInt;
      C t = C(0);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) t += row[i] * q[i];
      sT[n * TS + NF * kTP + h * NpEdge + pp] = t;
    }
  }
  if constexpr (!MF) __syncthreads();

  // ---- B: numerical flux at every face point ----
  // Shared faces from the tile's work list: a face with an out-of-tile
  // neighbour evaluates its own side; of an in-tile pair only the owner
  // evaluates F* and also fills the partner's slot (negated, at the reversed
  // point, scaled by the partner's ||n||), which is the value the reference
  // stores once per edge and reads back from both sides (edges.go:454-483).
  // With shock capturing (whose viscous penalty also reads the owner's trace
  // at the reversed point) results stay in registers until every trace of
  // the tile has been read; in FP64 until the thread's last item. Boundary
  // faces follow in a rolled loop: a boundary face point reads and writes
  // only its own LDS slot.
  C fl[SHOCK || sizeof(C) == 8 ? ITB : 1][4];
#pragma unroll
  for (int j = 0; j < ITB; ++j) {
    const int it = tid + j * kThreads;
    if (it < nfl) {
      const int w = it / NpEdge, pp = it - w * NpEdge;
      const int packed = fpk[j];
      const int e = packed & 255, f = (packed >> 8) & 3, p = f * NpEdge + pp;
      const size_t fi = size_t(f) * Kp + k0 + e;
      const V4 gq = sFG[w];
      C own[4], oth[4], F[4];
      const int base = packed >> 13;                    // the neighbour's slot
      const int stride = (packed & (1 << 11)) ? 1 : kTP;  // halo: point-contiguous
      const int rp = NpEdge - 1 - pp;
      const bool xt = HX && (packed & (1 << 11));  // neighbour traces in hx[j]
      const int idx = xt ? p * kTP + e : base + rp * stride;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        own[c] = sT[c * TS + p * kTP + e];
        oth[c] = sT[c * TS + idx];
        if (xt) oth[c] = hx[j][c];
      }
      const C nx = gq[0], ny = gq[1];
      const bool owner = packed & (1 << 12);
      if (owner) numflux<C, !SHOCK>(FLUX, g, own, oth, nx, ny, F);
      else numflux<C, !SHOCK>(FLUX, g, oth, own, nx, ny, F);
      if constexpr (SHOCK) {
        // Viscous edge flux at owner point jo (StoreEdgeViscousFlux,
        // edges.go:151-244): average of the two sides' eps*grad Q along
        // the owner normal, minus the penalty on the owner's trace jump.
        const int kk = k0 + e, nb = a.sh.nbr_elem[fi], nf = a.sh.nbr_face[fi];
        const int kL = owner ? kk : nb, kR = owner ? nb : kk;
        const int rowL = NpInt + (owner ? f * NpEdge + pp : nf * NpEdge + rp);
        const int rowR = NpInt + (owner ? nf * NpEdge + rp : f * NpEdge + pp);
        // with both sides inactive eps is zero on the edge and VF vanishes
        const bool aL = a.sh.active[kL] != 0, aR = a.sh.active[kR] != 0;
        if (aL || aR) {
          const double lam = 0.5 * ((aL ? a.sh.epsF[size_t(rowL) * Kp + kL] : 0.) +
                                    (aR ? a.sh.epsF[size_t(rowR) * Kp + kR] : 0.));
          const C coef = C((a.sh.omega * lam) * a.sh.vooL[fi]);
          const C *diss = static_cast<const C *>(a.sh.diss);
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            // owner traces at owner points jo and NpEdge-1-jo
            const C EQj = owner ? own[c] : oth[c];
            const C EQr = owner ? sT[c * TS + (f * NpEdge + rp) * kTP + e] : sT[c * TS + base + pp * stride];
            const C *dX = diss + size_t(c * 2 * NR) * Kp, *dY = dX + size_t(NR) * Kp;
            const C vFL = aL ? nx * dX[size_t(rowL) * Kp + kL] + ny * dY[size_t(rowL) * Kp + kL] : C(0);
            const C vFR = aR ? nx * dX[size_t(rowR) * Kp + kR] + ny * dY[size_t(rowR) * Kp + kR] : C(0);
            const C VF = C(0.5) * (vFL + vFR) - coef * (EQj - EQr);
            F[c] = F[c] - VF;
          }
        }
      }
      if constexpr (!SHOCK && sizeof(C) == 4) {
        // Without shock capturing each trace slot of the tile is read by
        // exactly one item (an in-tile pair by its owner item at point pp,
        // which reads and writes the partner's point NpEdge-1-pp; a
        // cross-tile face by its own item), so the item can overwrite the
        // slots it read with the scaled fluxes at once. Measured faster in
        // FP32 only (FP64 writes its results after its last item: 75
        // against 81 us at order 4, 80,000 elements).
        const C wn = gq[2];  // IInII, negated on the non-owner side
#pragma unroll
        for (int c = 0; c < 4; ++c) sT[c * TS + p * kTP + e] = F[c] * wn;
        if (packed & (1 << 10)) {
          const C w2 = gq[3];
#pragma unroll
          for (int c = 0; c < 4; ++c) sT[c * TS + idx] = F[c] * w2;
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) fl[j][c] = F[c];
      }
    }
  }
  // results held until every item of the thread is done (FP64), or until
  // every trace of the tile has been read (shock capturing: a barrier)
  constexpr bool HOLD = SHOCK || sizeof(C) == 8;
  if constexpr (SHOCK) __syncthreads();
#pragma unroll
  for (int j = 0; j < (HOLD ? ITB : 0); ++j) {
    const int it = tid + j * kThreads;
    if (it < nfl) {
      const int w = it / NpEdge, pp = it - w * NpEdge;
      const int packed = fpk[j];
      const int e = packed & 255, f = (packed >> 8) & 3, p = f * NpEdge + pp;
      const V4 gq = sFG[w];
      const C wn = gq[2];  // IInII, negated on the non-owner side
#pragma unroll
      for (int c = 0; c < 4; ++c) sT[c * TS + p * kTP + e] = fl[j][c] * wn;
      if (packed & (1 << 10)) {
        // the in-tile partner's slot (pf*NpEdge*E + e2) at the reversed
        // point, with the partner's IInII negated
        const int idx = (packed >> 13) + (NpEdge - 1 - pp) * kTP;
        const C w2 = gq[3];
#pragma unroll
        for (int c = 0; c < 4; ++c) sT[c * TS + idx] = fl[j][c] * w2;
      }
    }
  }
#pragma unroll 1
  for (int it = tid; it < (nbnd > 0 ? E * NF : 0); it += kThreads) {
    const int e = it % E, p = it / E, f = p / NpEdge, pp = p - f * NpEdge;
    const size_t fi = size_t(f) * Kp + k0 + e;
    const int code = a.g.code[fi];
    const int kind = code & 15;
    if (kind == FK_Shared) continue;
    C own[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) own[c] = sT[c * TS + p * kTP + e];
    const double nxd = a.g.Nx[fi], nyd = a.g.Ny[fi];
    const C nx = C(nxd), ny = C(nyd);
    C F[4] = {C(0), C(0), C(0), C(0)};
    if (kind == FK_Far || kind == FK_IVortex) {
      double qb[4];
      boundary_state(kind, ph, a.g.bxy, code >> 8, NpEdge, pp, time, own, nxd, nyd, qb);
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
    if constexpr (SHOCK) {
      // unshared edge: the own side's eps*grad Q along its normal
      const int row = NpInt + f * NpEdge + pp;
      if (a.sh.active[k0 + e] != 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const C *dX = static_cast<const C *>(a.sh.diss) + size_t(c * 2 * NR + row) * Kp + k0 + e;
          F[c] = F[c] - (nx * dX[0] + ny * dX[size_t(NR) * Kp]);
        }
      }
    }
    const C w = C(a.g.IInII[fi]);
#pragma unroll
    for (int c = 0; c < 4; ++c) sT[c * TS + p * kTP + e] = F[c] * w;
  }
  __syncthreads();

  // FP32 matrix-core path, stages 1-3: phase E's q0 operand, loaded under
  // phase C
  constexpr bool PF = MF && !LAST && sizeof(C) == 4;
  double q0p[PF ? 4 : 1][PF ? KI : 1];
  if constexpr (PF) {
    if (l.stage >= 1) {
      const double *q0e = static_cast<const double *>(a.q0) + size_t(k0 + me) * NV;
#pragma unroll
      for (int c = 0; c < 4; ++c) lane_load<N, C>(q0e + c * NpInt, krow, q0p[c]);
    }
  }
  // ---- C: interior flux + divergence ----
  // (matrix-core path: the divergence stays in the lanes' registers, rows
  // MS::point(krow, j) of element me, the points of qv)
  typename Mma<C>::vec rq[MF ? 4 : 1];
  if constexpr (MF) {
    // K-steps 2j and 2j + 1: the r and s components of the contravariant
    // interior flux at the lane's point j (all four variables from one flux
    // evaluation); then the scaled face fluxes from LDS
    using M = Mma<C>;
    C af[MS::KS];
#pragma unroll
    for (int s2 = 0; s2 < MS::KS; ++s2) af[s2] = Tabs<N, C>::div()[s2 * 64 + lane];
    const C Jdc = sJ[me], J0 = sJ[E + me], J1 = sJ[2 * E + me], J2 = sJ[3 * E + me], J3 = sJ[4 * E + me];
#pragma unroll
    for (int c = 0; c < 4; ++c) rq[c] = M::zero();
#pragma unroll
    for (int j = 0; j < KI; ++j) {
      const C q4[4] = {qv[0][j], qv[1][j], qv[2][j], qv[3][j]};
      C Fx[4], Fy[4];
      flux_xy(g, q4, Fx, Fy);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const C Fr = Jdc * (J0 * Fx[c] + J1 * Fy[c]), Fs = Jdc * (J2 * Fx[c] + J3 * Fy[c]);
        rq[c] = M::mma(af[2 * j], Fr, rq[c]);
        rq[c] = M::mma(af[2 * j + 1], Fs, rq[c]);
      }
    }
#pragma unroll
    for (int s2 = 0; s2 < Sh::KE; ++s2) {
      const int p = min(s2 + Sh::KE * krow, NF - 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) rq[c] = M::mma(af[2 * KI + s2], sT[c * TS + p * kTP + me], rq[c]);
    }
    const C moojd = sJ[5 * E + me];  // -1/Jdet
#pragma unroll
    for (int c = 0; c < 4; ++c) rq[c] = rq[c] * moojd;
  } else {
    // VALU: thread (element te, variable tn)
    const int kk = k0 + te;
    bool act = false;  // shock capturing: eps*grad Q is nonzero somewhere in the element
    if constexpr (SHOCK) act = a.sh.active[kk] != 0;
    const C Jdc = sJ[te], J0 = sJ[E + te], J1 = sJ[2 * E + te], J2 = sJ[3 * E + te], J3 = sJ[4 * E + te];
    C acc[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) acc[i] = C(0);
#pragma unroll 1
    for (int j = 0; j < NpInt; ++j) {
      const C r0 = sQ[(0 * NpInt + j) * P + te], r1 = sQ[(1 * NpInt + j) * P + te];
      const C r2 = sQ[(2 * NpInt + j) * P + te], r3 = sQ[(3 * NpInt + j) * P + te];
      C Fx, Fy;
      flux_component(tn, g, r0, r1, r2, r3, Fx, Fy);
      if constexpr (SHOCK) {
        if (act) {
          const C *dX = static_cast<const C *>(a.sh.diss) + size_t(tn * 2 * NR + j) * Kp + kk;
          Fx -= dX[0];
          Fy -= dX[size_t(NR) * Kp];
        }
      }
      const C Fr = Jdc * (J0 * Fx + J1 * Fy), Fs = Jdc * (J2 * Fx + J3 * Fy);
#pragma unroll
      for (int i = 0; i < NpInt; ++i) acc[i] += DvT[j * NpInt + i] * Fr + DvT[(NpInt + j) * NpInt + i] * Fs;
    }
#pragma unroll 2
    for (int p = 0; p < NF; ++p) {
      const C Fe = sT[tn * TS + p * kTP + te];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) acc[i] += DvT[(2 * NpInt + p) * NpInt + i] * Fe;
    }
    const C moojd = sJ[5 * E + te];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) acc[i] *= moojd;
    if constexpr (SHOCK) {
      // modal filter of the RHS (limitAndFilterSolution dissipation.go:606-622)
      ConstD *V = shock_ops_c(a.sh) + Sh::oV, *Vi = shock_ops_c(a.sh) + Sh::oVi, *mf = shock_ops_c(a.sh) + Sh::omf;
      const double fac = 1. - sin(0.5 * M_PI * a.sh.sigma[kk]);
      double m[NpInt];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < NpInt; ++j) t += Vi[i * NpInt + j] * double(acc[j]);
        m[i] = i == 0 ? t : t * (mf[i] * fac);
      }
#pragma unroll
      for (int i = 0; i < NpInt; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < NpInt; ++j) t += V[i * NpInt + j] * m[j];
        acc[i] = C(t);
      }
    }
    __syncthreads();  // every thread has read its fluxes; sR may now overwrite sT
#pragma unroll
    for (int i = 0; i < NpInt; ++i) sR[(tn * NpInt + i) * P + te] = acc[i];
  }
  if constexpr (!MF) __syncthreads();

  if (rhs_only) {
    double *out = l.rhs_out + size_t(k0) * NV;
    if constexpr (MF) {
#pragma unroll
      for (int j = 0; j < KI; ++j) {
        const int i = MS::point(krow, j);
        if (i < NpInt)
#pragma unroll
          for (int c = 0; c < 4; ++c) out[size_t(me) * NV + c * NpInt + i] = double(rq[c][j]);
      }
    } else {
      for (int idx = tid; idx < E * NV; idx += kThreads) {
        const int e = idx / NV, r = idx - e * NV;
        out[idx] = double(sR[r * P + e]);
      }
    }
    return;
  }

  constexpr int ITG = (3 * E * NpEdge + kThreads - 1) / kThreads;

  // ---- E: SSP54 stage combination (euler.go:511-565), FP64 arithmetic ----
  const int stage = l.stage;
  const bool lts = a.lts.on;
  double dt = 0.0;  // GlobalDT; stays 0 with local time stepping (never computed)
  if (!lts) {
    dt = ph.cfl / from_bits(wmax_bits);
    if constexpr (SHOCK) {
      const double dtv = a.sh.cdiff / from_bits(wmaxv_bits);
      dt = fmin(dt, dtv);
    }
    if (time + dt > ph.final_time) dt = ph.final_time - time;
  } else {
    // CalcElementMaxWaveSpeed (edges.go:291-323) + CalculateLocalDT: the
    // element DT is CFL over the max of its previous value (-100 at stage 0)
    // and the speeds of its three edges
    if (tid < E) {
      const int kk = k0 + tid;
      double d = 0.0;
      if (kk < ci.k_end) {
        const double *fw = a.lts.fw + size_t(l.par2) * 3 * Kp;
        double w = stage == 0 ? -100.0 : a.lts.DTk[kk];
#pragma unroll
        for (int f = 0; f < 3; ++f) w = fmax(w, fw[a.lts.wsrc[size_t(f) * Kp + kk]]);
        d = ph.cfl / w;
        if constexpr (SHOCK) {
          const double *fwv = a.lts.fwv + size_t(l.par2) * 3 * Kp;
          double dv = a.lts.DTv[kk];
#pragma unroll
          for (int f = 0; f < 3; ++f) dv = fmax(dv, fwv[a.lts.wsrc[size_t(f) * Kp + kk]]);
          if (dv > 1.e-9) {
            dv = a.sh.cdiff / dv;
            d = fmin(d, dv);
          }
          a.lts.DTv[kk] = dv;
        }
        a.lts.DTk[kk] = d;
      }
      sDT[tid] = d;
    }
    __syncthreads();
  }
  int bad = 0;
  {
    // The operands of this stage are loaded for all LPT values of the
    // thread first, then combined and stored.
    const size_t base = size_t(k0) * NV;
    double *q0 = static_cast<double *>(a.q0) + base;
    S *q1 = static_cast<S *>(a.q1) + base;
    S *q2 = static_cast<S *>(a.q2) + base;
    S *q3 = static_cast<S *>(a.q3) + base;
    S *q4 = static_cast<S *>(a.q4) + base;
    double *Rr = static_cast<double *>(a.R) + base;
    const int nreal = min(E, ci.k_end - k0);
    if constexpr (MF) {
      // the lane's values: variable c at point MS::point(krow, j) of element
      // me; the divergence is rq, the stage input qv, which takes the new
      // state for phase F (a repeated point keeps its finite old value)
      const size_t eb = size_t(me) * NV;
      const double dte = lts ? sDT[me] : dt;
      const bool real = me < nreal;
      if constexpr (LAST) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          double v0[KI], vR[KI];
          S v2[KI], v3[KI];
          lane_load<N, C>(q0 + eb + c * NpInt, krow, v0);
          lane_load<N, C>(q2 + eb + c * NpInt, krow, v2);
          lane_load<N, C>(q3 + eb + c * NpInt, krow, v3);
          lane_load<N, C>(Rr + eb + c * NpInt, krow, vR);
#pragma unroll
          for (int j = 0; j < KI; ++j) {
            const double dtR = dte * double(rq[c][j]);
            const double Rn = -v0[j] + 0.517231671970585 * double(v2[j]) + 0.096059710526146 * double(v3[j]) +
                              0.386708617503269 * double(qv[c][j]) + 0.063692468666290 * (dte * vR[j]) +
                              0.226007483236906 * dtR;
            vR[j] = Rn;
            v0[j] += Rn;
            if (MS::point(krow, j) < NpInt) {
              qv[c][j] = C(v0[j]);
              if (real && !(v0[j] == v0[j])) bad = 1;
            }
          }
          lane_store<N, C>(Rr + eb + c * NpInt, krow, vR);
          lane_store<N, C>(q0 + eb + c * NpInt, krow, v0);
        }
      } else {
        double v0[4][KI];
        if constexpr (PF) {
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int j = 0; j < KI; ++j) v0[c][j] = q0p[c][j];  // loaded under phase C
        } else if (stage >= 1) {
#pragma unroll
          for (int c = 0; c < 4; ++c) lane_load<N, C>(q0 + eb + c * NpInt, krow, v0[c]);
        }
        S *qn = stage == 0 ? q1 : stage == 1 ? q2 : stage == 2 ? q3 : q4;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          S nv[KI];
          double rh[KI];
#pragma unroll
          for (int j = 0; j < KI; ++j) {
            const double rhs = double(rq[c][j]);
            const double qsv = double(qv[c][j]);
            const double dtR = dte * rhs;
            rh[j] = rhs;
            switch (stage) {
              case 0: nv[j] = S(qsv + 0.391752226571890 * dtR); break;
              case 1: nv[j] = S(0.444370493651235 * v0[c][j] + 0.555629506348765 * qsv + 0.368410593050371 * dtR); break;
              case 2: nv[j] = S(0.620101851488403 * v0[c][j] + 0.379898148511597 * qsv + 0.251891774271694 * dtR); break;
              default: nv[j] = S(0.178079954393132 * v0[c][j] + 0.821920045606868 * qsv + 0.544974750228521 * dtR); break;
            }
            if (MS::point(krow, j) < NpInt) {
              qv[c][j] = C(nv[j]);
              if (real && !(nv[j] == nv[j])) bad = 1;
            }
          }
          lane_store<N, C>(qn + eb + c * NpInt, krow, nv);
          if (stage == 3) lane_store<N, C>(Rr + eb + c * NpInt, krow, rh);
        }
      }
    } else if constexpr (LAST) {
      // stage 4: R = -q0 + a q2 + b q3 + c q4 + d dt R3 + e dt RHS, q0 += R;
      // with FP32 stage states the four operand arrays are loaded in two
      // halves (the register budget of three waves per SIMD)
      constexpr int CH = sizeof(S) == 4 ? (LPT + 1) / 2 : LPT;
#pragma unroll
      for (int j0 = 0; j0 < LPT; j0 += CH) {
      double v0[CH], vR[CH];
      S v2[CH], v3[CH];
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        const int idx = tid + (j0 + jj) * kThreads;
        const bool ok = j0 + jj < LPT && idx < E * NV;
        v0[jj] = ok ? q0[idx] : 0.0;
        v2[jj] = ok ? q2[idx] : S(0);
        v3[jj] = ok ? q3[idx] : S(0);
        vR[jj] = ok ? Rr[idx] : 0.0;
      }
#pragma unroll
      for (int jj = 0; jj < CH; ++jj) {
        const int j = j0 + jj;
        const int idx = tid + j * kThreads;
        if (j >= LPT) continue;
        if (idx >= E * NV) continue;
        const int e = idx / NV, r = idx - e * NV;
        const double dte = lts ? sDT[e] : dt;
        const double dtR = dte * double(sR[r * P + e]);
        const double qsv = double(sQ[r * P + e]);
        const double q0v = double(v0[jj]);
        const double dtR3 = dte * double(vR[jj]);
        const double Rn = -q0v + 0.517231671970585 * double(v2[jj]) + 0.096059710526146 * double(v3[jj]) +
                          0.386708617503269 * qsv + 0.063692468666290 * dtR3 + 0.226007483236906 * dtR;
        Rr[idx] = Rn;
        const double nv = q0v + Rn;
        q0[idx] = nv;
        sQ[r * P + e] = C(nv);
        if (e < nreal && !(nv == nv)) bad = 1;
      }
      }
    } else {
      double v0[LPT];
      if (stage >= 1) {
#pragma unroll
        for (int j = 0; j < LPT; ++j) {
          const int idx = tid + j * kThreads;
          v0[j] = (idx < E * NV) ? q0[idx] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < LPT; ++j) {
        const int idx = tid + j * kThreads;
        if (idx >= E * NV) continue;
        const int e = idx / NV, r = idx - e * NV;
        const double rhs = double(sR[r * P + e]);
        const double qsv = double(sQ[r * P + e]);
        const double dtR = (lts ? sDT[e] : dt) * rhs;
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
            Rr[idx] = rhs;
            nv = S(0.178079954393132 * double(v0[j]) + 0.821920045606868 * qsv + 0.544974750228521 * dtR);
            q4[idx] = nv;
            break;
        }
        sQ[r * P + e] = C(nv);
        if (e < nreal && !(nv == nv)) bad = 1;
      }
    }
  }
  const double tnext = LAST ? time + dt : time;
  double wl = 0.0;
  if constexpr (!SHOCK) {
    // With shock capturing the next stage state may be filtered before its
    // traces are taken, so its wave speeds come from the gradient pass.
    if (lts && tid < 3 * E) sWf[tid] = 0ull;
    __syncthreads();

    // ---- F: traces of the new state (own faces) ----
    if constexpr (MF) {
      traces_mfma_regs<N, C>(qv, sT, lane, krow, me, TS);
    } else {
      C q[NpInt];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) q[i] = sQ[(tn * NpInt + i) * P + te];
#pragma unroll 2
      for (int p = 0; p < NF; ++p) {
        C t = C(0);
#pragma unroll
        for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * q[i];
        sT[tn * TS + p * kTP + te] = t;
      }
    }
    __syncthreads();

    // ---- G: owner-side wave speed of the next stage, max over the tile,
    // and the traces of the cross-tile faces to the neighbours' slots of the
    // next halo-trace buffer (at the neighbour's point order); both lists
    // are in LDS ----
    const int nw = nwc * NpEdge, nxp = nxc * NpEdge;
#pragma unroll
    for (int j = 0; j < ITG; ++j) {
      const int it = tid + j * kThreads;
      if (it < nw) {
        const int w = it / NpEdge, pp = it - w * NpEdge;
        const int packed = sWP[w];
        const int e = packed & 255, f = packed >> 8, p = f * NpEdge + pp;
        const int kk = k0 + e;
        const size_t fi = size_t(f) * Kp + kk;
        const int code = sWC[w];
        const int kind = code & 15;
        C own[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) own[c] = sT[c * TS + p * kTP + e];
        double q[4] = {double(own[0]), double(own[1]), double(own[2]), double(own[3])};
        if (kind == FK_Far || kind == FK_IVortex)
          boundary_state(kind, ph, a.g.bxy, code >> 8, NpEdge, pp, tnext, own, a.g.Nx[fi], a.g.Ny[fi], q);
        const double wsp = point_speed(ph.gamma, q, sOK[e]);
        if (!(wsp == wsp)) bad = 1;
        wl = fmax(wl, wsp);
        if (lts) atomicMax(&sWf[e * 3 + f], bits(wsp));
      }
    }
    V4 *hout = static_cast<V4 *>(a.ht) + size_t(l.hpar ^ 1) * a.ntiles * a.Fmax * NpEdge;
#pragma unroll
    for (int j = 0; j < ITG; ++j) {
      const int it = tid + j * kThreads;
      if (it < nxp) {
        const int w = it / NpEdge, pp = it - w * NpEdge;
        const int xk = sXP[w];
        const int e = xk & 255, f = (xk >> 8) & 3, p = f * NpEdge + pp;
        hout[size_t(sXD[w]) * NpEdge + (NpEdge - 1 - pp)] =
            V4{sT[p * kTP + e], sT[TS + p * kTP + e], sT[2 * TS + p * kTP + e], sT[3 * TS + p * kTP + e]};
      }
    }
  }
  wl = wave_max(wl);
  if ((tid & 63) == 0) red[tid >> 6] = wl;
  bad = __syncthreads_or(bad);
  if constexpr (!SHOCK) {
    if (lts) {
      double *fw = a.lts.fw + size_t(l.par2 ^ 1) * 3 * Kp;
      const int *wlist = a.g.wave_list + size_t(tile) * a.Wmax;
      for (int w = tid; w < a.g.wave_cnt[tile]; w += kThreads) {
        const int packed = wlist[w];
        const int e = packed & 255, f = packed >> 8;
        fw[size_t(f) * Kp + k0 + e] = from_bits(sWf[e * 3 + f]);
      }
    }
  }
  if (tid == 0) {
    if constexpr (!SHOCK) {
      double m = red[0];
      for (int w = 1; w < kThreads / 64; ++w) m = fmax(m, red[w]);
      atomic_max_bits(&ctl->wmax[(l.ring + 1) % 3], m);
    }
    if (bad) atomicOr(&ctl->nan, 1);
    if (first) {
      ctl->wmax[(l.ring + 2) % 3] = 0ull;
      ctl->wmaxv[(l.ring + 2) % 3] = 0ull;
      ctl->dt[stage] = dt;
      if (LAST) {
        ctl->time[par ^ 1] = tnext;
        ctl->steps += 1;
      }
    }
  }
}

template <int N, class S, class SIN, class C, int FLUX, bool LAST, bool SHOCK>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(StageWaves<S, LAST, SHOCK>::value)))
k_stage(StageArgs) {
  // The argument is read through a pointer into the kernel-argument segment
  // (the struct is the only explicit argument, at offset 0): each field is
  // then loaded where it is used rather than all of them at kernel entry,
  // which halves the scalar registers spilled to vector lanes.
  const StageArgs &a = *(const StageArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const int tile = tile_of_block(a.xcd_tiles);
  if (tile >= a.ntiles) return;
  stage_tile<N, S, SIN, C, FLUX, LAST, SHOCK>(a, launch_of(a), tile);
}

// Wave speeds of the current state (start of a step sequence): one thread per
// (element, face), traces straight from HBM; each face's maximum goes to its
// case's control block (an atomic per face: this runs once per sequence) and,
// for local time stepping, to the per-face speed buffer.
template <int N, class S>
__global__ void __launch_bounds__(256) k_prologue(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NV = Sh::NV;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  const int kk = it / 3, f = it - 3 * kk;
  if (kk >= a.Kp) return;
  const int cs = a.tile_case[kk / kTile];
  const CaseInfo &ci = a.cases[cs];
  if (kk >= ci.k_end) return;
  Ctl *ctl = a.ctl + cs;
  const double *Itp = a.Interp64;
  const size_t fi = size_t(f) * a.Kp + kk;
  const int code = a.g.code[fi];
  if (!(code & FC_WAVE)) return;
  const S *q = static_cast<const S *>(a.qs) + size_t(kk) * NV;
  double wl = 0.0;
  int bad = 0;
  for (int pp = 0; pp < NpEdge; ++pp) {
    const int p = f * NpEdge + pp;
    double tr[4];
    for (int n = 0; n < 4; ++n) {
      double t = 0;
      for (int i = 0; i < NpInt; ++i) t += Itp[p * NpInt + i] * double(q[n * NpInt + i]);
      tr[n] = t;
    }
    const int kind = code & 15;
    if (kind == FK_Far || kind == FK_IVortex) {
      const double own[4] = {tr[0], tr[1], tr[2], tr[3]};
      boundary_state(kind, ci.ph, a.g.bxy, code >> 8, NpEdge, pp, ctl->time[a.step_index & 1], own,
                     a.g.Nx[fi], a.g.Ny[fi], tr);
    }
    const double w = point_speed(ci.ph.gamma, tr, a.g.oohK[kk]);
    if (!(w == w)) bad = 1;
    wl = fmax(wl, w);
  }
  if (a.lts.on) a.lts.fw[size_t(a.par2) * 3 * a.Kp + fi] = wl;
  atomic_max_bits(&ctl->wmax[a.ring], wl);
  if (bad) atomicOr(&ctl->nan, 1);
}

// Halo traces of a state set outside the step sequence (a restart, an RHS
// evaluation), into halo-trace buffer a.hpar: thread (tile, export item,
// variable) interpolates one variable of one element to the NpEdge points of
// one cross-tile face, in the stage kernel's arithmetic type C, and stores
// them at the neighbour's point order. The stage kernel does the same for
// its new state after every stage.
template <int N, class SIN, class C>
__global__ void __launch_bounds__(256) k_halo_export(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NV = Sh::NV;
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = it & 3, item = it >> 2;
  const int tile = item / a.Xmax, w = item - tile * a.Xmax;
  if (tile >= a.ntiles || w >= a.g.xp_cnt[tile]) return;
  const int packed = a.g.xp_list[size_t(tile) * a.Xmax + w];
  const size_t dst = size_t(a.g.xp_dst[size_t(tile) * a.Xmax + w]);
  const int e = packed & 255, f = (packed >> 8) & 3;
  const SIN *qn = static_cast<const SIN *>(a.qs) + size_t(tile * kTile + e) * NV + n * NpInt;
  C q[NpInt];
#pragma unroll
  for (int i = 0; i < NpInt; ++i) q[i] = C(qn[i]);
  const C *Itp = Tabs<N, C>::ops() + Sh::NpFlux * NpInt;
  C *hout = static_cast<C *>(a.ht) + size_t(a.hpar) * a.ntiles * a.Fmax * NpEdge * 4;
#pragma unroll
  for (int pp = 0; pp < NpEdge; ++pp) {
    const C *row = Itp + (f * NpEdge + pp) * NpInt;
    C t = C(0);
#pragma unroll
    for (int i = 0; i < NpInt; ++i) t += row[i] * q[i];
    hout[(dst * NpEdge + (NpEdge - 1 - pp)) * 4 + n] = t;
  }
}

// ---- Shock capturing ----
// SIN: storage of the stage state the passes read; the sensor, the element
// fields and the state filter compute in FP64; the gradient pass computes
// and stores in D, the stage kernel's arithmetic type.
template <class D>
__device__ __forceinline__ const D *shock_ops(const ShockArgs &s);
template <>
__device__ __forceinline__ const double *shock_ops<double>(const ShockArgs &s) { return s.ops; }
template <>
__device__ __forceinline__ const float *shock_ops<float>(const ShockArgs &s) { return s.opsF; }

// Pass 0: Persson sensor on the density (UpdateSeMoment,
// DG2D/dfr_shock_capturing.go:142-166), sigma (UpdateShockFinderSigma,
// dissipation.go:491-518, S0 = 4/(N+1)^4), element viscosity
// (CalculateElementViscosity :397-412), and the max-merge of both to the
// vertices (MergeElementScalarToVertices euler.go:1048-1065) by atomics on
// the IEEE bits of the non-negative values. One thread per element.
template <int N, class SIN>
__global__ void __launch_bounds__(256) k_shock_sensor(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NV = Sh::NV;
  const int kk = blockIdx.x * blockDim.x + threadIdx.x;
  // the element-fields pass of this stage builds the active-element list
  if (kk == 0) *a.sh.act_cnt = 0u;
  if (kk >= a.Kp) return;
  // the element's density and size go out first, under the chain element ->
  // case -> control block of the skip check (kk < Kp keeps them in bounds)
  const SIN *rho = static_cast<const SIN *>(a.qs) + size_t(kk) * NV;
  double r[NpInt];
#pragma unroll
  for (int i = 0; i < NpInt; ++i) r[i] = rho[i];
  const double hK = a.sh.hK[kk];
  // the element's case (a batch holds several, each with its own control)
  const int cs = a.tile_case[kk / kTile];
  const CaseInfo &ci = a.cases[cs];
  if (kk >= ci.k_end || step_skipped(a, a.ctl + cs, ci.ph.final_time)) return;
  ConstD *Pm = shock_ops_c(a.sh) + Sh::oP, *Mm = shock_ops_c(a.sh) + Sh::oM, *Dm = shock_ops_c(a.sh) + Sh::oD;
  double num = 0, den = 0;
#pragma unroll
  for (int i = 0; i < NpInt; ++i) {
    double X = 0, Y = 0, DQ = 0;
#pragma unroll
    for (int j = 0; j < NpInt; ++j) {
      X += Pm[i * NpInt + j] * r[j];
      Y += Mm[i * NpInt + j] * r[j];
      DQ += Dm[i * NpInt + j] * r[j];
    }
    num += DQ * X;
    den += r[i] * Y;
  }
  const double se = log10(num / den);
  const double kappa = a.sh.kappa, S0 = 4. / double((N + 1) * (N + 1) * (N + 1) * (N + 1));
  const double left = S0 - kappa, right = S0 + kappa, ook = 0.5 / kappa;
  double sg = 0.;
  if (se < left) sg = 0.;
  else if (se >= left && se <= right) sg = 0.5 * (1. + sin(M_PI * ook * (se - S0)));
  else if (se > right) sg = 1.;
  const double eps = a.sh.eps0 * hK * sg;
  a.sh.Se[kk] = se;
  a.sh.eps[kk] = eps;
  unsigned long long *sv = a.sh.sigV + size_t(a.sh.vbuf) * a.sh.Nv;
  unsigned long long *ev = a.sh.epsV + size_t(a.sh.vbuf) * a.sh.Nv;
  // the buffers start at zero: a zero sigma (and so eps) changes no maximum
  if (sg > 0.) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int v = a.sh.evert[size_t(c) * a.Kp + kk];
      atomicMax(&sv[v], bits(sg));
      atomicMax(&ev[v], bits(eps));
    }
  }
}

// Pass 1: sigma = mean of the vertex maxima (MergeVertexScalarToElement),
// eps at the RT points by barycentric interpolation of the vertex maxima
// (InterpolateEpsilonSigma dissipation.go:219-242), the modal filter of the
// stage state at rkStep 2 (euler.go:603-608), and clearing of the other
// vertex buffer for the next stage. One thread per element (padded range).
template <int N, class SIN>
__global__ void __launch_bounds__(256) k_shock_prep(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NV = Sh::NV, NR = Sh::NR;
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  {
    unsigned long long *so = a.sh.sigV + size_t(a.sh.vbuf ^ 1) * a.sh.Nv;
    unsigned long long *eo = a.sh.epsV + size_t(a.sh.vbuf ^ 1) * a.sh.Nv;
    for (int v = gid; v < a.sh.Nv; v += gridDim.x * blockDim.x) {
      so[v] = 0ull;
      eo[v] = 0ull;
    }
  }
  const int kk = gid;
  if (kk >= a.Kp) return;
  // the element's vertices go out first, under the chain element -> case ->
  // control block of the skip check (kk < Kp keeps them in bounds)
  const int v0 = a.sh.evert[kk], v1 = a.sh.evert[size_t(a.Kp) + kk], v2 = a.sh.evert[2 * size_t(a.Kp) + kk];
  const int cs = a.tile_case[kk / kTile];
  const CaseInfo &ci = a.cases[cs];
  if (step_skipped(a, a.ctl + cs, ci.ph.final_time)) return;
  double *epsF = a.sh.epsF + kk;  // [NR][Kp]
  if (kk >= ci.k_end) {  // padding
    a.sh.sigma[kk] = 0.;
    a.sh.active[kk] = 0;
    return;
  }
  const unsigned long long *sv = a.sh.sigV + size_t(a.sh.vbuf) * a.sh.Nv;
  const unsigned long long *ev = a.sh.epsV + size_t(a.sh.vbuf) * a.sh.Nv;
  const double sg = ((from_bits(sv[v0]) + from_bits(sv[v1])) + from_bits(sv[v2])) / 3.;
  a.sh.sigma[kk] = sg;
  const double e0 = from_bits(ev[v0]), e1 = from_bits(ev[v1]), e2 = from_bits(ev[v2]);
  const bool act = e0 > 0. || e1 > 0. || e2 > 0.;
  a.sh.active[kk] = act ? 1 : 0;
  {
    // append to the active-element list: one atomic per wave
    const unsigned long long m = __ballot(act);
    if (m) {
      const int lane = threadIdx.x & 63, leader = __ffsll(m) - 1;
      unsigned base = 0;
      if (lane == leader) base = atomicAdd(a.sh.act_cnt, unsigned(__popcll(m)));
      base = __shfl(base, leader, 64);
      if (act) a.sh.act_list[base + unsigned(__popcll(m & ((1ull << lane) - 1ull)))] = kk;
    }
  }
  // an inactive element's eps is zero at every RT point; readers check the
  // flag instead of reading epsF
  if (act) {
    ConstD *B = shock_ops_c(a.sh) + Sh::oBary;
#pragma unroll
    for (int r = 0; r < NR; ++r) epsF[size_t(r) * a.Kp] = B[3 * r] * e0 + B[3 * r + 1] * e1 + B[3 * r + 2] * e2;
  }
  if (a.sh.filter_state) {
    ConstD *V = shock_ops_c(a.sh) + Sh::oV, *Vi = shock_ops_c(a.sh) + Sh::oVi, *mf = shock_ops_c(a.sh) + Sh::omf;
    const double fac = 1. - sin(0.5 * M_PI * sg);
    SIN *q = static_cast<SIN *>(a.qs) + size_t(kk) * NV;
    for (int n = 0; n < 4; ++n) {
      double u[NpInt], m[NpInt];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) u[i] = q[n * NpInt + i];
#pragma unroll
      for (int i = 0; i < NpInt; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < NpInt; ++j) t += Vi[i * NpInt + j] * u[j];
        m[i] = i == 0 ? t : t * (mf[i] * fac);
      }
#pragma unroll
      for (int i = 0; i < NpInt; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < NpInt; ++j) t += V[i * NpInt + j] * m[j];
        q[n * NpInt + i] = SIN(t);
      }
    }
  }
}

// Pass 2: the stage's owner-side wave speeds (|U|+c)/hK and viscous speeds
// eps/hK^2 (StoreEdgeAggregates edges.go:246-289) from the own traces, one
// tile of 64 elements x 4 variables per workgroup; the tile's state is
// staged through LDS with coalesced loads.
template <int N, class SIN, class D>
__global__ void __launch_bounds__(kThreads) k_shock_grad(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux, NF = Sh::NF, NV = Sh::NV,
                E = kTile, P = kTile + 1;
  // the tile state [NV][P], then (aliased) the traces [4][NF][E]
  constexpr int LW = NV * P > 4 * NF * E ? NV * P : 4 * NF * E;
  __shared__ D sbuf[LW];
  __shared__ double red[2][kThreads / 64];
  const int tile = blockIdx.x, tid = threadIdx.x;
  const int te = tid % E, tn = tid / E, k0 = tile * E, Kp = a.Kp;
  const SIN *Q = static_cast<const SIN *>(a.qs);
  const D *Itp = Tabs<N, D>::ops() + NpFlux * NpInt;
  // The tile's state and the owned faces' codes, 1/hK and active flags do
  // not depend on the control block (every tile lies inside the padded
  // range): their loads go out first, under the chain tile -> case ->
  // control block of the skip check.
  constexpr int LPT = (E * NV + kThreads - 1) / kThreads;
  SIN v[LPT];
  {
    const SIN *qt = Q + size_t(k0) * NV;
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      v[j] = idx < E * NV ? qt[idx] : SIN(0);
    }
  }
  const int we = tid / 3, wf = tid - 3 * we;  // (element, face) of the wave-speed threads
  int wcode = 0;
  double woohK = 0.0;
  bool wact = false;
  if (tid < 3 * E) {
    wcode = a.g.code[size_t(wf) * Kp + k0 + we];
    woohK = a.g.oohK[k0 + we];
    wact = a.sh.active[k0 + we] != 0;
  }
  // the tile's case (a batch holds several, each with its own control)
  const int cs = a.tile_case[tile];
  const CaseInfo &ci = a.cases[cs];
  Ctl *ctl = a.ctl + cs;
  if (a.stage < 0 || step_skipped(a, ctl, ci.ph.final_time)) return;
  {
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int idx = tid + j * kThreads;
      if (idx < E * NV) {
        const int e = idx / NV, r = idx - e * NV;
        sbuf[r * P + e] = D(v[j]);
      }
    }
  }
  __syncthreads();
  D t[NF];
  {
    D u[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) u[i] = sbuf[(tn * NpInt + i) * P + te];
#pragma unroll
    for (int p = 0; p < NF; ++p) {
      D acc = 0;
#pragma unroll
      for (int j = 0; j < NpInt; ++j) acc += Itp[p * NpInt + j] * u[j];
      t[p] = acc;
    }
  }
  __syncthreads();  // sbuf now holds the traces
  D *sTr = sbuf;    // [4][NF][E]
#pragma unroll
  for (int p = 0; p < NF; ++p) sTr[(tn * NF + p) * E + te] = t[p];
  __syncthreads();
  // wave speeds of the owned faces: thread (element, face)
  double wl = 0.0, wv = 0.0;
  int bad = 0;
  const double time = ctl->time[a.step_index & 1];
  if (tid < 3 * E) {
    const int e = we, f = wf, k = k0 + e;
    const size_t fi = size_t(f) * Kp + k;
    const int code = wcode;
    if (k < ci.k_end && (code & FC_WAVE)) {
      const int kind = code & 15;
      const double oohK = woohK;
      const bool actk = wact;  // else eps = 0 on the face
      const double *ef = a.sh.epsF + size_t(NpInt + f * NpEdge) * Kp + k;
      double fwl = 0.0, fwv = 0.0;
      for (int pp = 0; pp < NpEdge; ++pp) {
        const int p = f * NpEdge + pp;
        double q[4] = {double(sTr[(0 * NF + p) * E + e]), double(sTr[(1 * NF + p) * E + e]),
                       double(sTr[(2 * NF + p) * E + e]), double(sTr[(3 * NF + p) * E + e])};
        if (kind == FK_Far || kind == FK_IVortex) {
          const double own[4] = {q[0], q[1], q[2], q[3]};
          boundary_state(kind, ci.ph, a.g.bxy, code >> 8, NpEdge, pp, time, own, a.g.Nx[fi], a.g.Ny[fi], q);
        }
        const double sp = point_speed(ci.ph.gamma, q, oohK);
        if (!(sp == sp)) bad = 1;
        fwl = fmax(fwl, sp);
        if (actk) fwv = fmax(fwv, oohK * oohK * ef[size_t(pp) * Kp]);
      }
      if (a.lts.on) {
        a.lts.fw[size_t(a.par2) * 3 * Kp + fi] = fwl;
        a.lts.fwv[size_t(a.par2) * 3 * Kp + fi] = fwv;
      }
      wl = fwl;
      wv = fwv;
    }
  }
  wl = wave_max(wl);
  wv = wave_max(wv);
  if ((tid & 63) == 0) {
    red[0][tid >> 6] = wl;
    red[1][tid >> 6] = wv;
  }
  bad = __syncthreads_or(bad);
  if (tid == 0) {
    double m = red[0][0], mv = red[1][0];
    for (int w = 1; w < kThreads / 64; ++w) {
      m = fmax(m, red[0][w]);
      mv = fmax(mv, red[1][w]);
    }
    atomic_max_bits(&ctl->wmax[a.ring], m);
    atomic_max_bits(&ctl->wmaxv[a.ring], mv);
    if (bad) atomicOr(&ctl->nan, 1);
  }
}

// Pass 3: eps * grad Q by the RT element (CalculateEpsilonGradient
// dissipation.go:244-272 via GetSolutionGradientUsingRTElement
// euler.go:864-918) at the interior and edge RT points of the active
// elements (elsewhere eps is zero at every RT point and the stage kernel does
// not read it), from interior DOFs Jinv*Q and edge DOFs n*IInII/Jdet times
// the owner's trace (read reversed by the non-owner). Thread = (list entry,
// variable, direction); a grid-stride loop over the list that the
// element-fields pass built. In FP32, blockIdx.y picks one of three blocks
// of RT points: a few percent of the elements are active, so one thread per
// (entry, variable, direction) would leave about one wave per SIMD to wait
// on its loads; the row blocks triple the waves in flight (each recomputes
// the element's DOFs) and keep the table rows uniform per workgroup, so they
// stay scalar loads. In FP64 the recomputed DOFs cost more than the waves
// gain (measured), so one block holds every row.
constexpr int kGradActBlocks = 1024;
template <class D>
constexpr int grad_act_row_groups() { return sizeof(D) == 4 ? 3 : 1; }
template <int N, class SIN, class D>
__global__ void __launch_bounds__(256) k_shock_grad_act(StageArgs a) {
  using Sh = Shape<N>;
  constexpr int NpInt = Sh::NpInt, NpEdge = Sh::NpEdge, NpFlux = Sh::NpFlux, NF = Sh::NF, NV = Sh::NV,
                NR = Sh::NR;
  const size_t n = size_t(*a.sh.act_cnt) * 8;
  const int Kp = a.Kp;
  const SIN *Q = static_cast<const SIN *>(a.qs);
  const D *Itp = Tabs<N, D>::ops() + NpFlux * NpInt;
  // the RT divergence rows through the constant address space (read-only
  // here): uniform scalar loads, where a generic pointer, which the stores
  // below may alias, leaves them vector loads awaited row block by row block
  using CD = const __attribute__((address_space(4))) D;
  CD *Div = (CD *)(shock_ops<D>(a.sh) + Sh::oDiv);
  for (size_t it = size_t(blockIdx.x) * blockDim.x + threadIdx.x; it < n; it += size_t(gridDim.x) * blockDim.x) {
    const int kk = a.sh.act_list[it >> 3], tn = int(it >> 1) & 3, dir = int(it) & 1;
    D u[NpInt];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) u[i] = D(Q[size_t(kk) * NV + tn * NpInt + i]);
    // edge values: the owner's trace at this face's points, in this face's
    // order. The three faces' neighbour entries, then the three source
    // states (the element's own where it owns the face, so every address is
    // valid), are loaded before any face branches: one latency each instead
    // of two per face one after another.
    bool own[3];
    int nbf[3];
    const SIN *qsrc[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const size_t fi = size_t(f) * Kp + kk;
      const int code = a.g.code[fi];
      const int nb = a.sh.nbr_elem[fi];
      nbf[f] = a.sh.nbr_face[fi];
      // unshared faces (boundary, padding) are always their own edge's owner
      own[f] = (code & FC_FLUX_OWNER) || (code & 15) != FK_Shared;
      qsrc[f] = Q + size_t(own[f] ? kk : nb) * NV + tn * NpInt;
    }
    D un[3][NpInt];
#pragma unroll
    for (int f = 0; f < 3; ++f)
#pragma unroll
      for (int j = 0; j < NpInt; ++j) un[f][j] = D(qsrc[f][j]);
    D ue[NF];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      if (own[f]) {
#pragma unroll
        for (int i = 0; i < NpEdge; ++i) {
          D t = 0;
#pragma unroll
          for (int j = 0; j < NpInt; ++j) t += Itp[(f * NpEdge + i) * NpInt + j] * u[j];
          ue[f * NpEdge + i] = t;
        }
      } else {
        const int nf = nbf[f];
#pragma unroll
        for (int i = 0; i < NpEdge; ++i) {
          const int p = nf * NpEdge + (NpEdge - 1 - i);
          D t = 0;
#pragma unroll
          for (int j = 0; j < NpInt; ++j) t += Itp[p * NpInt + j] * un[f][j];
          ue[f * NpEdge + i] = t;
        }
      }
    }
    const double Jd = a.g.Jdet[kk], ooJd = 1. / Jd;
    const D jr = D(a.g.Jinv[size_t(dir) * Kp + kk]), js = D(a.g.Jinv[size_t(2 + dir) * Kp + kk]);
    D dof[NpFlux];
#pragma unroll
    for (int i = 0; i < NpInt; ++i) {
      dof[i] = jr * u[i];
      dof[NpInt + i] = js * u[i];
    }
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const size_t fi = size_t(f) * Kp + kk;
      const double nrm = dir == 0 ? a.sh.Nxo[fi] : a.sh.Nyo[fi];
      const D w = D(ooJd * nrm * a.g.IInII[fi]);
#pragma unroll
      for (int i = 0; i < NpEdge; ++i) dof[2 * NpInt + f * NpEdge + i] = w * ue[f * NpEdge + i];
    }
    const double *epsF = a.sh.epsF + kk;                                   // [NR][Kp]
    D *dout = static_cast<D *>(a.sh.diss) + size_t(tn * 2 * NR) * Kp + kk;  // [4][2][NR][Kp]
    constexpr int RG = grad_act_row_groups<D>();
    if constexpr (RG == 1) {
#pragma unroll 3
      for (int r = 0; r < NR; ++r) {
        D acc = 0;
#pragma unroll
        for (int j = 0; j < NpFlux; ++j) acc += Div[r * NpFlux + j] * dof[j];
        dout[size_t(dir * NR + r) * Kp] = acc * D(epsF[size_t(r) * Kp]);
      }
    } else {
      constexpr int RPG = (NR + RG - 1) / RG;
      const int r0 = int(blockIdx.y) * RPG;
#pragma unroll
      for (int rr = 0; rr < RPG; ++rr) {
        const int r = r0 + rr;
        if (r >= NR) break;
        D acc = 0;
#pragma unroll
        for (int j = 0; j < NpFlux; ++j) acc += Div[r * NpFlux + j] * dof[j];
        dout[size_t(dir * NR + r) * Kp] = acc * D(epsF[size_t(r) * Kp]);
      }
    }
  }
}

// ---- launchers ----

template <int N>
size_t lds_bytes_n(size_t csize, int Hmax, int Fmax, int WXmax, bool lts, bool mf) {
  return lds_tile_bytes<N>(csize, Hmax, Fmax, WXmax, mf) + (lts ? lds_lts_bytes : 0);
}

template <int N, class S, class SIN, class C, int FLUX, bool LAST, bool SHOCK>
hipError_t launch_stage_t(const StageArgs &a, hipStream_t s) {
  // one thread stages each list item into LDS
  if (a.Fmax > kThreads || (!SHOCK && (a.Wmax > kThreads || a.Xmax > kThreads))) return hipErrorInvalidValue;
  const size_t lds = lds_bytes_n<N>(sizeof(C), SHOCK ? a.Hmax : 0, a.Fmax, SHOCK ? 0 : a.Wmax + a.Xmax,
                                    a.lts.on != 0, UseMfma<N, C, SHOCK>::value);
  static size_t attr_bytes = 0;  // dynamic LDS the kernel is currently allowed
  if (lds > 64 * 1024 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage<N, S, SIN, C, FLUX, LAST, SHOCK>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    attr_bytes = lds;
  }
  const int q = (a.ntiles + 7) / 8;
  hipLaunchKernelGGL((k_stage<N, S, SIN, C, FLUX, LAST, SHOCK>), dim3(8 * q), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

template <int N, class S, class C, int FLUX, bool SHOCK>
hipError_t launch_stage_l(const StageArgs &a, hipStream_t s) {
  if (a.stage == 4) return launch_stage_t<N, S, S, C, FLUX, true, SHOCK>(a, s);
  if (a.stage <= 0) return launch_stage_t<N, S, double, C, FLUX, false, SHOCK>(a, s);  // input q0
  return launch_stage_t<N, S, S, C, FLUX, false, SHOCK>(a, s);
}

template <int N, class S, class C, bool SHOCK>
hipError_t launch_stage_f(const StageArgs &a, hipStream_t s) {
  switch (a.ph.flux) {
    case 0: return launch_stage_l<N, S, C, 0, SHOCK>(a, s);
    case 1: return launch_stage_l<N, S, C, 1, SHOCK>(a, s);
    case 2: return launch_stage_l<N, S, C, 2, SHOCK>(a, s);
    case 3: return launch_stage_l<N, S, C, 3, SHOCK>(a, s);
  }
  return hipErrorInvalidValue;
}

// Arithmetic type of the FP32-bulk precision: FP32 from N=2 on. At N<=1 the
// RT edge points sit 0.056 apart and |DivInt| reaches 1.2e3, so the
// divergence cancels three digits and runs in FP64 (the elements are small
// there).
template <int N>
using BulkC = typename std::conditional<(N <= 1), double, float>::type;

template <int N>
hipError_t launch_stage_n(int prec, int shock, const StageArgs &a, hipStream_t s) {
  if (prec == 0) {
    if (shock) return launch_stage_f<N, double, double, true>(a, s);
    return launch_stage_f<N, double, double, false>(a, s);
  }
  if (shock) return launch_stage_f<N, float, BulkC<N>, true>(a, s);
  return launch_stage_f<N, float, BulkC<N>, false>(a, s);
}

template <int N>
hipError_t launch_prologue_n(int prec, const StageArgs &a, hipStream_t s) {
  const long items = long(a.Kp) * 3;
  const int blocks = int((items + 255) / 256);
  (void)prec;  // the prologue reads q0, which is FP64 in every precision
  hipLaunchKernelGGL((k_prologue<N, double>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int N>
hipError_t launch_halo_export_n(int prec, const StageArgs &a, hipStream_t s) {
  if (!a.ht) return hipErrorInvalidValue;
  const long items = long(a.ntiles) * a.Xmax * 4;
  const int blocks = int((items + 255) / 256);
  if (prec == 0) hipLaunchKernelGGL((k_halo_export<N, double, double>), dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_halo_export<N, double, BulkC<N>>), dim3(blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <int N, class SIN, class D>
hipError_t launch_shock_t(int pass, const StageArgs &a, hipStream_t s) {
  switch (pass) {
    case 0:
      hipLaunchKernelGGL((k_shock_sensor<N, SIN>), dim3((a.K + 255) / 256), dim3(256), 0, s, a);
      break;
    case 1:
      hipLaunchKernelGGL((k_shock_prep<N, SIN>), dim3((a.Kp + 255) / 256), dim3(256), 0, s, a);
      break;
    case 2:
      hipLaunchKernelGGL((k_shock_grad<N, SIN, D>), dim3(a.ntiles), dim3(kThreads), 0, s, a);
      break;
    case 3:
      hipLaunchKernelGGL((k_shock_grad_act<N, SIN, D>),
                         dim3(std::max(1, std::min(kGradActBlocks, (a.Kp * 8 + 255) / 256)), grad_act_row_groups<D>()),
                         dim3(256), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// The passes read the stage input: q0 (FP64) at stage 0 and for the RHS
// evaluation, else the stage state in the precision's storage.
template <int N>
hipError_t launch_shock_n(int prec, int pass, const StageArgs &a, hipStream_t s) {
  if (prec == 0) return launch_shock_t<N, double, double>(pass, a, s);
  if (a.stage <= 0) return launch_shock_t<N, double, BulkC<N>>(pass, a, s);
  return launch_shock_t<N, float, BulkC<N>>(pass, a, s);
}

// The order-N tables of the including translation unit (internal linkage:
// every unit that instantiates a stage body uploads its own copies)
template <int N>
static hipError_t upload_tables(const double *divT, const double *interp) {
  using Sh = Shape<N>;
  // A fragments of the matrix-core contractions in the point order of the
  // arithmetic type (MfShape): the FluxEdgeInterp [row block][K-step][lane]
  // and the DivInt [K-step][lane] (r components, s components, face DOFs)
  auto fragments = [&](auto tag, std::vector<double> &ai, std::vector<double> &af) {
    using MS = MfShape<N, decltype(tag)>;
    static_assert(MS::KS * 64 <= kMfMax && Sh::RB * MS::KI * 64 <= kMfIMax, "MFMA operand tables");
    ai.assign(size_t(Sh::RB) * MS::KI * 64, 0.);
    for (int rb = 0; rb < Sh::RB; ++rb)
      for (int s = 0; s < MS::KI; ++s)
        for (int l = 0; l < 64; ++l) {
          const int p = 16 * rb + l % 16, i = MS::point(l / 16, s);
          if (p < Sh::NF && i < Sh::NpInt) ai[(size_t(rb) * MS::KI + s) * 64 + l] = interp[size_t(p) * Sh::NpInt + i];
        }
    af.assign(size_t(MS::KS) * 64, 0.);
    for (int s = 0; s < MS::KS; ++s)
      for (int l = 0; l < 64; ++l) {
        const int row = l % 16, kr = l / 16;
        int j = -1;
        if (s < 2 * MS::KI) {
          const int i = MS::point(kr, s / 2);
          if (i < Sh::NpInt) j = (s % 2) * Sh::NpInt + i;
        } else {
          const int p = (s - 2 * MS::KI) + Sh::KE * kr;
          if (p < Sh::NF) j = 2 * Sh::NpInt + p;
        }
        if (row < Sh::NpInt && j >= 0) af[size_t(s) * 64 + l] = divT[size_t(j) * Sh::NpInt + row];
      }
  };
  std::vector<double> ai, af;
  fragments(float(), ai, af);
  {
    const std::vector<float> aif(ai.begin(), ai.end()), aff(af.begin(), af.end());
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cMfI<N>), aif.data(), aif.size() * sizeof(float), 0,
                                     hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(cMfA<N>), aff.data(), aff.size() * sizeof(float), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  fragments(double(), ai, af);
  {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cMfID<N>), ai.data(), ai.size() * sizeof(double), 0,
                                     hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    e = hipMemcpyToSymbol(HIP_SYMBOL(cMfAD<N>), af.data(), af.size() * sizeof(double), 0, hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  }
  std::vector<double> d(static_cast<size_t>(Sh::ops));
  std::copy(divT, divT + size_t(Sh::NpFlux) * Sh::NpInt, d.begin());
  std::copy(interp, interp + size_t(Sh::NF) * Sh::NpInt, d.begin() + size_t(Sh::NpFlux) * Sh::NpInt);
  std::vector<float> f(d.begin(), d.end());
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cOpsD<N>), d.data(), d.size() * sizeof(double), 0,
                                   hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(cOpsF<N>), f.data(), f.size() * sizeof(float), 0, hipMemcpyHostToDevice);
}

template <int N>
hipError_t upload_operators_n(const double *divT, const double *interp) {
  return upload_tables<N>(divT, interp);
}

}  // namespace dev
}  // namespace gocfd

// Explicit instantiation of the per-order entry points of one order.
#define GOCFD_INSTANTIATE_ORDER(NN)                                                                      \
  namespace gocfd {                                                                                      \
  namespace dev {                                                                                        \
  template hipError_t launch_stage_n<NN>(int, int, const StageArgs &, hipStream_t);                      \
  template hipError_t launch_prologue_n<NN>(int, const StageArgs &, hipStream_t);                        \
  template hipError_t launch_halo_export_n<NN>(int, const StageArgs &, hipStream_t);                     \
  template hipError_t launch_shock_n<NN>(int, int, const StageArgs &, hipStream_t);                      \
  template hipError_t upload_operators_n<NN>(const double *, const double *);                           \
  }                                                                                                      \
  }

// ---- All non-shock order groups of an ensemble in one launch per stage ----
// (a translation unit per flux includes this header with
// GOCFD_MULTI_FLUX defined and instantiates its entry points)
#ifdef GOCFD_MULTI_FLUX
namespace gocfd {
namespace dev {

template <int N, class S>
using MultiC = typename std::conditional<(sizeof(S) == 8 || N <= 1), double, float>::type;

// KIND 0: stage 0 (input q0), 1: stages 1-3, 2: stage 4
template <class S, int FLUX, int KIND>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(StageWaves<S, KIND == 2, false>::value)))
k_stage_multi(MultiArgs m) {
  const int t = tile_of_block(m.xcd_tiles);
  if (t >= m.ntiles) return;
  int g = 0;
  while (g + 1 < m.ngroups && t >= m.g[g + 1].tile0) ++g;
  const StageArgs &a = *m.g[g].a;
  const int st = m.l.stage;
  const StageLaunch l{st, m.l.step_index, m.l.ring, m.l.par2, m.l.hpar,
                      st == 0 ? a.q0 : st == 1 ? a.q1 : st == 2 ? a.q2 : st == 3 ? a.q3 : a.q4, 0., nullptr};
  const int tile = t - m.g[g].tile0;
  using SIN = typename std::conditional<KIND == 0, double, S>::type;
  constexpr bool LAST = KIND == 2;
  switch (m.g[g].N) {
    case 0: stage_tile<0, S, SIN, MultiC<0, S>, FLUX, LAST, false>(a, l, tile); break;
    case 1: stage_tile<1, S, SIN, MultiC<1, S>, FLUX, LAST, false>(a, l, tile); break;
    case 2: stage_tile<2, S, SIN, MultiC<2, S>, FLUX, LAST, false>(a, l, tile); break;
    case 3: stage_tile<3, S, SIN, MultiC<3, S>, FLUX, LAST, false>(a, l, tile); break;
    default: stage_tile<4, S, SIN, MultiC<4, S>, FLUX, LAST, false>(a, l, tile); break;
  }
}

template <class S, int FLUX, int KIND>
static hipError_t launch_multi_k(const MultiArgs &m, hipStream_t s) {
  size_t lds = 0;
  for (int g = 0; g < m.ngroups; ++g) {
    const MultiGroup &G = m.g[g];
    const size_t cs = sizeof(S) == 8 || G.N <= 1 ? sizeof(double) : sizeof(float);
    size_t b = 0;
    switch (G.N) {
      case 0: b = lds_bytes_n<0>(cs, 0, G.Fmax, G.WXmax, G.lts != 0, false); break;
      case 1: b = lds_bytes_n<1>(cs, 0, G.Fmax, G.WXmax, G.lts != 0, false); break;
      case 2: b = lds_bytes_n<2>(cs, 0, G.Fmax, G.WXmax, G.lts != 0, false); break;
      case 3: b = lds_bytes_n<3>(cs, 0, G.Fmax, G.WXmax, G.lts != 0, true); break;
      default: b = lds_bytes_n<4>(cs, 0, G.Fmax, G.WXmax, G.lts != 0, true); break;
    }
    lds = std::max(lds, b);
  }
  static size_t attr_bytes = 0;  // dynamic LDS the kernel is currently allowed
  if (lds > 64 * 1024 && lds > attr_bytes) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_stage_multi<S, FLUX, KIND>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    attr_bytes = lds;
  }
  const int q = (m.ntiles + 7) / 8;
  hipLaunchKernelGGL((k_stage_multi<S, FLUX, KIND>), dim3(8 * q), dim3(kThreads), lds, s, m);
  return hipGetLastError();
}

template <>
hipError_t launch_stage_multi_f<GOCFD_MULTI_FLUX>(int prec, const MultiArgs &m, hipStream_t s) {
  // FP64 only: the FP32 variant faulted the GPU in its first run and stays
  // off until that is explained
  if (prec != 0) return hipErrorNotSupported;
  for (int g = 0; g < m.ngroups; ++g)
    if (m.g[g].Fmax > kThreads || m.g[g].WXmax > 2 * kThreads) return hipErrorInvalidValue;
  const int kind = m.l.stage == 0 ? 0 : m.l.stage == 4 ? 2 : 1;
  if (prec == 0) {
    if (kind == 0) return launch_multi_k<double, GOCFD_MULTI_FLUX, 0>(m, s);
    if (kind == 2) return launch_multi_k<double, GOCFD_MULTI_FLUX, 2>(m, s);
    return launch_multi_k<double, GOCFD_MULTI_FLUX, 1>(m, s);
  }
  return hipErrorNotSupported;
}

template <>
hipError_t upload_operators_multi_f<GOCFD_MULTI_FLUX>(int N, const double *divT, const double *interp) {
  switch (N) {
    case 0: return upload_tables<0>(divT, interp);
    case 1: return upload_tables<1>(divT, interp);
    case 2: return upload_tables<2>(divT, interp);
    case 3: return upload_tables<3>(divT, interp);
    case 4: return upload_tables<4>(divT, interp);
  }
  return hipErrorInvalidValue;
}

}  // namespace dev
}  // namespace gocfd
#endif
