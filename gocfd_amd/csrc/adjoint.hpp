// Tangent-linear and discrete-adjoint kernels of the FP64 SSP54 step
// (global or local time stepping; PerssonC0 with global time stepping, or
// with local time stepping while its sensor is quiet):
// interface between the host driver
// (adjoint_runner.cpp) and the per-order kernel translation units.
//
// The differentiated map is one SSP54 step of the reference (rkAdvance,
// model_problems/Euler2D/euler.go:460-566) with the per-stage global time
// step dt_s = CFL / max wave speed (calculateGlobalDT euler.go:945-971,
// StoreEdgeAggregates edges.go:246-289) of the stage input, or the element
// dts of local time stepping (CalculateLocalDT euler.go:973-1002), the DFR
// right-hand side (RHSInternalPoints euler.go:665-699, SetRTFluxInternal
// euler.go:701-726, CalculateEdgeEulerFlux edges.go:325-452) and the
// Riemann-invariant far field (RiemannBC bcs.go:70-133). Its inputs are the
// step-start state and the far-field state FSFar.Qinf.
//
// The forward kernels are templated on the arithmetic type: double (the
// recomputation of the stages for the adjoint) or ad::Dual (the
// tangent-linear step). The adjoint kernels apply the transposed pointwise
// Jacobians (numerical flux, boundary state, physical flux, wave speed),
// which they form one input direction at a time in dual arithmetic, and the
// transposed DivInt / FluxEdgeInterp contractions and SSP combinations.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace gocfd {
namespace dev {

// Device storage of one dual value (ad::Dual in the kernels)
struct ad_storage {
  double v, d;
};

// SSP54 coefficients (rkAdvance, model_problems/Euler2D/euler.go:511-565)
constexpr double kSspA0 = 0.391752226571890;
constexpr double kSspB1 = 0.444370493651235, kSspC1 = 0.555629506348765, kSspA1 = 0.368410593050371;
constexpr double kSspB2 = 0.620101851488403, kSspC2 = 0.379898148511597, kSspA2 = 0.251891774271694;
constexpr double kSspB3 = 0.178079954393132, kSspC3 = 0.821920045606868, kSspA3 = 0.544974750228521;
constexpr double kSspD2 = 0.517231671970585, kSspD3 = 0.096059710526146, kSspD4 = 0.386708617503269;
constexpr double kSspE3 = 0.063692468666290, kSspE4 = 0.226007483236906;

struct AdjArgs {
  int N, K, Kp;
  const double *Jdet, *Jinv, *IInII, *Nx, *Ny, *oohK;  // [Kp], [4][Kp], [3][Kp] x3, [Kp]
  const int *code, *nbr_elem, *nbr_face;               // [3][Kp]
  const double *Interp;                                // FluxEdgeInterp [3 NpEdge][NpInt]
  const double *DivInt;                                // [NpInt][NpFlux]
  Phys ph;
  int io_tied;     // In/Out faces use the far state (every case but the shock tube)
  double dfar[4];  // tangent of the far state (tangent-linear step)
  // local time stepping (CalculateLocalDT, euler.go:973-1002): element dt
  // = CFL / max(previous stage's element dt, speeds of its faces), the face
  // (f Kp + kk) speed read from face wsrc[f Kp + kk] (the edge's owner side)
  int lts;
  const int *wsrc;
  // unshared face points of real elements (boundary kinds), t = (f Kp + kk) NpEdge + pp
  const int *bnd;
  int nbnd;
  // PerssonC0 with local time stepping, while its sensor is quiet (sigma =
  // 0 everywhere): the modal
  // filter F = V diag(1, mf_1, mf_2, ...) Vinv [NpInt][NpInt] of the RHS at
  // every stage and of the stage state at stage 2 (limitAndFilterSolution
  // dissipation.go:606-622, euler.go:603-608); the sensor's matrices
  // P, M, D [NpInt][NpInt] each and its threshold S0 - Kappa (sigma > 0 from
  // there on, UpdateShockFinderSigma dissipation.go:491-518)
  int shock;
  const double *filt;
  const double *sens;
  double sens_left;
  // PerssonC0 with an active sensor (global time stepping): the sensor
  // ramp sigma(Se) (UpdateShockFinderSigma dissipation.go:491-518), the
  // element viscosity eps = eps0 hK sigma (CalculateElementViscosity
  // :397-412), the vertex max-merges (euler.go:1048-1065), the vertex-mean
  // sigma's modal filters (dissipation.go:606-622), eps at the RT points
  // (InterpolateEpsilonSigma :219-242), the RT gradient times eps
  // (CalculateEpsilonGradient :244-272, euler.go:864-918), the viscous edge
  // flux (StoreEdgeViscousFlux edges.go:151-244), AddDissipation
  // (dissipation.go:274-346) and the viscous dt, all differentiated
  int act;
  const double *shops;          // the shock operator block (P, M, D, V, Vinv, mf, DivR, Bary)
  const double *hK;             // [Kp] EdgeLenMax / (N+1)^2
  const double *Nxo, *Nyo;      // [3][Kp] own outward normals
  const double *vooL;           // [3][Kp] 1 / owner edge length
  const int *evert;             // [3][Kp] vertex ids
  const int *vptr, *vinc;       // vertex -> (kk << 2 | local vertex) of its real elements, CSR
  int Nv;
  double kappa, S0, eps0, cdiff, omega;
};

// Buffers of one launch. Element arrays are element-major [Kp][4][NpInt];
// face-point arrays are [3][Kp][NpEdge] records of 4 values.
struct AdjBufs {
  int dual;             // forward kernels in dual arithmetic
  int stage;            // 0..4
  const void *Us;       // stage input
  const void *U0, *U2, *U3, *R3;  // SSP operands (step start, stage inputs 2 and 3, stage-3 RHS)
  const void *Rs;       // this stage's RHS
  void *out;            // next stage input (or the step's result)
  void *FE;             // face-point edge fluxes, scaled by ||n|| and side
  void *R;              // RHS output
  void *SP;             // face-point wave speeds
  void *dts;            // [5] stage time steps (double or dual)
  int *argmax;          // [5] face point of each stage's maximum wave speed
  double *Wv;           // [5] the maximum wave speeds
  double *time;         // step-start time (advanced by the last stage's dt)
  int *flag;            // set when a stage's dt would be clamped to FinalTime
  double *part_v;       // reduction partials (4 nparts)
  int *part_i;
  int nparts;
  // adjoint
  double *lam;          // cotangent of this stage's output
  double *lamU0;        // cotangent of the step-start state
  double *lamQ[5];      // cotangents of the stage inputs 1..4 (index 1..4)
  double *lamR;         // cotangent of this stage's RHS
  double *lamR3;        // stage-3 RHS cotangent from the last stage
  double *lamOwn;       // face-point trace cotangents
  double *lamFarPts;    // face-point cotangents of the far state (accumulated)
  double *lamFarW;      // [4] far-state cotangent from the dt terms (accumulated)
  double *lam_dt;       // [5] cotangents of the stage dts
  // local time stepping: per element of this stage
  void *dtk;            // element dt (double or dual) [Kp]
  const void *dtk_prev; // the previous stage's element dt (stage >= 1)
  int *srck;            // face point whose speed sets the dt, -1: the previous dt
  double *wk;           // the maximum (CFL / dt)
  double *carry;        // cotangent of the previous stage's element dt [Kp]
  double *lamSpeed;     // cotangents of the face-point speeds [3 Kp NpEdge]
  double *target;       // cotangent of this stage's input (accumulated)
  // active PerssonC0: forward fields of this stage (T; the records double
  // in the double sweeps) and their cotangents
  void *sig, *eps;      // [Kp] element sigma and eps
  void *sigv, *epsv;    // [Nv] vertex maxima
  int *args, *arge;     // [Nv] the elements that attain them
  void *sigb;           // [Kp] vertex-mean sigma
  void *epsf;           // [NR][Kp] eps at the used RT points
  void *dis;            // [4][2][NR][Kp] eps grad Q there
  void *SPV;            // [3 Kp NpEdge] viscous speeds of owner-side face points, -1 elsewhere
  double *Rpre;         // this stage's RHS before its modal filter (double sweeps)
  const double *Xpre;   // the pre-filter values a transposed filter differentiates at
  int *argmaxv, *vbr;   // [5] the viscous arg-max point; 1 where the viscous dt binds
  double *Wvv;          // [5] the maximum viscous speed
  double *part_v2;      // viscous arg-max partials
  int *part_i2;
  int accumulate;       // transposed filter: add to lamSigb (else set)
  const double *Usens;  // the sensor's input (the stage input before the stage-2 filter)
  double *lamSigb, *lamEpsf, *lamDis, *lamEQv, *lamEQg, *lamSigv, *lamEpsv;
  const double *pack_v, *pack_d;  // value and tangent arrays (pack/unpack)
  double *unpack_v, *unpack_d;
  size_t n;             // element-array length (Kp * 4 * NpInt) or reduction length
};

enum AdjOp : int {
  ADJ_FWD_FACE = 0,  // FE <- edge values of Us; SP <- owner-side wave speeds of Us
  ADJ_ARGMAX,        // dts[stage], argmax[stage], Wv[stage] <- from SP; time, flag
  ADJ_FWD_RHS,       // R <- RHS(Us, FE); out <- the stage's SSP54 combination
  ADJ_ADJ_SSP,       // cotangents through the SSP combination; lam_dt[stage]
  ADJ_ADJ_FACE,      // lamOwn, lamFarPts <- from lamR at Us
  ADJ_ADJ_ELEM,      // target += interior-flux and trace transposes
  ADJ_ADJ_WAVE,      // target, lamFarW += the dt term of the stage
  ADJ_FAR_SUM,       // target[0..3] <- sum of lamFarPts + lamFarW
  ADJ_FWD_LTS,       // dtk, srck, wk <- local dt of each element from SP and dtk_prev
  ADJ_ADJ_SSP_LTS,   // the SSP transpose with element dts; lamSpeed, carry <- the dt terms
  ADJ_ADJ_LTS_WAVE,  // lamOwn, lamFarPts += the face-point speed cotangents (then cleared)
  ADJ_SENSOR,        // flag |= 4 where the shock sensor of Us reaches its ramp
  ADJ_FILTER,        // out <- F out per element and variable (in place; dual or double)
  ADJ_FILTER_T,      // lam <- F^T lam per element and variable (in place)
  ADJ_FWD_SHOCK,     // sig, eps, sigv, epsv, args, arge, sigb, epsf, SPV <- the sensor path of Us
  ADJ_FILTER_SIG,    // out <- F(sigb) out per element and variable (in place; dual or double)
  ADJ_FWD_VISC,      // dis <- eps grad Q of Us; FE -= the viscous edge values
  ADJ_ADJ_FILTER_SIG_T,  // lam <- F(sigb)^T lam; lamSigb (+)= its sigma term at Xpre
  ADJ_ADJ_VISC,      // lamDis, lamEpsf, lamEQv, lamEQg, target <- viscous flux and gradient transposes;
                     // lamOwn += the owner-trace terms
  ADJ_ADJ_VDT,       // lamEpsf += the viscous dt term where it binds
  ADJ_ADJ_SENSOR,    // target += the vertex-merge, ramp and sensor transposes at Usens
  ADJ_PACK,          // out (dual) <- (pack_v, pack_d)
  ADJ_UNPACK,        // (unpack_v, unpack_d) <- Us (dual)
};

hipError_t adj_launch(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s);
template <int N> hipError_t adj_launch_n(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s);
hipError_t adj_launch_common(int op, const AdjArgs &a, const AdjBufs &b, hipStream_t s);

}  // namespace dev
}  // namespace gocfd
