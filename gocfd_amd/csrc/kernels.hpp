// Device-side interface of the 2D Euler DFR stage kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gocfd {
namespace dev {

#ifndef GOCFD_TILE
#define GOCFD_TILE 64
#endif
constexpr int kTile = GOCFD_TILE;  // elements per workgroup tile
constexpr int kThreads = 4 * kTile; // threads per workgroup: one per (element, variable)
// Row pitch (elements) of the stage kernel's LDS trace/flux rows [point][element]:
// four more than the tile, so that the face points of one element, read by
// neighbouring lanes in the Riemann and wave-speed phases, fall on different
// LDS banks (a 64-element pitch put them all on one)
constexpr int kTP = kTile + 4;
static_assert(kThreads % 64 == 0, "whole wavefronts per workgroup");

// Device control block. Time is double-buffered by step parity and the wave
// speed maxima live in rings of three slots indexed by the global stage
// counter, so that no kernel writes a value another block of the same launch
// still reads.
struct Ctl {
  double time[2];
  double dt[5];
  unsigned long long wmax[3];   // max (|U|+c)/hK, IEEE bits of a positive double
  unsigned long long wmaxv[3];  // max eps/hK^2 (shock capturing)
  int nan;
  int steps;
  int pad[2];
};

struct Geom {
  const double *Jdet, *Jinv, *IInII, *Nx, *Ny, *oohK;  // [Kp], [4][Kp], [3][Kp] x3, [Kp]
  const int *code, *nbr_slot;                          // [3][Kp]
  const int *halo_elem, *halo_face;                    // [ntiles][Hmax]
  const int *wave_list, *wave_cnt;                     // [ntiles][Wmax], [ntiles]
  const int *wave_code;                                // [ntiles][Wmax]: code of each wave_list face
  const int *flux_list, *flux_cnt;                     // [ntiles][Fmax], [ntiles]
  // per flux-list item, in the stage arithmetic type: owner normal (nx, ny),
  // this side's signed IInII, the in-tile partner's IInII negated
  const void *flux_geo;                                // [ntiles][Fmax][4]
  const int *bnd_cnt;                                  // [ntiles] non-shared faces
  const double *bxy;                                   // IVortex boundary points
  // Faces whose neighbour lies in another tile, per tile: after a stage the
  // tile writes its traces of the new state there into the neighbour's face
  // slots of the halo-trace buffer. [ntiles][Xmax]: e | face << 8, and the
  // neighbour's face index (f * Kp + kk); [ntiles] counts
  const int *xp_list, *xp_dst, *xp_cnt;
};

// A far-field state of the Riemann-invariant boundary (FreeStream,
// fluids.go:245-273)
struct FarState {
  double q[4];
  double gamma, minf, pinf, cinf;
};

struct Phys {
  double gamma, minf, pinf, cinf;  // FSFar
  double qinf[4];
  double cfl, final_time;
  FarState io[2];                  // FSIn, FSOut (In / Out boundary edges)
  int flux;
};

// One case of a batch (a single solver is a batch of one): its physics
// constants, the end of its element range and its first tile. Tiles never
// straddle cases.
struct CaseInfo {
  Phys ph;
  int k_end;   // one past the case's last element
  int tile0;   // the case's first tile (does the case's control bookkeeping)
};

// Local time stepping (CalculateLocalDT, euler.go:973-1002)
struct LtsArgs {
  int on;
  double *fw;        // [2][3][Kp] per-face max wave speed, buffer by stage parity
  double *fwv;       // [2][3][Kp] per-face max viscous speed (shock capturing)
  const int *wsrc;   // [3][Kp] face (f*Kp + kk) whose edge speed bounds this face
  double *DTk;       // [Kp] element DT of the last stage
  double *DTv;       // [Kp] element DTVisc (never reset, as in the reference)
};

// Shock capturing (PerssonC0): per-order operators and per-element fields.
// Operator block: Pm, Mass, Dm, V, Vinv [NpInt^2] each, mf [NpInt],
// DivR [NR][NpFlux] (rows 0..NpInt-1 and 2NpInt.. of Div), Bary [NR][3],
// with NR = NpInt + 3 NpEdge; in FP64 (ops) and FP32 (opsF).
// The dissipation eps*grad Q (diss) is stored in the stage kernel's
// arithmetic type: FP64, or FP32 for the FP32-bulk precision from N=2 on.
struct ShockArgs {
  const double *ops;
  const float *opsF;
  const int *evert;                  // [3][Kp] vertex ids (original numbering)
  const double *hK;                  // [Kp] EdgeLenMax / (N+1)^2
  unsigned long long *sigV, *epsV;   // [2][Nv] vertex maxima (IEEE bits)
  int Nv, vbuf;                      // vertex buffer used by this stage
  double *Se, *sigma, *eps;          // [Kp]
  double *epsF;                      // [NR][Kp] viscosity at the used RT points
  void *diss;                        // [4][2][NR][Kp] eps * grad at the used RT points
  int *active;                       // [Kp] 1 if a vertex of the element carries eps > 0;
                                     // else eps = 0 at all its RT points and diss is
                                     // neither written nor read (it is zero)
  int *act_list;                     // [Kp] the active elements of this stage (any order)
  unsigned *act_cnt;                 // their number (cleared by the sensor pass)
  const int *nbr_elem, *nbr_face;    // [3][Kp]
  const double *Nxo, *Nyo;           // [3][Kp] own outward normals
  const double *vooL;                // [3][Kp] 1 / owner edge length (shared faces)
  double kappa, eps0, cdiff, omega;
  int filter_state;                  // filter the stage state (rkStep == 2)
};

struct StageArgs {
  int K, Kp, ntiles, Hmax, Wmax, Fmax, Xmax;
  int xcd_tiles;       // 1: each XCD takes a contiguous tile range; 0: tile = workgroup id
  int stage;           // 0..4, or -1 for the RHS-only evaluation
  int step_index;      // steps since the state was set (selects time parity)
  int ring;            // global stage counter mod 3
  int par2;            // global stage counter mod 2
  int hpar;            // halo-trace buffer holding this stage's input traces
  // Halo-trace buffers [2][ntiles Fmax][NpEdge][4] in the stage arithmetic
  // type (without shock capturing): by tile and flux-list item, the traces
  // of each cross-tile face's neighbour at the face's own points, written by
  // the neighbour's tile
  void *ht;
  double t_rhs;        // time for the RHS-only evaluation
  void *qs;            // stage input state (element-major)
  void *q0, *q1, *q2, *q3, *q4, *R;
  double *rhs_out;     // RHS-only output (element-major, double)
  Ctl *ctl;            // control blocks, one per case
  const CaseInfo *cases;
  const int *tile_case;  // [ntiles] case of each tile
  Geom g;
  const double *Interp64;       // FluxEdgeInterp in FP64
  Phys ph;             // case 0's constants (launch dispatch; shock-capturing passes)
  LtsArgs lts;
  ShockArgs sh;
};

// Precision variants: 0 = FP64 storage and arithmetic; 1 = FP32 storage of
// the stage states q1..q4 and FP32 bulk arithmetic (fluxes, divergence),
// with FP64 geometry, dt reduction, SSP combination and FP64 storage of the
// accumulated state q0 and the stage-3 RHS R.
// shock != 0 selects the PerssonC0 stage kernel.
hipError_t launch_stage(int N, int prec, int shock, const StageArgs &a, hipStream_t s);
hipError_t launch_prologue(int N, int prec, const StageArgs &a, hipStream_t s);
// Halo traces of the FP64 state a.qs into buffer a.hpar (a state set outside
// the step sequence; stage launches write them for their new state)
hipError_t launch_halo_export(int N, int prec, const StageArgs &a, hipStream_t s);
// Shock-capturing passes of one stage: 0 sensor + vertex max, 1 element
// fields (+ stage-state filter, active-element list), 2 wave speeds,
// 3 eps * gradient of the active elements.
constexpr int kShockPasses = 4;
hipError_t launch_shock(int N, int prec, int pass, const StageArgs &a, hipStream_t s);
hipError_t launch_ctl_init(Ctl *ctl, double time, hipStream_t s, int ncases = 1);
hipError_t launch_fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s);
// Copy the order-N operators (DivInt^T [NpFlux][NpInt], FluxEdgeInterp
// [3NpEdge][NpInt]) into the current device's constant memory.
hipError_t upload_operators(int N, const double *divT, const double *interp);
size_t stage_lds_bytes(int N, int prec, int Hmax);

// All non-shock order groups of an ensemble in one stage launch: tiles are
// numbered group after group; each workgroup finds its group and runs that
// group's order. The group arguments live in device memory; the launch
// values (stage, step, rings) are the same for every group.
constexpr int kMultiMax = 10;
struct MultiGroup {
  const StageArgs *a;   // the group's arguments (device memory)
  int tile0;            // its first tile in the launch's numbering
  int N;
  int Fmax, WXmax, lts; // its LDS layout (host side)
};
struct MultiLaunch {
  int stage, step_index, ring, par2, hpar;
};
struct MultiArgs {
  int ngroups, ntiles, xcd_tiles;
  MultiLaunch l;
  MultiGroup g[kMultiMax];
};
hipError_t launch_stage_multi(int prec, int flux, const MultiArgs &m, hipStream_t s);
// The order-N operators in the tables of the multi-order kernels of a flux
hipError_t upload_operators_multi(int flux, int N, const double *divT, const double *interp);

// Per-order entry points (one translation unit per order)
template <int N> hipError_t launch_stage_n(int prec, int shock, const StageArgs &a, hipStream_t s);
template <int N> hipError_t launch_prologue_n(int prec, const StageArgs &a, hipStream_t s);
template <int N> hipError_t launch_halo_export_n(int prec, const StageArgs &a, hipStream_t s);
template <int N> hipError_t launch_shock_n(int prec, int pass, const StageArgs &a, hipStream_t s);
template <int N> hipError_t upload_operators_n(const double *divT, const double *interp);
// Per-flux entry points of the multi-order kernels (one translation unit per flux)
template <int FLUX> hipError_t launch_stage_multi_f(int prec, const MultiArgs &m, hipStream_t s);
template <int FLUX> hipError_t upload_operators_multi_f(int N, const double *divT, const double *interp);

}  // namespace dev
}  // namespace gocfd
