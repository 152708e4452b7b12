// Device-side interface of the 2D Euler DFR stage kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gocfd {
namespace dev {

constexpr int kTile = 64;      // elements per workgroup tile (one wavefront wide)
constexpr int kThreads = 256;  // threads per workgroup

// Device control block. Time is double-buffered by step parity and the wave
// speed maximum lives in a ring of three slots indexed by the global stage
// counter, so that no kernel writes a value another block of the same launch
// still reads.
struct Ctl {
  double time[2];
  double dt[5];
  unsigned long long wmax[3];
  int nan;
  int steps;
  int pad[2];
};

struct Geom {
  const double *Jdet, *Jinv, *IInII, *Nx, *Ny, *oohK;  // [Kp], [4][Kp], [3][Kp] x3, [Kp]
  const int *code, *nbr_slot;                          // [3][Kp]
  const int *halo_elem, *halo_face;                    // [ntiles][Hmax]
  const int *wave_list, *wave_cnt;                     // [ntiles][Wmax], [ntiles]
  const double *bxy;                                   // IVortex boundary points
};

struct Phys {
  double gamma, minf, pinf, cinf;
  double qinf[4];
  double cfl, final_time;
  int flux;
};

struct StageArgs {
  int K, Kp, ntiles, Hmax, Wmax;
  int stage;           // 0..4, or -1 for the RHS-only evaluation
  int step_index;      // steps since the state was set (selects time parity)
  int ring;            // global stage counter mod 3
  double t_rhs;        // time for the RHS-only evaluation
  const void *qs;      // stage input state (element-major)
  void *q0, *q1, *q2, *q3, *q4, *R;
  double *rhs_out;     // RHS-only output (element-major, double)
  Ctl *ctl;
  Geom g;
  const double *Interp64;       // FluxEdgeInterp in FP64 (wave-speed prologue)
  Phys ph;
};

// Precision variants: 0 = FP64 storage and arithmetic; 1 = FP32 storage and
// bulk arithmetic with FP64 geometry, dt reduction and SSP combination.
hipError_t launch_stage(int N, int prec, const StageArgs &a, hipStream_t s);
hipError_t launch_prologue(int N, int prec, const StageArgs &a, hipStream_t s);
hipError_t launch_ctl_init(Ctl *ctl, double time, hipStream_t s);
// Copy the order-N operators (DivInt^T [NpFlux][NpInt], FluxEdgeInterp
// [3NpEdge][NpInt]) into the current device's constant memory.
hipError_t upload_operators(int N, const double *divT, const double *interp);
size_t stage_lds_bytes(int N, int prec, int Hmax);

}  // namespace dev
}  // namespace gocfd
