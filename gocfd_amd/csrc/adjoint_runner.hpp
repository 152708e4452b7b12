// Host driver of the tangent-linear and discrete-adjoint step (kernels in
// adjoint_kernels.hpp / adjoint_common.hip, scope in adjoint.hpp).
//
// vjp: one forward sweep stores the state every `every` steps; the reverse
// sweep recomputes each block of steps from its checkpoint, then each step's
// five stage inputs, right-hand sides, dts and arg-max points, and applies
// the transposed stages from the last to the first.
// jvp: the forward sweep in dual arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "adjoint.hpp"
#include "capi_common.hpp"
#include "dfr_setup.hpp"

namespace gocfd {
namespace capi {

struct AdjointGeometry {
  const double *Jdet, *Jinv, *IInII, *Nx, *Ny, *oohK, *Interp;
  const int *code;
};

class AdjointRunner {
 public:
  // shock: PerssonC0 with this Kappa. With global time stepping the whole
  // sensor path is differentiated (an active sensor included); with local
  // time stepping only while the sensor is quiet (sigma = 0 everywhere; a
  // window where it is not is refused)
  AdjointRunner(const Discretisation &d, const DeviceLayout &L, const dev::Phys &ph, bool io_tied, bool lts,
                bool shock, double kappa, const AdjointGeometry &geo, hipStream_t stream);
  // lam0 = (dq_n/dq_0)^T lamN and far = (dq_n/dqinf)^T lamN for nsteps steps
  // from q0 (device, element-major FP64) at time t0; every = 0 picks
  // ceil(sqrt(nsteps)) steps per checkpoint
  void vjp(const double *q0, double t0, int nsteps, int every, const double *lamN, double *lam0, double far[4]);
  // qn and dqn = (dq_n/dq_0) dq0 + (dq_n/dqinf) dfar
  void jvp(const double *q0, const double *dq0, const double dfar[4], double t0, int nsteps, double *qn,
           double *dqn);
  int last_every() const { return every_; }
  bool last_stored_records() const { return stored_records_; }

 private:
  // Where a step leaves its stage inputs U[1..4], right-hand sides R[0..4],
  // dts, arg-max points and wave maxima
  struct StageRecord {
    void *U[5];
    void *R[5];
    void *dts;
    int *argmax;
    double *Wv;
    // local time stepping: per stage, the element dts, the points that set
    // them and the maxima
    void *dtk[5];
    int *srck[5];
    double *wk[5];
    // active PerssonC0, per stage: vertex-mean sigma, eps at the used RT
    // points, the vertex maxima's elements, the RHS before its filter; the
    // stage-2 state before its filter; the viscous dt's point and branch
    void *sigb[5], *epsf[5];
    int *args[5], *arge[5];
    double *Rpre[5];
    double *U2pre;
    int *argmaxv, *vbr;
    double *Wvv;
  };
  void launch(int op, dev::AdjBufs &b);
  void forward_step(bool dual, const void *Ustart, void *Uend, double *time, const StageRecord &rec);
  void adjoint_step(const double *Ustart, const StageRecord &rec, dev::AdjBufs &b);
  StageRecord scratch_record();
  void check_flag();
  // Buffers kept from call to call (grown on demand, freed with the solver):
  // element-array states (checkpoints, block starts, stage records) and the
  // small per-step record arrays
  double *state(size_t i);
  struct SmallRecord {
    DevBuf<dev::ad_storage> dts, dtk;  // dtk, srck, wk: [5][Kp] with local time stepping
    DevBuf<int> argmax, srck;
    DevBuf<double> Wv, wk;
  };
  void fill_small(StageRecord &r, SmallRecord &sm);
  SmallRecord &small(size_t i);
  std::vector<std::unique_ptr<DevBuf<double>>> states_;
  std::vector<std::unique_ptr<SmallRecord>> smalls_;
  DevBuf<double> lamU0_, lamQ_[5], lamR_, lamR3_, lamOwn_, lamFarPts_, lamFarW_, lam_dt_, far_d_;
  DevBuf<dev::ad_storage> dualA_, dualB_;
  SmallRecord scratch_small_;
  DevBuf<int> wsrc_, bnd_;
  DevBuf<double> filt_, sens_;
  DevBuf<double> carry_, lamSpeed_;
  dev::AdjArgs a_{};
  hipStream_t s_;
  size_t n_ = 0, nfp_ = 0;  // element-array and face-point counts
  int every_ = 0;
  bool stored_records_ = false;
  DevBuf<double> divint_;
  DevBuf<int> nbr_elem_, nbr_face_;
  // stage record (sized for dual values; the double sweeps use the front half)
  DevBuf<dev::ad_storage> U_[5], R_[5];
  DevBuf<dev::ad_storage> FE_, SP_, tmp_;
  DevBuf<dev::ad_storage> dts_;
  DevBuf<int> argmax_, flag_, part_i_;
  DevBuf<double> Wv_, time_, part_v_;
  int nparts_ = 1024;
  // active PerssonC0 (global time stepping)
  DevBuf<double> shops_, hK_, Nxo_, Nyo_, vooL_;
  DevBuf<int> evert_, vptr_, vinc_;
  DevBuf<dev::ad_storage> sig_, eps_, sigv_, epsv_, dis_, SPV_, sigbR_, epsfR_;
  DevBuf<int> argsR_, argeR_, argmaxv_, vbr_, part_i2_;
  DevBuf<double> RpreR_, U2pre_, Wvv_, part_v2_;
  DevBuf<double> lamSigb_, lamEpsf_, lamDis_, lamEQv_, lamEQg_, lamSigv_, lamEpsv_;
  int NR_ = 0;
};

}  // namespace capi
}  // namespace gocfd
