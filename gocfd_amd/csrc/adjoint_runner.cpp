// Host driver of the tangent-linear and discrete-adjoint step.
#include "adjoint_runner.hpp"

#include <algorithm>
#include <cmath>

namespace gocfd {
namespace capi {

AdjointRunner::AdjointRunner(const Discretisation &d, const DeviceLayout &L, const dev::Phys &ph, bool io_tied,
                             const AdjointGeometry &geo, hipStream_t stream)
    : s_(stream) {
  if (L.nbfaces > 0)
    throw SetupError(GOCFD_ERR_ARG,
                     "the tangent-linear and adjoint steps do not handle IVortex boundary faces (their analytic "
                     "state depends on the time)");
  const int NpInt = d.ops.NpInt, NpEdge = d.ops.NpEdge;
  a_.N = d.N;
  a_.K = L.K;
  a_.Kp = L.Kp;
  a_.Jdet = geo.Jdet; a_.Jinv = geo.Jinv; a_.IInII = geo.IInII; a_.Nx = geo.Nx; a_.Ny = geo.Ny;
  a_.oohK = geo.oohK; a_.code = geo.code; a_.Interp = geo.Interp;
  nbr_elem_.upload(L.nbr_elem);
  nbr_face_.upload(L.nbr_face);
  divint_.upload(d.ops.DivInt);
  a_.nbr_elem = nbr_elem_.p;
  a_.nbr_face = nbr_face_.p;
  a_.DivInt = divint_.p;
  a_.ph = ph;
  a_.io_tied = io_tied ? 1 : 0;
  n_ = size_t(L.Kp) * 4 * NpInt;
  nfp_ = size_t(3) * L.Kp * NpEdge;
  for (int s = 0; s < 5; ++s) {
    U_[s].alloc(n_);
    R_[s].alloc(n_);
  }
  FE_.alloc(4 * nfp_);
  SP_.alloc(nfp_);
  tmp_.alloc(n_);
  dts_.alloc(5);
  argmax_.alloc(5);
  flag_.alloc(1);
  Wv_.alloc(5);
  time_.alloc(1);
  part_v_.alloc(std::max<size_t>(size_t(nparts_), (n_ + 255) / 256));
  part_i_.alloc(size_t(nparts_));
}

void AdjointRunner::launch(int op, dev::AdjBufs &b) {
  b.nparts = nparts_;
  HIPCHECK(dev::adj_launch(op, a_, b, s_));
}

// One SSP54 step from Ustart into Uend, leaving the stage inputs (U_[1..4]),
// right-hand sides (R_[0..4]), dts, arg-max points and wave maxima in the
// stage record. With `time`, the FinalTime check and the time advance.
void AdjointRunner::forward_step(bool dual, const void *Ustart, void *Uend, double *time) {
  dev::AdjBufs b{};
  b.dual = dual ? 1 : 0;
  b.dts = dts_.p;
  b.argmax = argmax_.p;
  b.Wv = Wv_.p;
  b.time = time;
  b.flag = flag_.p;
  b.part_v = part_v_.p;
  b.part_i = part_i_.p;
  b.n = n_;
  for (int s = 0; s < 5; ++s) {
    const void *Us = s == 0 ? Ustart : static_cast<const void *>(U_[s].p);
    b.stage = s;
    b.Us = Us;
    b.FE = FE_.p;
    b.R = R_[s].p;
    b.SP = SP_.p;
    launch(dev::ADJ_FWD_FACE, b);
    launch(dev::ADJ_FWD_RHS, b);
    launch(dev::ADJ_FWD_WAVE, b);
    launch(dev::ADJ_ARGMAX, b);
    b.U0 = Ustart;
    b.Rs = R_[s].p;
    b.U2 = U_[2].p;
    b.U3 = U_[3].p;
    b.R3 = R_[3].p;
    b.out = s < 4 ? static_cast<void *>(U_[s + 1].p) : Uend;
    launch(dev::ADJ_FWD_SSP, b);
  }
}

void AdjointRunner::check_flag() {
  int f = 0;
  HIPCHECK(hipMemcpyAsync(&f, flag_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
  if (f & 1)
    throw SetupError(GOCFD_ERR_ARG,
                     "the steps reach FinalTime (a stage dt would be clamped); the derivative of the clamp is not "
                     "formed");
  if (f & 2) throw SetupError(GOCFD_ERR_ARG, "the mesh has no face that bounds the time step");
}

void AdjointRunner::vjp(const double *q0, double t0, int nsteps, int every, const double *lamN, double *lam0,
                        double far[4]) {
  if (nsteps < 1) throw SetupError(GOCFD_ERR_ARG, "nsteps must be positive");
  every_ = every > 0 ? every : int(std::ceil(std::sqrt(double(nsteps))));
  every_ = std::min(every_, nsteps);
  const int nblocks = (nsteps + every_ - 1) / every_;
  const size_t bytes = n_ * sizeof(double);
  std::vector<DevBuf<double>> ck(nblocks), S(every_);
  for (auto &c : ck) c.alloc(n_);
  for (int i = 1; i < every_; ++i) S[i].alloc(n_);
  DevBuf<double> end;
  end.alloc(n_);
  HIPCHECK(hipMemsetAsync(flag_.p, 0, sizeof(int), s_));
  HIPCHECK(hipMemcpyAsync(time_.p, &t0, sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHECK(hipMemcpyAsync(ck[0].p, q0, bytes, hipMemcpyDeviceToDevice, s_));
  // forward sweep: checkpoints at the block starts
  for (int bk = 0; bk < nblocks; ++bk) {
    const int j0 = bk * every_, cnt = std::min(every_, nsteps - j0);
    const double *src = ck[bk].p;
    for (int i = 0; i < cnt; ++i) {
      double *dst = i + 1 < cnt ? S[i + 1].p : (bk + 1 < nblocks ? ck[bk + 1].p : end.p);
      forward_step(false, src, dst, time_.p);
      src = dst;
    }
  }
  check_flag();
  // reverse sweep
  DevBuf<double> lamU0, lamQ[5], lamR, lamR3, lamOwn, lamOth, lamFarPts, lamFarW, lam_dt;
  lamU0.alloc(n_);
  for (int s = 1; s < 5; ++s) lamQ[s].alloc(n_);
  lamR.alloc(n_);
  lamR3.alloc(n_);
  lamOwn.alloc(4 * nfp_);
  lamOth.alloc(4 * nfp_);
  lamFarPts.alloc(4 * nfp_);
  lamFarW.alloc(4);
  lam_dt.alloc(5);
  HIPCHECK(hipMemcpyAsync(lamU0.p, lamN, bytes, hipMemcpyDeviceToDevice, s_));
  HIPCHECK(hipMemsetAsync(lamFarPts.p, 0, 4 * nfp_ * sizeof(double), s_));
  HIPCHECK(hipMemsetAsync(lamFarW.p, 0, 4 * sizeof(double), s_));
  dev::AdjBufs b{};
  b.dual = 0;
  b.dts = dts_.p;
  b.argmax = argmax_.p;
  b.Wv = Wv_.p;
  b.part_v = part_v_.p;
  b.part_i = part_i_.p;
  b.n = n_;
  b.lamU0 = lamU0.p;
  for (int s = 1; s < 5; ++s) b.lamQ[s] = lamQ[s].p;
  b.lamR = lamR.p;
  b.lamR3 = lamR3.p;
  b.lamOwn = lamOwn.p;
  b.lamOth = lamOth.p;
  b.lamFarPts = lamFarPts.p;
  b.lamFarW = lamFarW.p;
  b.lam_dt = lam_dt.p;
  for (int bk = nblocks - 1; bk >= 0; --bk) {
    const int j0 = bk * every_, cnt = std::min(every_, nsteps - j0);
    std::vector<const double *> starts(cnt);
    starts[0] = ck[bk].p;
    for (int i = 1; i < cnt; ++i) {
      forward_step(false, starts[i - 1], S[i].p, nullptr);
      starts[i] = S[i].p;
    }
    for (int i = cnt - 1; i >= 0; --i) {
      forward_step(false, starts[i], tmp_.p, nullptr);  // the stage record of this step
      for (int s = 4; s >= 0; --s) {
        b.stage = s;
        b.lam = s == 4 ? lamU0.p : lamQ[s + 1].p;
        b.Rs = R_[s].p;
        b.R3 = R_[3].p;
        launch(dev::ADJ_ADJ_SSP, b);
        b.Us = s == 0 ? static_cast<const void *>(starts[i]) : static_cast<const void *>(U_[s].p);
        b.target = s == 0 ? lamU0.p : lamQ[s].p;
        launch(dev::ADJ_ADJ_FACE, b);
        launch(dev::ADJ_ADJ_ELEM, b);
        launch(dev::ADJ_ADJ_WAVE, b);
      }
    }
  }
  launch(dev::ADJ_FAR_SUM, b);
  HIPCHECK(hipMemcpyAsync(lam0, lamU0.p, bytes, hipMemcpyDeviceToDevice, s_));
  HIPCHECK(hipMemcpyAsync(far, part_v_.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
}

void AdjointRunner::jvp(const double *q0, const double *dq0, const double dfar[4], double t0, int nsteps,
                        double *qn, double *dqn) {
  if (nsteps < 1) throw SetupError(GOCFD_ERR_ARG, "nsteps must be positive");
  for (int c = 0; c < 4; ++c) a_.dfar[c] = dfar ? dfar[c] : 0.;
  DevBuf<dev::ad_storage> A, B;
  A.alloc(n_);
  B.alloc(n_);
  HIPCHECK(hipMemsetAsync(flag_.p, 0, sizeof(int), s_));
  HIPCHECK(hipMemcpyAsync(time_.p, &t0, sizeof(double), hipMemcpyHostToDevice, s_));
  dev::AdjBufs b{};
  b.n = n_;
  b.pack_v = q0;
  b.pack_d = dq0;
  b.out = A.p;
  launch(dev::ADJ_PACK, b);
  for (int j = 0; j < nsteps; ++j) {
    forward_step(true, A.p, B.p, time_.p);
    std::swap(A.p, B.p);
  }
  b.Us = A.p;
  b.unpack_v = qn;
  b.unpack_d = dqn;
  launch(dev::ADJ_UNPACK, b);
  for (int c = 0; c < 4; ++c) a_.dfar[c] = 0.;
  check_flag();
}

}  // namespace capi
}  // namespace gocfd
