// Host driver of the tangent-linear and discrete-adjoint step.
#include "adjoint_runner.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace gocfd {
namespace capi {

AdjointRunner::AdjointRunner(const Discretisation &d, const DeviceLayout &L, const dev::Phys &ph, bool io_tied,
                             bool lts, bool shock, double kappa, const AdjointGeometry &geo, hipStream_t stream)
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
  a_.lts = lts ? 1 : 0;
  a_.shock = shock ? 1 : 0;
  if (shock) {
    // the shock operator block: P, Mass, D, V, Vinv [Np^2] each, then mf [Np]
    const std::vector<double> ops = build_shock_operators(d.ops);
    const size_t np = size_t(NpInt), n2 = np * np;
    const double *V = ops.data() + 3 * n2, *Vi = V + n2, *mf = Vi + n2;
    std::vector<double> F(n2, 0.0);
    for (size_t i = 0; i < np; ++i)
      for (size_t k = 0; k < np; ++k) {
        double acc = 0.;
        for (size_t m = 0; m < np; ++m) acc += V[i * np + m] * (m == 0 ? 1.0 : mf[m]) * Vi[m * np + k];
        F[i * np + k] = acc;
      }
    filt_.upload(F);
    sens_.upload(std::vector<double>(ops.begin(), ops.begin() + 3 * n2));
    a_.filt = filt_.p;
    a_.sens = sens_.p;
    const double np1 = double(d.N + 1);
    a_.sens_left = 4. / (np1 * np1 * np1 * np1) - kappa;
    if (!lts) {
      // the whole sensor path, an active sensor included
      a_.act = 1;
      shops_.upload(ops);
      hK_.upload(L.hK);
      Nxo_.upload(L.Nxo);
      Nyo_.upload(L.Nyo);
      vooL_.upload(L.vooL);
      evert_.upload(L.evert);
      // vertex -> (element << 2 | local vertex) of the real elements, in
      // element order
      std::vector<int> vptr(size_t(d.Nv) + 1, 0), vinc(3 * size_t(L.K));
      for (int kk = 0; kk < L.K; ++kk)
        for (int c = 0; c < 3; ++c) ++vptr[size_t(L.evert[size_t(c) * L.Kp + kk]) + 1];
      for (int v = 0; v < d.Nv; ++v) vptr[v + 1] += vptr[v];
      std::vector<int> fill(vptr.begin(), vptr.end() - 1);
      for (int kk = 0; kk < L.K; ++kk)
        for (int c = 0; c < 3; ++c) vinc[size_t(fill[L.evert[size_t(c) * L.Kp + kk]]++)] = (kk << 2) | c;
      vptr_.upload(vptr);
      vinc_.upload(vinc);
      a_.shops = shops_.p;
      a_.hK = hK_.p;
      a_.Nxo = Nxo_.p;
      a_.Nyo = Nyo_.p;
      a_.vooL = vooL_.p;
      a_.evert = evert_.p;
      a_.vptr = vptr_.p;
      a_.vinc = vinc_.p;
      a_.Nv = d.Nv;
      a_.kappa = kappa;
      a_.S0 = 4. / (np1 * np1 * np1 * np1);
      a_.eps0 = 5. / 1.5;
      a_.cdiff = 1. / (np1 * np1);
      a_.omega = double(d.N * d.N);
      NR_ = NpInt + 3 * NpEdge;
    }
  }
  if (lts) {
    wsrc_.upload(L.wsrc);
    a_.wsrc = wsrc_.p;
  }
  n_ = size_t(L.Kp) * 4 * NpInt;
  nfp_ = size_t(3) * L.Kp * NpEdge;
  {
    std::vector<int> bnd;
    for (int f = 0; f < 3; ++f)
      for (int kk = 0; kk < L.K; ++kk) {
        const size_t fi = size_t(f) * L.Kp + kk;
        if ((L.face_code[fi] & 15) == FK_Shared) continue;
        for (int pp = 0; pp < NpEdge; ++pp) bnd.push_back(int(fi * NpEdge + pp));
      }
    bnd_.upload(bnd);
    a_.bnd = bnd_.p;
    a_.nbnd = int(bnd.size());
  }
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
  part_v_.alloc(4 * size_t(nparts_));
  part_i_.alloc(size_t(nparts_));
  if (a_.act) {
    const size_t Kp = size_t(L.Kp), Nv = size_t(d.Nv), nr = size_t(NR_) * Kp;
    sig_.alloc(Kp);
    eps_.alloc(Kp);
    sigv_.alloc(Nv);
    epsv_.alloc(Nv);
    dis_.alloc(8 * nr);
    SPV_.alloc(3 * Kp * NpEdge);
    sigbR_.alloc(5 * Kp);
    epsfR_.alloc(5 * nr);
    argsR_.alloc(5 * Nv);
    argeR_.alloc(5 * Nv);
    argmaxv_.alloc(5);
    vbr_.alloc(5);
    Wvv_.alloc(5);
    RpreR_.alloc(5 * n_);
    U2pre_.alloc(n_);
    part_v2_.alloc(size_t(nparts_));
    part_i2_.alloc(size_t(nparts_));
    lamSigb_.alloc(Kp);
    lamEpsf_.alloc(nr);
    lamDis_.alloc(8 * nr);
    lamEQv_.alloc(4 * nfp_);
    lamEQg_.alloc(4 * nfp_);
    lamSigv_.alloc(Nv);
    lamEpsv_.alloc(Nv);
  }
  if (lts) {
    scratch_small_.dtk.alloc(5 * size_t(L.Kp));
    scratch_small_.srck.alloc(5 * size_t(L.Kp));
    scratch_small_.wk.alloc(5 * size_t(L.Kp));
    carry_.alloc(L.Kp);
    lamSpeed_.alloc(nfp_);
  }
}

double *AdjointRunner::state(size_t i) {
  while (states_.size() <= i) {
    states_.push_back(std::make_unique<DevBuf<double>>());
    states_.back()->alloc(n_);
  }
  return states_[i]->p;
}

AdjointRunner::SmallRecord &AdjointRunner::small(size_t i) {
  while (smalls_.size() <= i) {
    smalls_.push_back(std::make_unique<SmallRecord>());
    SmallRecord &sm = *smalls_.back();
    sm.dts.alloc(5);
    sm.argmax.alloc(5);
    sm.Wv.alloc(5);
    if (a_.lts) {
      sm.dtk.alloc(5 * size_t(a_.Kp));
      sm.srck.alloc(5 * size_t(a_.Kp));
      sm.wk.alloc(5 * size_t(a_.Kp));
    }
  }
  return *smalls_[i];
}

void AdjointRunner::fill_small(StageRecord &r, SmallRecord &sm) {
  r.dts = sm.dts.p;
  r.argmax = sm.argmax.p;
  r.Wv = sm.Wv.p;
  for (int s = 0; s < 5; ++s) {
    const size_t o = size_t(s) * a_.Kp;
    r.dtk[s] = sm.dtk.p ? static_cast<void *>(sm.dtk.p + o) : nullptr;
    r.srck[s] = sm.srck.p ? sm.srck.p + o : nullptr;
    r.wk[s] = sm.wk.p ? sm.wk.p + o : nullptr;
  }
}

void AdjointRunner::launch(int op, dev::AdjBufs &b) {
  b.nparts = nparts_;
  HIPCHECK(dev::adj_launch(op, a_, b, s_));
}

AdjointRunner::StageRecord AdjointRunner::scratch_record() {
  StageRecord r{};
  for (int s = 0; s < 5; ++s) {
    r.U[s] = U_[s].p;
    r.R[s] = R_[s].p;
  }
  fill_small(r, scratch_small_);
  r.dts = dts_.p;
  r.argmax = argmax_.p;
  r.Wv = Wv_.p;
  if (a_.act) {
    const size_t Kp = size_t(a_.Kp), Nv = size_t(a_.Nv), nr = size_t(NR_) * Kp;
    for (int s = 0; s < 5; ++s) {
      r.sigb[s] = sigbR_.p + s * Kp;
      r.epsf[s] = epsfR_.p + s * nr;
      r.args[s] = argsR_.p + s * Nv;
      r.arge[s] = argeR_.p + s * Nv;
      r.Rpre[s] = RpreR_.p + s * n_;
    }
    r.U2pre = U2pre_.p;
    r.argmaxv = argmaxv_.p;
    r.vbr = vbr_.p;
    r.Wvv = Wvv_.p;
  }
  return r;
}

// One SSP54 step from Ustart into Uend, leaving the stage record in `rec`.
// With `time`, the FinalTime check and the time advance.
void AdjointRunner::forward_step(bool dual, const void *Ustart, void *Uend, double *time, const StageRecord &rec) {
  dev::AdjBufs b{};
  b.dual = dual ? 1 : 0;
  b.dts = rec.dts;
  b.argmax = rec.argmax;
  b.Wv = rec.Wv;
  b.time = time;
  b.flag = flag_.p;
  b.part_v = part_v_.p;
  b.part_i = part_i_.p;
  b.n = n_;
  for (int s = 0; s < 5; ++s) {
    const void *Us = s == 0 ? Ustart : static_cast<const void *>(rec.U[s]);
    b.stage = s;
    b.Us = Us;
    if (a_.act) {
      // the sensor fields of the stage input, then the stage-2 state filter
      // in place (its input kept for the adjoint)
      b.sig = sig_.p;
      b.eps = eps_.p;
      b.sigv = sigv_.p;
      b.epsv = epsv_.p;
      b.args = rec.args[s];
      b.arge = rec.arge[s];
      b.sigb = rec.sigb[s];
      b.epsf = rec.epsf[s];
      b.SPV = SPV_.p;
      b.dis = dis_.p;
      launch(dev::ADJ_FWD_SHOCK, b);
      if (s == 2) {
        if (!dual) HIPCHECK(hipMemcpyAsync(rec.U2pre, rec.U[2], n_ * sizeof(double), hipMemcpyDeviceToDevice, s_));
        b.out = rec.U[2];
        launch(dev::ADJ_FILTER_SIG, b);
      }
    } else if (a_.shock) {
      // the sensor on the stage input (checked in the sweeps that check the
      // time), then the stage-2 state filter in place
      if (time) launch(dev::ADJ_SENSOR, b);
      if (s == 2) {
        b.out = rec.U[2];
        launch(dev::ADJ_FILTER, b);
      }
    }
    b.FE = FE_.p;
    b.R = rec.R[s];
    b.SP = SP_.p;
    b.U0 = Ustart;
    b.U2 = rec.U[2];
    b.U3 = rec.U[3];
    b.R3 = rec.R[3];
    b.out = s < 4 ? rec.U[s + 1] : Uend;
    launch(dev::ADJ_FWD_FACE, b);
    if (a_.act) launch(dev::ADJ_FWD_VISC, b);
    b.argmaxv = rec.argmaxv;
    b.vbr = rec.vbr;
    b.Wvv = rec.Wvv;
    b.part_v2 = part_v2_.p;
    b.part_i2 = part_i2_.p;
    b.Rpre = (a_.act && !dual) ? rec.Rpre[s] : nullptr;
    if (a_.lts) {
      b.dtk = rec.dtk[s];
      b.dtk_prev = s > 0 ? rec.dtk[s - 1] : nullptr;
      b.srck = rec.srck[s];
      b.wk = rec.wk[s];
      launch(dev::ADJ_FWD_LTS, b);  // no time advance or FinalTime clamp, as in the solver
    } else {
      launch(dev::ADJ_ARGMAX, b);
    }
    launch(dev::ADJ_FWD_RHS, b);
  }
}

// The transposed stages of one step, last to first, from its stage record:
// the cotangent of the step's result in b.lamU0 becomes that of its start
void AdjointRunner::adjoint_step(const double *Ustart, const StageRecord &rec, dev::AdjBufs &b) {
  b.dts = rec.dts;
  b.argmax = rec.argmax;
  b.Wv = rec.Wv;
  for (int s = 4; s >= 0; --s) {
    b.stage = s;
    b.lam = s == 4 ? b.lamU0 : b.lamQ[s + 1];
    b.Rs = rec.R[s];
    b.R3 = rec.R[3];
    b.dtk = rec.dtk[s];
    b.srck = rec.srck[s];
    b.wk = rec.wk[s];
    launch(a_.lts ? dev::ADJ_ADJ_SSP_LTS : dev::ADJ_ADJ_SSP, b);
    if (a_.act) {  // the RHS filter, transposed; its sigma term sets lamSigb
      dev::AdjBufs bf = b;
      bf.lam = b.lamR;
      bf.Xpre = rec.Rpre[s];
      bf.sigb = rec.sigb[s];
      bf.accumulate = 0;
      launch(dev::ADJ_ADJ_FILTER_SIG_T, bf);
    } else if (a_.shock) {  // the RHS filter, transposed
      dev::AdjBufs bf = b;
      bf.lam = b.lamR;
      launch(dev::ADJ_FILTER_T, bf);
    }
    b.Us = s == 0 ? static_cast<const void *>(Ustart) : static_cast<const void *>(rec.U[s]);
    b.target = s == 0 ? b.lamU0 : b.lamQ[s];
    launch(dev::ADJ_ADJ_FACE, b);
    if (a_.act) {  // the viscous edge flux and eps grad Q, transposed
      b.epsf = rec.epsf[s];
      launch(dev::ADJ_ADJ_VISC, b);
    }
    if (a_.lts) launch(dev::ADJ_ADJ_LTS_WAVE, b);
    launch(dev::ADJ_ADJ_ELEM, b);
    if (!a_.lts) {
      if (a_.act) {  // the viscous dt where it bound, into eps
        b.vbr = rec.vbr;
        b.argmaxv = rec.argmaxv;
        b.Wvv = rec.Wvv;
        launch(dev::ADJ_ADJ_VDT, b);
      }
      launch(dev::ADJ_ADJ_WAVE, b);
    }
    if (a_.shock && s == 2) {  // the stage-2 state filter, transposed
      dev::AdjBufs bf = b;
      bf.lam = b.lamQ[2];
      if (a_.act) {
        bf.Xpre = rec.U2pre;
        bf.sigb = rec.sigb[2];
        bf.accumulate = 1;
        launch(dev::ADJ_ADJ_FILTER_SIG_T, bf);
      } else {
        launch(dev::ADJ_FILTER_T, bf);
      }
    }
    if (a_.act) {  // the vertex merges, the ramp and the sensor, at the sensor's input
      b.Usens = s == 2 ? rec.U2pre : static_cast<const double *>(b.Us);
      b.args = rec.args[s];
      b.arge = rec.arge[s];
      launch(dev::ADJ_ADJ_SENSOR, b);
    }
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
  if (f & 4)
    throw SetupError(GOCFD_ERR_ARG,
                     "the PerssonC0 shock sensor is active in the window (an element's Se reaches S0 - Kappa); "
                     "the derivative through the sensor and the artificial dissipation is not formed");
}

void AdjointRunner::vjp(const double *q0, double t0, int nsteps, int every, const double *lamN, double *lam0,
                        double far[4]) {
  if (nsteps < 1) throw SetupError(GOCFD_ERR_ARG, "nsteps must be positive");
  every_ = every > 0 ? every : int(std::ceil(std::sqrt(double(nsteps))));
  every_ = std::min(every_, nsteps);
  const int nblocks = (nsteps + every_ - 1) / every_;
  const size_t bytes = n_ * sizeof(double);
  // pooled states: checkpoints [0, nblocks), block starts, the result
  std::vector<double *> ck(nblocks), S(every_, nullptr);
  for (int i = 0; i < nblocks; ++i) ck[i] = state(size_t(i));
  for (int i = 1; i < every_; ++i) S[i] = state(size_t(nblocks + i));
  double *const end = state(size_t(nblocks + every_));
  const size_t rec0 = size_t(nblocks + every_ + 1);  // first stage-record state
  HIPCHECK(hipMemsetAsync(flag_.p, 0, sizeof(int), s_));
  HIPCHECK(hipMemcpyAsync(time_.p, &t0, sizeof(double), hipMemcpyHostToDevice, s_));
  HIPCHECK(hipMemcpyAsync(ck[0], q0, bytes, hipMemcpyDeviceToDevice, s_));
  // The stage records of a whole block are kept when the states they still
  // need fit in a quarter of the free device memory; otherwise each step's
  // record is recomputed from its start just before its transposed stages.
  size_t free_b = 0, total_b = 0;
  HIPCHECK(hipMemGetInfo(&free_b, &total_b));
  const size_t want = rec0 + 9 * size_t(every_), have = states_.size();
  stored_records_ = (want > have ? (want - have) * bytes : 0) <= free_b / 4;
  // GOCFD_ADJOINT_KEEP_RECORDS=0 forces the per-step recomputation
  if (const char *e = std::getenv("GOCFD_ADJOINT_KEEP_RECORDS")) stored_records_ = stored_records_ && e[0] != '0';
  // the active sensor path keeps one step's record (its fields per stage)
  if (a_.act) stored_records_ = false;
  // With one block and its records kept, the record sweep is the forward
  // sweep (with the FinalTime check); otherwise a forward sweep first stores
  // the checkpoints at the block starts.
  const bool one_sweep = nblocks == 1 && stored_records_;
  const StageRecord scratch = scratch_record();
  if (!one_sweep) {
    for (int bk = 0; bk < nblocks; ++bk) {
      const int j0 = bk * every_, cnt = std::min(every_, nsteps - j0);
      const double *src = ck[bk];
      for (int i = 0; i < cnt; ++i) {
        double *dst = i + 1 < cnt ? S[i + 1] : (bk + 1 < nblocks ? ck[bk + 1] : end);
        forward_step(false, src, dst, time_.p, scratch);
        src = dst;
      }
    }
    check_flag();
  }
  std::vector<StageRecord> recs(stored_records_ ? size_t(every_) : 0);
  for (size_t i = 0; i < recs.size(); ++i) {
    StageRecord &r = recs[i];
    r.U[0] = nullptr;
    for (int s = 1; s < 5; ++s) r.U[s] = state(rec0 + 9 * i + size_t(s - 1));
    for (int s = 0; s < 5; ++s) r.R[s] = state(rec0 + 9 * i + 4 + size_t(s));
    fill_small(r, small(i));
  }
  // reverse sweep
  if (!lamU0_.p) {
    lamU0_.alloc(n_);
    for (int s = 1; s < 5; ++s) lamQ_[s].alloc(n_);
    lamR_.alloc(n_);
    lamR3_.alloc(n_);
    lamOwn_.alloc(4 * nfp_);
    lamFarPts_.alloc(4 * nfp_);
    lamFarW_.alloc(4);
    lam_dt_.alloc(5);
    far_d_.alloc(4);
  }
  DevBuf<double> &lamU0 = lamU0_, *lamQ = lamQ_, &lamR = lamR_, &lamR3 = lamR3_, &lamOwn = lamOwn_,
                 &lamFarPts = lamFarPts_, &lamFarW = lamFarW_, &lam_dt = lam_dt_;
  HIPCHECK(hipMemcpyAsync(lamU0.p, lamN, bytes, hipMemcpyDeviceToDevice, s_));
  HIPCHECK(hipMemsetAsync(lamFarPts.p, 0, 4 * nfp_ * sizeof(double), s_));
  HIPCHECK(hipMemsetAsync(lamFarW.p, 0, 4 * sizeof(double), s_));
  if (a_.lts) {
    HIPCHECK(hipMemsetAsync(carry_.p, 0, carry_.n * sizeof(double), s_));
    HIPCHECK(hipMemsetAsync(lamSpeed_.p, 0, lamSpeed_.n * sizeof(double), s_));
  }
  dev::AdjBufs b{};
  b.dual = 0;
  b.part_v = part_v_.p;
  b.part_i = part_i_.p;
  b.n = n_;
  b.lamU0 = lamU0.p;
  for (int s = 1; s < 5; ++s) b.lamQ[s] = lamQ[s].p;
  b.lamR = lamR.p;
  b.lamR3 = lamR3.p;
  b.lamOwn = lamOwn.p;
  b.lamFarPts = lamFarPts.p;
  b.lamFarW = lamFarW.p;
  b.lam_dt = lam_dt.p;
  b.carry = carry_.p;
  b.lamSpeed = lamSpeed_.p;
  b.lamSigb = lamSigb_.p;
  b.lamEpsf = lamEpsf_.p;
  b.lamDis = lamDis_.p;
  b.lamEQv = lamEQv_.p;
  b.lamEQg = lamEQg_.p;
  b.lamSigv = lamSigv_.p;
  b.lamEpsv = lamEpsv_.p;
  for (int bk = nblocks - 1; bk >= 0; --bk) {
    const int j0 = bk * every_, cnt = std::min(every_, nsteps - j0);
    std::vector<const double *> starts(cnt);
    starts[0] = ck[bk];
    for (int i = 0; i < cnt; ++i) {
      void *next = i + 1 < cnt ? static_cast<void *>(S[i + 1]) : static_cast<void *>(tmp_.p);
      if (stored_records_) forward_step(false, starts[i], next, one_sweep ? time_.p : nullptr, recs[i]);
      else if (i + 1 < cnt) forward_step(false, starts[i], next, nullptr, scratch);
      if (i + 1 < cnt) starts[i + 1] = S[i + 1];
    }
    if (one_sweep) check_flag();
    for (int i = cnt - 1; i >= 0; --i) {
      if (stored_records_) {
        adjoint_step(starts[i], recs[i], b);
      } else {
        forward_step(false, starts[i], tmp_.p, nullptr, scratch);
        adjoint_step(starts[i], scratch, b);
      }
    }
  }
  b.target = far_d_.p;
  launch(dev::ADJ_FAR_SUM, b);
  HIPCHECK(hipMemcpyAsync(lam0, lamU0.p, bytes, hipMemcpyDeviceToDevice, s_));
  HIPCHECK(hipMemcpyAsync(far, far_d_.p, 4 * sizeof(double), hipMemcpyDeviceToHost, s_));
  HIPCHECK(hipStreamSynchronize(s_));
}

void AdjointRunner::jvp(const double *q0, const double *dq0, const double dfar[4], double t0, int nsteps,
                        double *qn, double *dqn) {
  if (nsteps < 1) throw SetupError(GOCFD_ERR_ARG, "nsteps must be positive");
  for (int c = 0; c < 4; ++c) a_.dfar[c] = dfar ? dfar[c] : 0.;
  if (!dualA_.p) {
    dualA_.alloc(n_);
    dualB_.alloc(n_);
  }
  dev::ad_storage *A = dualA_.p, *B = dualB_.p;
  HIPCHECK(hipMemsetAsync(flag_.p, 0, sizeof(int), s_));
  HIPCHECK(hipMemcpyAsync(time_.p, &t0, sizeof(double), hipMemcpyHostToDevice, s_));
  dev::AdjBufs b{};
  b.n = n_;
  b.pack_v = q0;
  b.pack_d = dq0;
  b.out = A;
  launch(dev::ADJ_PACK, b);
  const StageRecord scratch = scratch_record();
  for (int j = 0; j < nsteps; ++j) {
    forward_step(true, A, B, time_.p, scratch);
    std::swap(A, B);
  }
  b.Us = A;
  b.unpack_v = qn;
  b.unpack_d = dqn;
  launch(dev::ADJ_UNPACK, b);
  for (int c = 0; c < 4; ++c) a_.dfar[c] = 0.;
  check_flag();
}

}  // namespace capi
}  // namespace gocfd
