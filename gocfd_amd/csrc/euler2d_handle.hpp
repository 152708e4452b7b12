// The single-solver handle of the C ABI (include/gocfd_euler2d.h): the
// device-resident state, geometry and per-stage orchestration, shared by the
// translation units that implement its entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "adjoint_runner.hpp"
#include "capi_common.hpp"
#include "dfr_setup.hpp"
#include "gocfd_euler2d.h"
#include "kernels.hpp"

// included only by the translation units of the handle's entry points
using namespace gocfd;
using namespace gocfd::capi;

struct gocfd_euler2d {
  gocfd_params prm{};
  Discretisation d;
  DeviceLayout L;
  CaseStreams fs;
  int device = 0;
  int prec = 0;
  size_t ssize = 8;  // bytes per stored stage-state value (q1..q4); q0 and R are FP64
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // state (element-major, Kp * NV values of the storage type)
  void *q[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  void *R = nullptr;
  void *scratch_in = nullptr;
  DevBuf<double> rhs_out;
  DevBuf<double> Jdet, Jinv, IInII, Nx, Ny, oohK, bxy, Interp64;
  DevBuf<int> code, nbr_slot, halo_elem, halo_face, wave_list, wave_cnt, wave_code, flux_list, flux_cnt, bnd_cnt;
  DevBuf<int> perm;  // [Kp] device element -> reference element (-1: padding)
  StageTypeBuf flux_geo;
  // halo traces (without shock capturing): export lists and the two buffers
  DevBuf<int> xp_list, xp_dst, xp_cnt;
  DevBuf<double> ht;
  // local time stepping
  bool lts = false;
  DevBuf<double> fw, fwv, DTk, DTv;
  DevBuf<int> wsrc;
  // shock capturing (PerssonC0)
  bool shock = false;
  double kappa = 0, eps0 = 0;
  DevBuf<double> sh_ops, hK, Se, sigma, eps, epsF, diss, Nxo, Nyo, vooL;
  DevBuf<float> sh_opsF;
  DevBuf<int> evert, nbr_elem, nbr_face, active, act_list;
  DevBuf<unsigned> act_cnt;
  DevBuf<unsigned long long> sigV, epsV;
  // a single solver is a batch of one case
  DevBuf<dev::CaseInfo> cases;
  DevBuf<int> tile_case;
  dev::Ctl *ctl = nullptr;
  long long gstage = 0;  // stage launches since the control block was initialised
  int step_index = 0;    // steps since the control block was initialised
  int steps_base = 0;    // RK.StepCount before the control block was initialised
  double last_batch_ms = 0;
  int last_batch_launches = 0;
  int NV = 0;
  std::unique_ptr<capi::AdjointRunner> adj;  // tangent-linear / adjoint step, built on first use

  ~gocfd_euler2d() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &p : q) if (p) (void)hipFree(p);
    if (R) (void)hipFree(R);
    if (scratch_in) (void)hipFree(scratch_in);
    if (ctl) (void)hipFree(ctl);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  dev::StageArgs args() const {
    dev::StageArgs a{};
    a.K = L.K; a.Kp = L.Kp; a.ntiles = L.ntiles; a.Hmax = L.Hmax; a.Wmax = L.Wmax; a.Fmax = L.Fmax;
    a.Xmax = L.Xmax;
    a.xcd_tiles = 1;
    a.q0 = q[0]; a.q1 = q[1]; a.q2 = q[2]; a.q3 = q[3]; a.q4 = q[4]; a.R = R;
    a.ctl = ctl;
    a.g.Jdet = Jdet.p; a.g.Jinv = Jinv.p; a.g.IInII = IInII.p; a.g.Nx = Nx.p; a.g.Ny = Ny.p;
    a.g.oohK = oohK.p; a.g.code = code.p; a.g.nbr_slot = nbr_slot.p;
    a.g.halo_elem = halo_elem.p; a.g.halo_face = halo_face.p; a.g.bxy = bxy.p;
    a.g.wave_list = wave_list.p; a.g.wave_cnt = wave_cnt.p; a.g.wave_code = wave_code.p;
    a.g.flux_list = flux_list.p; a.g.flux_cnt = flux_cnt.p; a.g.bnd_cnt = bnd_cnt.p;
    a.g.flux_geo = flux_geo.ptr();
    a.g.xp_list = xp_list.p; a.g.xp_dst = xp_dst.p; a.g.xp_cnt = xp_cnt.p;
    a.ht = ht.p;
    a.Interp64 = Interp64.p;
    a.lts.on = lts ? 1 : 0;
    a.lts.fw = fw.p; a.lts.fwv = fwv.p; a.lts.wsrc = wsrc.p; a.lts.DTk = DTk.p; a.lts.DTv = DTv.p;
    if (shock) {
      a.sh.ops = sh_ops.p; a.sh.opsF = sh_opsF.p; a.sh.evert = evert.p; a.sh.hK = hK.p;
      a.sh.sigV = sigV.p; a.sh.epsV = epsV.p; a.sh.Nv = d.Nv;
      a.sh.Se = Se.p; a.sh.sigma = sigma.p; a.sh.eps = eps.p; a.sh.epsF = epsF.p; a.sh.diss = diss.p;
      a.sh.active = active.p;
      a.sh.act_list = act_list.p;
      a.sh.act_cnt = act_cnt.p;
      a.sh.nbr_elem = nbr_elem.p; a.sh.nbr_face = nbr_face.p;
      a.sh.Nxo = Nxo.p; a.sh.Nyo = Nyo.p; a.sh.vooL = vooL.p;
      a.sh.kappa = kappa; a.sh.eps0 = eps0;
      const double np1 = double(d.N + 1);
      a.sh.cdiff = 1. / (np1 * np1);
      a.sh.omega = 1. * double(d.N * d.N);
    }
    a.ph = phys_of(fs, prm);
    a.cases = cases.p;
    a.tile_case = tile_case.p;
    return a;
  }

  // reference layout [4][NpInt][K] (double) <-> element-major FP64 device
  // buffer (q0 and the RHS-evaluation input)
  void to_device(const double *Q, void *dst) {
    std::vector<double> h(size_t(L.Kp) * NV);
    fill_element_major(L, d.ops.NpInt, Q, h.data());
    HIPCHECK(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  // reference-layout cotangent or tangent -> element-major device buffer,
  // zero on the padding elements
  void to_device_zero_pad(const double *Q, void *dst) {
    std::vector<double> h(size_t(L.Kp) * NV);
    fill_element_major(L, d.ops.NpInt, Q, h.data());
    for (int kk = L.K; kk < L.Kp; ++kk) std::fill(h.begin() + size_t(kk) * NV, h.begin() + size_t(kk + 1) * NV, 0.0);
    HIPCHECK(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  // element-major FP64 device buffer -> reference layout
  void from_device(const void *src, double *Q) {
    std::vector<double> h(size_t(L.Kp) * NV);
    HIPCHECK(hipMemcpy(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    read_element_major(L, d.ops.NpInt, h.data(), Q);
  }

  dev::Ctl read_ctl() {
    dev::Ctl c;
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipMemcpy(&c, ctl, sizeof(c), hipMemcpyDeviceToHost));
    return c;
  }

  // The tangent-linear / adjoint driver (FP64, global or local time
  // stepping; PerssonC0 with global time stepping, or with local time
  // stepping while its sensor is quiet), on the solver's geometry and
  // stream
  capi::AdjointRunner &adjoint() {
    if (prec != 0) throw SetupError(GOCFD_ERR_ARG, "the tangent-linear and adjoint steps differentiate the FP64 step");
    if (shock && lts) {
      // the element viscous dt (DTv, never reset in the reference) must be
      // unset: it stays zero while the sensor has been quiet since the start
      std::vector<double> dtv(DTv.n);
      HIPCHECK(hipMemcpy(dtv.data(), DTv.p, dtv.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (double x : dtv)
        if (x != 0.)
          throw SetupError(GOCFD_ERR_ARG,
                           "LocalTimeStepping with PerssonC0: the element viscous time step is already set (the "
                           "sensor was active before the window); its derivative is not formed");
    }
    if (!adj) {
      capi::AdjointGeometry geo{Jdet.p, Jinv.p, IInII.p, Nx.p, Ny.p, oohK.p, Interp64.p, code.p};
      adj = std::make_unique<capi::AdjointRunner>(d, L, phys_of(fs, prm), prm.init_type != GOCFD_INIT_SHOCKTUBE,
                                                  lts, shock, kappa, geo, stream);
    }
    return *adj;
  }
  double current_time() {
    dev::Ctl c = read_ctl();
    return c.time[c.steps & 1];
  }

  // (Re)start the step sequence from the state in q[0] at `time`
  void restart(double time) {
    HIPCHECK(dev::launch_ctl_init(ctl, time, stream));
    gstage = 0;
    step_index = 0;
    if (shock) {
      HIPCHECK(dev::launch_fill_u64(sigV.p, sigV.n, 0ull, stream));
      HIPCHECK(dev::launch_fill_u64(epsV.p, epsV.n, 0ull, stream));
    }
    if (lts) HIPCHECK(dev::launch_fill_f64(DTv.p, DTv.n, 0.0, stream));
    if (!shock) {  // with shock capturing each stage's gradient pass measures its wave speeds
      dev::StageArgs a = args();
      a.qs = q[0];
      a.ring = 0;
      a.par2 = 0;
      a.hpar = 0;
      a.step_index = 0;
      HIPCHECK(dev::launch_prologue(prm.polynomial_order, prec, a, stream));
      HIPCHECK(dev::launch_halo_export(prm.polynomial_order, prec, a, stream));
    }
  }

  // Arguments of an RHS evaluation of the element-major FP64 state qs into
  // out, with its halo traces in the buffer the next stage does not read
  // (the one the last stage read); on stream s
  dev::StageArgs rhs_args(const void *qs, double t, double *out, hipStream_t s) {
    dev::StageArgs a = args();
    a.stage = -1;
    a.t_rhs = t;
    a.qs = const_cast<void *>(qs);
    a.rhs_out = out;
    a.ring = int(gstage % 3);
    a.par2 = int(gstage & 1);
    a.hpar = a.par2 ^ 1;
    a.step_index = step_index;
    if (!shock) HIPCHECK(dev::launch_halo_export(prm.polynomial_order, prec, a, s));
    return a;
  }

  // One stage: the shock-capturing passes (if any), then the stage kernel,
  // on stream s (the handle's stream when null)
  void launch_one_stage(dev::StageArgs &a, hipStream_t s = nullptr) {
    const int N = prm.polynomial_order;
    if (!s) s = stream;
    if (shock) {
      a.sh.vbuf = int(gstage & 1);
      a.sh.filter_state = a.stage == 2 ? 1 : 0;
      for (int pass = 0; pass < dev::kShockPasses; ++pass) HIPCHECK(dev::launch_shock(N, prec, pass, a, s));
    }
    HIPCHECK(dev::launch_stage(N, prec, shock ? 1 : 0, a, s));
  }

  void enqueue_steps(int nsteps) {
    dev::StageArgs a = args();
    HIPCHECK(hipEventRecord(ev0, stream));
    for (int j = 0; j < nsteps; ++j) {
      for (int s = 0; s < 5; ++s) {
        a.stage = s;
        a.step_index = step_index;
        a.ring = int(gstage % 3);
        a.par2 = int(gstage & 1);
        a.hpar = a.par2;
        a.qs = q[s];
        launch_one_stage(a);
        ++gstage;
      }
      ++step_index;
    }
    HIPCHECK(hipEventRecord(ev1, stream));
    last_batch_launches = 5 * nsteps;
  }

  int finish_batch() {
    HIPCHECK(hipEventSynchronize(ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
    last_batch_ms = ms;
    dev::Ctl c = read_ctl();
    if (c.nan) return fail(GOCFD_ERR_NAN, "NAN found");
    return GOCFD_OK;
  }
};

#define NEED_DEVICE(h) \
  if ((h)->device < 0) return fail(GOCFD_ERR_DEVICE, "handle was created without a device");

namespace gocfd {
namespace capi {
// Per-device set-up of the no-allocation RHS entry point (noalloc.hip):
// resolves the device address of its static scratch array on the current
// device, once per device.
void noalloc_setup(int device);
}  // namespace capi
}  // namespace gocfd
