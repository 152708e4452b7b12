// Batched ensembles: many independent cases (meshes of different sizes,
// orders, Mach numbers and angles) stepped together. The reference runs one
// case per process; this has no counterpart there.
//
// Cases of the same order, limiter (PerssonC0 with one Kappa) and
// time-stepping mode form a group whose device layout is the concatenation
// of the cases' layouts (every case starts on a tile boundary, so a tile
// belongs to one case); each stage is one launch per group (after the three
// shock-capturing passes in a PerssonC0 group), or one launch for all the
// groups without shock capturing (a multi-order kernel: each workgroup runs
// its tile's order) plus one per PerssonC0 group. Every
// case keeps its own control block (time, dt, wave-speed maxima) and
// physics constants, selected per tile. A run of 6 steps -- the period of
// the per-stage ring and time-parity arguments -- or a multiple of it is
// captured once as a HIP graph and replayed. The launch configuration (the
// groups serialized on one stream, as parallel graph branches or in one
// multi-order launch, the steps per graph, the XCD tile order) can be picked
// by timing the candidates
// (gocfd_ensemble_tune); it does not change the results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <vector>

#include "capi_common.hpp"
#include "dfr_setup.hpp"
#include "gocfd_euler2d.h"
#include "kernels.hpp"

using namespace gocfd;
using namespace gocfd::capi;

namespace {

constexpr int kGraphPeriod = 6;  // steps after which every per-stage argument repeats

// rocprofv3 exports these to the program it runs (kernel tracing, counter
// collection). A captured graph replayed under its kernel-trace interception
// ended in a segmentation fault inside the HIP runtime's graph launch; the
// kernels are the same when launched one by one, so with the profiler
// attached an ensemble launches eagerly (graph_steps 0) unless told otherwise.
bool profiler_attached() {
  for (const char *v : {"ROCPROF_KERNEL_TRACE", "ROCPROF_COUNTER_COLLECTION"}) {
    const char *e = std::getenv(v);
    if (e && *e && std::strcmp(e, "0") != 0) return true;
  }
  return false;
}

// Test hook: with GOCFD_TEST_FAIL_CAPTURE set (and not "0"), a graph
// capture throws after its first captured step, as an error from a launch
// inside the capture would; the tests use it to check that the handle stays
// usable. No kernel is affected.
bool capture_failure_requested() {
  const char *e = std::getenv("GOCFD_TEST_FAIL_CAPTURE");
  return e && *e && std::strcmp(e, "0") != 0;
}

// Ends a stream capture that an exception interrupts: branch streams forked
// into the capture are joined back (an unjoined fork would make the end of
// the capture fail), the stream leaves capture mode, the partial graph is
// destroyed and the stage counters that the capture advanced are restored,
// so that the handle stays usable.
struct CaptureGuard {
  hipStream_t stream;
  long long *gstage;
  int *step_index;
  long long g0;
  int s0;
  std::vector<std::pair<hipStream_t, hipEvent_t>> joins;  // forked branch streams, their join events
  bool active = true;
  ~CaptureGuard() {
    if (!active) return;
    for (auto &j : joins) {
      (void)hipEventRecord(j.second, j.first);
      (void)hipStreamWaitEvent(stream, j.second, 0);
    }
    hipGraph_t partial = nullptr;
    (void)hipStreamEndCapture(stream, &partial);
    if (partial) (void)hipGraphDestroy(partial);
    (void)hipGetLastError();
    *gstage = g0;
    *step_index = s0;
  }
};

struct EnsCase {
  gocfd_params prm{};
  Discretisation d;
  DeviceLayout L;
  CaseStreams fs;
  std::vector<double> Q0;
  int group = 0, off = 0, boff = 0;
};

struct EnsGroup {
  int N = 0, Kp = 0, ntiles = 0, Hmax = 1, Wmax = 1, Fmax = 1, Xmax = 1, NV = 0;
  bool shock = false, lts = false;  // PerssonC0 limiter, LocalTimeStepping (the same for the group)
  double kappa = 0;
  int Nv = 0;                       // vertices of the group's cases
  std::vector<int> cases;
  DevBuf<double> Jdet, Jinv, IInII, Nx, Ny, oohK, bxy, Interp64;
  DevBuf<int> code, nbr_slot, halo_elem, halo_face, wave_list, wave_cnt, wave_code, flux_list, flux_cnt, bnd_cnt, tile_case;
  // halo traces (without shock capturing): export lists and the two buffers
  DevBuf<int> xp_list, xp_dst, xp_cnt;
  DevBuf<double> ht;
  StageTypeBuf flux_geo;
  void *q[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  void *R = nullptr;
  std::vector<double> q0_init;         // initial element-major state (tuning restores it)
  // local time stepping
  DevBuf<double> fw, fwv, DTk, DTv;
  DevBuf<int> wsrc;
  // shock capturing
  DevBuf<double> sh_ops, hK, Se, sigma, eps, epsF, diss, Nxo, Nyo, vooL;
  DevBuf<float> sh_opsF;
  DevBuf<int> evert, nbr_elem, nbr_face, active, act_list;
  DevBuf<unsigned> act_cnt;
  DevBuf<unsigned long long> sigV, epsV;
  hipStream_t stream = nullptr;        // the group's branch of a parallel graph
  hipEvent_t done = nullptr;
  ~EnsGroup() {
    for (auto &p : q) if (p) (void)hipFree(p);
    if (R) (void)hipFree(R);
    if (done) (void)hipEventDestroy(done);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace

struct gocfd_ensemble {
  int device = 0, prec = 0, flux = 0;
  size_t ssize = 8;
  std::vector<std::unique_ptr<EnsCase>> cases;
  std::vector<std::unique_ptr<EnsGroup>> groups;
  DevBuf<dev::CaseInfo> cinfo;
  dev::Ctl *ctl = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long long gstage = 0;
  int step_index = 0;
  double last_ms = 0;
  int last_steps = 0;
  // launch configuration
  int parallel = 0;              // 0 groups in sequence, 1 parallel graph branches, 2 one multi-order launch
  // the groups without shock capturing as one multi-order launch per stage
  DevBuf<dev::StageArgs> margs;  // their arguments (device)
  std::vector<int> mgroups;      // their indices
  bool multi_ok = false;
  int graph_steps = kGraphPeriod;
  int xcd_tiles = 1;

  ~gocfd_ensemble() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    groups.clear();
    if (ctl) (void)hipFree(ctl);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (fork) (void)hipEventDestroy(fork);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  dev::StageArgs args(const EnsGroup &g) const {
    dev::StageArgs a{};
    a.K = g.Kp; a.Kp = g.Kp; a.ntiles = g.ntiles; a.Hmax = g.Hmax; a.Wmax = g.Wmax; a.Fmax = g.Fmax;
    a.Xmax = g.Xmax;
    a.xcd_tiles = xcd_tiles;
    a.q0 = g.q[0]; a.q1 = g.q[1]; a.q2 = g.q[2]; a.q3 = g.q[3]; a.q4 = g.q[4]; a.R = g.R;
    a.ctl = ctl;
    a.cases = cinfo.p;
    a.tile_case = g.tile_case.p;
    a.g.Jdet = g.Jdet.p; a.g.Jinv = g.Jinv.p; a.g.IInII = g.IInII.p; a.g.Nx = g.Nx.p; a.g.Ny = g.Ny.p;
    a.g.oohK = g.oohK.p; a.g.code = g.code.p; a.g.nbr_slot = g.nbr_slot.p;
    a.g.halo_elem = g.halo_elem.p; a.g.halo_face = g.halo_face.p; a.g.bxy = g.bxy.p;
    a.g.wave_list = g.wave_list.p; a.g.wave_cnt = g.wave_cnt.p; a.g.wave_code = g.wave_code.p;
    a.g.flux_list = g.flux_list.p; a.g.flux_cnt = g.flux_cnt.p; a.g.bnd_cnt = g.bnd_cnt.p;
    a.g.flux_geo = g.flux_geo.ptr();
    a.g.xp_list = g.xp_list.p; a.g.xp_dst = g.xp_dst.p; a.g.xp_cnt = g.xp_cnt.p;
    a.ht = g.ht.p;
    a.Interp64 = g.Interp64.p;
    const EnsCase &c0 = *cases[g.cases[0]];
    a.ph = phys_of(c0.fs, c0.prm);
    a.lts.on = g.lts ? 1 : 0;
    a.lts.fw = g.fw.p; a.lts.fwv = g.fwv.p; a.lts.wsrc = g.wsrc.p; a.lts.DTk = g.DTk.p; a.lts.DTv = g.DTv.p;
    if (g.shock) {
      a.sh.ops = g.sh_ops.p; a.sh.opsF = g.sh_opsF.p; a.sh.evert = g.evert.p; a.sh.hK = g.hK.p;
      a.sh.sigV = g.sigV.p; a.sh.epsV = g.epsV.p; a.sh.Nv = g.Nv;
      a.sh.Se = g.Se.p; a.sh.sigma = g.sigma.p; a.sh.eps = g.eps.p; a.sh.epsF = g.epsF.p; a.sh.diss = g.diss.p;
      a.sh.active = g.active.p;
      a.sh.act_list = g.act_list.p;
      a.sh.act_cnt = g.act_cnt.p;
      a.sh.nbr_elem = g.nbr_elem.p; a.sh.nbr_face = g.nbr_face.p;
      a.sh.Nxo = g.Nxo.p; a.sh.Nyo = g.Nyo.p; a.sh.vooL = g.vooL.p;
      a.sh.kappa = g.kappa; a.sh.eps0 = 5. / 1.5;
      const double np1 = double(g.N + 1);
      a.sh.cdiff = 1. / (np1 * np1);
      a.sh.omega = double(g.N * g.N);
    }
    return a;
  }

  void restart() {
    HIPCHECK(dev::launch_ctl_init(ctl, 0.0, stream, int(cases.size())));
    gstage = 0;
    step_index = 0;
    for (auto &g : groups) {
      if (g->shock) {
        HIPCHECK(dev::launch_fill_u64(g->sigV.p, g->sigV.n, 0ull, stream));
        HIPCHECK(dev::launch_fill_u64(g->epsV.p, g->epsV.n, 0ull, stream));
      }
      if (g->lts) HIPCHECK(dev::launch_fill_f64(g->DTv.p, g->DTv.n, 0.0, stream));
      // with shock capturing each stage's gradient pass measures its wave speeds
      if (g->shock) continue;
      dev::StageArgs a = args(*g);
      a.qs = g->q[0];
      a.ring = 0;
      a.par2 = 0;
      a.hpar = 0;
      a.step_index = 0;
      HIPCHECK(dev::launch_prologue(g->N, prec, a, stream));
      HIPCHECK(dev::launch_halo_export(g->N, prec, a, stream));
    }
  }

  // One SSP54 step of every group; with `branches` each group's stages go
  // to the group's own stream (inside a graph capture: parallel branches).
  void enqueue_step(bool branches) {
    const bool multi = parallel == 2 && multi_ok;
    for (int s = 0; s < 5; ++s) {
      if (multi) {
        dev::MultiArgs m{};
        m.ngroups = int(mgroups.size());
        m.xcd_tiles = xcd_tiles;
        m.l = dev::MultiLaunch{s, step_index, int(gstage % 3), int(gstage & 1), int(gstage & 1)};
        int t0 = 0;
        for (size_t i = 0; i < mgroups.size(); ++i) {
          const EnsGroup &g = *groups[size_t(mgroups[i])];
          m.g[i] = dev::MultiGroup{margs.p + i, t0, g.N, g.Fmax, g.Wmax + g.Xmax, g.lts ? 1 : 0};
          t0 += g.ntiles;
        }
        m.ntiles = t0;
        HIPCHECK(dev::launch_stage_multi(prec, flux, m, stream));
      }
      for (auto &g : groups) {
        if (multi && !g->shock) continue;
        dev::StageArgs a = args(*g);
        a.stage = s;
        a.step_index = step_index;
        a.ring = int(gstage % 3);
        a.par2 = int(gstage & 1);
        a.hpar = a.par2;
        a.qs = g->q[s];
        hipStream_t st = branches ? g->stream : stream;
        if (g->shock) {
          a.sh.vbuf = int(gstage & 1);
          a.sh.filter_state = s == 2 ? 1 : 0;
          for (int pass = 0; pass < dev::kShockPasses; ++pass) HIPCHECK(dev::launch_shock(g->N, prec, pass, a, st));
        }
        HIPCHECK(dev::launch_stage(g->N, prec, g->shock ? 1 : 0, a, st));
      }
      ++gstage;
    }
    ++step_index;
  }

  // Steps run eagerly (one stream) until the step counter is 1 mod
  // graph_steps; from there whole blocks replay the captured graph (captured
  // on first use; the groups fork from the ensemble's stream and join it
  // again inside the graph when `parallel`), and the remainder runs eagerly.
  // graph_steps 0: every step runs eagerly.
  void enqueue_steps(int nsteps) {
    HIPCHECK(hipEventRecord(ev0, stream));
    int left = nsteps;
    while (left > 0) {
      if (graph_steps > 0 && step_index % graph_steps == 1 && left >= graph_steps) {
        if (!gexec) capture_block();
        HIPCHECK(hipGraphLaunch(gexec, stream));
        gstage += 5LL * graph_steps;
        step_index += graph_steps;
        left -= graph_steps;
      } else {
        enqueue_step(false);
        --left;
      }
    }
    HIPCHECK(hipEventRecord(ev1, stream));
    last_steps = nsteps;
  }

  // Captures graph_steps steps into `graph` and instantiates `gexec`; the
  // stage counters are left where they were.
  void capture_block() {
    const long long g0 = gstage;
    const int s0 = step_index;
    hipGraph_t g = nullptr;
    {
      HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      CaptureGuard guard{stream, &gstage, &step_index, g0, s0};
      const bool br = parallel == 1 && groups.size() > 1;
      if (br) {
        HIPCHECK(hipEventRecord(fork, stream));
        for (auto &gr : groups) {
          HIPCHECK(hipStreamWaitEvent(gr->stream, fork, 0));
          guard.joins.push_back({gr->stream, gr->done});
        }
      }
      for (int j = 0; j < graph_steps; ++j) {
        enqueue_step(br);
        if (j == 0 && capture_failure_requested())
          throw SetupError(GOCFD_ERR_DEVICE, "graph capture interrupted (GOCFD_TEST_FAIL_CAPTURE is set)");
      }
      if (br) {
        for (auto &gr : groups) {
          HIPCHECK(hipEventRecord(gr->done, gr->stream));
          HIPCHECK(hipStreamWaitEvent(stream, gr->done, 0));
        }
      }
      guard.active = false;
      const hipError_t e = hipStreamEndCapture(stream, &g);
      gstage = g0;
      step_index = s0;
      if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        HIPCHECK(e);
      }
    }
    const hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      gexec = nullptr;
      HIPCHECK(e);
    }
    graph = g;
  }

  // Back to the initial state and time 0 (the stage states and R are
  // written before they are read in every step).
  void reset_state() {
    HIPCHECK(hipStreamSynchronize(stream));
    for (auto &g : groups)
      HIPCHECK(hipMemcpy(g->q[0], g->q0_init.data(), g->q0_init.size() * sizeof(double), hipMemcpyHostToDevice));
    restart();
    HIPCHECK(hipStreamSynchronize(stream));
  }

  int set_launch(int par, int gsteps, int xcd) {
    if (gsteps < 0 || gsteps % kGraphPeriod != 0 || gsteps > 600)
      return fail(GOCFD_ERR_ARG, "graph_steps must be 0 (eager launches) or a multiple of 6 (at most 600)");
    if (par < 0 || par > 2) return fail(GOCFD_ERR_ARG, "parallel must be 0, 1 or 2");
    if (par == 2 && !multi_ok)
      return fail(GOCFD_ERR_ARG, "no multi-order launch for this ensemble (FP32, no group without shock capturing, "
                                 "more than " + std::to_string(dev::kMultiMax) +
                                     " such groups, or a work list longer than a workgroup)");
    HIPCHECK(hipStreamSynchronize(stream));
    drop_graph();
    parallel = par;
    graph_steps = gsteps;
    xcd_tiles = xcd ? 1 : 0;
    return GOCFD_OK;
  }

  // Times every candidate configuration over `blocks` graph replays from
  // the initial state, keeps the fastest and restores the initial state.
  int tune(int blocks) {
    struct Cfg {
      int par, gs, xcd;
    };
    std::vector<Cfg> cand;
    if (profiler_attached()) {  // eager launches only (see profiler_attached)
      for (int par : {0, 2})
        if (par == 0 || multi_ok)
          for (int xcd : {1, 0}) cand.push_back({par, 0, xcd});
    } else {
      for (int par = 0; par <= 2; ++par) {
        if ((par == 1 && groups.size() < 2) || (par == 2 && !multi_ok)) continue;
        for (int gs : {kGraphPeriod, 4 * kGraphPeriod})
          for (int xcd : {1, 0}) cand.push_back({par, gs, xcd});
      }
    }
    double best = 1e300;
    Cfg pick = cand[0];
    for (const Cfg &c : cand) {
      int rc = set_launch(c.par, c.gs, c.xcd);
      if (rc != GOCFD_OK) return rc;
      const int bsteps = c.gs > 0 ? c.gs : kGraphPeriod;
      reset_state();
      enqueue_steps(1 + bsteps);  // eager first step, then the block that captures the graph
      if ((rc = finish()) != GOCFD_OK) return rc;
      enqueue_steps(blocks * bsteps);
      if ((rc = finish()) != GOCFD_OK) return rc;
      const double ms = last_ms / (blocks * bsteps);
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    int rc = set_launch(pick.par, pick.gs, pick.xcd);
    if (rc != GOCFD_OK) return rc;
    reset_state();
    last_ms = best;
    last_steps = 1;
    return GOCFD_OK;
  }

  int finish() {
    HIPCHECK(hipEventSynchronize(ev1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
    last_ms = ms;
    std::vector<dev::Ctl> c(cases.size());
    HIPCHECK(hipMemcpy(c.data(), ctl, c.size() * sizeof(dev::Ctl), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < c.size(); ++i)
      if (c[i].nan) return fail(GOCFD_ERR_NAN, "NAN found in case " + std::to_string(i));
    return GOCFD_OK;
  }
};

namespace {

// Concatenate the cases of one order into a group's device layout.
void build_group(gocfd_ensemble &h, EnsGroup &g) {
  const int E = dev::kTile;
  int off = 0, boff = 0;
  for (int ci : g.cases) {
    EnsCase &c = *h.cases[ci];
    c.off = off;
    c.boff = boff;
    off += c.L.Kp;
    boff += c.L.nbfaces;
    g.Hmax = std::max(g.Hmax, c.L.Hmax);
    g.Wmax = std::max(g.Wmax, c.L.Wmax);
    g.Fmax = std::max(g.Fmax, c.L.Fmax);
    g.Xmax = std::max(g.Xmax, c.L.Xmax);
  }
  g.Kp = off;
  g.ntiles = off / E;
  const int N = g.N, NpEdge = N + 2, NpInt = (N + 1) * (N + 2) / 2;
  g.NV = 4 * NpInt;
  const size_t Kp = g.Kp;
  std::vector<double> Jdet(Kp), Jinv(4 * Kp), IInII(3 * Kp), Nx(3 * Kp), Ny(3 * Kp), oohK(Kp), bxy;
  std::vector<int> code(3 * Kp), nbr_slot(3 * Kp), tile_case(g.ntiles);
  std::vector<int> halo_elem(size_t(g.ntiles) * g.Hmax, -1), halo_face(size_t(g.ntiles) * g.Hmax, 0);
  std::vector<int> wave_list(size_t(g.ntiles) * g.Wmax, 0), wave_cnt(g.ntiles, 0);
  std::vector<int> flux_list(size_t(g.ntiles) * g.Fmax, 0), flux_cnt(g.ntiles, 0), bnd_cnt(g.ntiles, 0);
  std::vector<int> xp_list(size_t(g.ntiles) * g.Xmax, 0), xp_dst(size_t(g.ntiles) * g.Xmax, 0), xp_cnt(g.ntiles, 0);
  std::vector<double> flux_geo(size_t(g.ntiles) * g.Fmax * 4, 0.0);
  std::vector<double> q0(Kp * g.NV);
  for (int ci : g.cases) {
    const EnsCase &c = *h.cases[ci];
    const DeviceLayout &L = c.L;
    const size_t o = c.off, kp = L.Kp;
    for (size_t kk = 0; kk < kp; ++kk) {
      Jdet[o + kk] = L.Jdet[kk];
      oohK[o + kk] = L.oohK[kk];
      for (int j = 0; j < 4; ++j) Jinv[j * Kp + o + kk] = L.Jinv[j * kp + kk];
      for (int f = 0; f < 3; ++f) {
        const size_t src = f * kp + kk, dst = f * Kp + o + kk;
        IInII[dst] = L.IInII[src];
        Nx[dst] = L.Nx[src];
        Ny[dst] = L.Ny[src];
        nbr_slot[dst] = L.nbr_slot[src];  // tile-relative
        int cd = L.face_code[src];
        if ((cd & 15) == FK_IVortex) cd = (cd & 0xff) | (((cd >> 8) + c.boff) << 8);
        code[dst] = cd;
      }
    }
    bxy.insert(bxy.end(), L.bpoint_xy.begin(), L.bpoint_xy.end());
    const int t0 = c.off / E;
    for (int t = 0; t < L.ntiles; ++t) {
      tile_case[t0 + t] = ci;
      for (int i = 0; i < L.Hmax; ++i) {
        const int e = L.halo_elem[size_t(t) * L.Hmax + i];
        halo_elem[size_t(t0 + t) * g.Hmax + i] = e < 0 ? -1 : e + c.off;
        halo_face[size_t(t0 + t) * g.Hmax + i] = L.halo_face[size_t(t) * L.Hmax + i];
      }
      wave_cnt[t0 + t] = L.wave_cnt[t];
      for (int i = 0; i < L.wave_cnt[t]; ++i) wave_list[size_t(t0 + t) * g.Wmax + i] = L.wave_list[size_t(t) * L.Wmax + i];
      flux_cnt[t0 + t] = L.flux_cnt[t];
      bnd_cnt[t0 + t] = L.bnd_cnt[t];
      for (int i = 0; i < L.flux_cnt[t]; ++i) {
        flux_list[size_t(t0 + t) * g.Fmax + i] = L.flux_list[size_t(t) * L.Fmax + i];
        for (int c = 0; c < 4; ++c)
          flux_geo[(size_t(t0 + t) * g.Fmax + i) * 4 + c] = L.flux_geo[(size_t(t) * L.Fmax + i) * 4 + c];
      }
      xp_cnt[t0 + t] = L.xp_cnt[t];
      for (int i = 0; i < L.xp_cnt[t]; ++i) {
        // the item slot tile * Fmax + item of the case -> of the group
        const int d = L.xp_dst[size_t(t) * L.Xmax + i], dt = d / L.Fmax, di = d - dt * L.Fmax;
        xp_list[size_t(t0 + t) * g.Xmax + i] = L.xp_list[size_t(t) * L.Xmax + i];
        xp_dst[size_t(t0 + t) * g.Xmax + i] = (t0 + dt) * g.Fmax + di;
      }
    }
    fill_element_major(L, NpInt, c.Q0.data(), q0.data() + o * g.NV);
  }
  g.Jdet.upload(Jdet); g.Jinv.upload(Jinv); g.IInII.upload(IInII); g.Nx.upload(Nx); g.Ny.upload(Ny);
  g.oohK.upload(oohK); g.bxy.upload(bxy); g.code.upload(code); g.nbr_slot.upload(nbr_slot);
  g.halo_elem.upload(halo_elem); g.halo_face.upload(halo_face);
  // the (group-offset) code of each wave-list face
  std::vector<int> wave_code(size_t(g.ntiles) * g.Wmax, 0);
  for (int t = 0; t < g.ntiles; ++t)
    for (int i = 0; i < wave_cnt[t]; ++i) {
      const int w = wave_list[size_t(t) * g.Wmax + i];
      wave_code[size_t(t) * g.Wmax + i] = code[size_t(w >> 8) * Kp + size_t(t) * E + (w & 255)];
    }
  g.wave_list.upload(wave_list); g.wave_cnt.upload(wave_cnt); g.wave_code.upload(wave_code);
  g.flux_list.upload(flux_list); g.flux_cnt.upload(flux_cnt); g.bnd_cnt.upload(bnd_cnt);
  g.flux_geo.upload(flux_geo, h.prec, N);
  g.tile_case.upload(tile_case);
  if (!g.shock) {
    g.xp_list.upload(xp_list); g.xp_dst.upload(xp_dst); g.xp_cnt.upload(xp_cnt);
    g.ht.alloc(size_t(2) * g.ntiles * g.Fmax * NpEdge * 4);  // FP64 size, enough for either arithmetic
  }
  const EnsCase &c0 = *h.cases[g.cases[0]];
  g.Interp64.upload(c0.d.ops.Interp);
  std::vector<double> divT(c0.d.ops.DivInt.size());
  for (int i = 0; i < NpInt; ++i)
    for (int j = 0; j < c0.d.ops.NpFlux; ++j)
      divT[size_t(j) * NpInt + i] = c0.d.ops.DivInt[size_t(i) * c0.d.ops.NpFlux + j];
  HIPCHECK(dev::upload_operators(N, divT.data(), c0.d.ops.Interp.data()));
  if (!g.shock) HIPCHECK(dev::upload_operators_multi(h.flux, N, divT.data(), c0.d.ops.Interp.data()));
  const size_t nval = Kp * g.NV;
  HIPCHECK(hipMalloc(&g.q[0], nval * sizeof(double)));
  HIPCHECK(hipMemcpy(g.q[0], q0.data(), nval * sizeof(double), hipMemcpyHostToDevice));
  g.q0_init = std::move(q0);
  if (g.lts || g.shock) {
    // per-element and per-face arrays of the cases, concatenated like the
    // geometry; element, face and vertex indices offset per case
    std::vector<int> wsrc(3 * Kp, 0), evert(3 * Kp, 0), nbr_elem(3 * Kp, -1), nbr_face(3 * Kp, 0);
    std::vector<double> hK(Kp, 0.0), Nxo(3 * Kp, 0.0), Nyo(3 * Kp, 0.0), vooL(3 * Kp, 0.0);
    int voff = 0;
    for (int ci : g.cases) {
      const EnsCase &c = *h.cases[ci];
      const DeviceLayout &L = c.L;
      const size_t o = c.off, kp = L.Kp;
      for (size_t kk = 0; kk < kp; ++kk) {
        if (g.shock) hK[o + kk] = L.hK[kk];
        for (int f = 0; f < 3; ++f) {
          const size_t src = f * kp + kk, dst = f * Kp + o + kk;
          if (g.lts) {
            const int w = L.wsrc[src], wf = int(w / kp), wk = int(w % kp);
            wsrc[dst] = int(wf * Kp + o + wk);
          }
          if (g.shock) {
            evert[dst] = L.evert[src] + voff;
            nbr_elem[dst] = L.nbr_elem[src] < 0 ? -1 : L.nbr_elem[src] + int(o);
            nbr_face[dst] = L.nbr_face[src];
            Nxo[dst] = L.Nxo[src];
            Nyo[dst] = L.Nyo[src];
            vooL[dst] = L.vooL[src];
          }
        }
      }
      voff += c.d.Nv;
    }
    if (g.lts) {
      g.fw.alloc(2 * 3 * Kp);
      g.fwv.alloc(2 * 3 * Kp);
      g.DTk.alloc(Kp);
      g.DTv.alloc(Kp);
      g.wsrc.upload(wsrc);
    }
    if (g.shock) {
      g.Nv = voff;
      const int NR = NpInt + 3 * NpEdge;
      const std::vector<double> sops = build_shock_operators(c0.d.ops);
      g.sh_ops.upload(sops);
      g.sh_opsF.upload(std::vector<float>(sops.begin(), sops.end()));
      g.evert.upload(evert); g.hK.upload(hK);
      g.nbr_elem.upload(nbr_elem); g.nbr_face.upload(nbr_face);
      g.Nxo.upload(Nxo); g.Nyo.upload(Nyo); g.vooL.upload(vooL);
      g.sigV.alloc(2 * size_t(std::max(voff, 1)));
      g.epsV.alloc(2 * size_t(std::max(voff, 1)));
      g.Se.alloc(Kp); g.sigma.alloc(Kp); g.eps.alloc(Kp);
      g.epsF.alloc(Kp * NR);
      g.diss.alloc(Kp * 4 * 2 * NR);  // FP64 or FP32 values (the larger size)
      g.active.alloc(Kp);
      g.act_list.alloc(Kp);
      g.act_cnt.alloc(1);
    }
  }
  HIPCHECK(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  for (int s = 1; s < 5; ++s) {
    HIPCHECK(hipMalloc(&g.q[s], nval * h.ssize));
    HIPCHECK(hipMemset(g.q[s], 0, nval * h.ssize));
  }
  HIPCHECK(hipMalloc(&g.R, nval * sizeof(double)));
  HIPCHECK(hipMemset(g.R, 0, nval * sizeof(double)));
}

}  // namespace

extern "C" {

int gocfd_ensemble_create(int32_t ncases, const gocfd_mesh *const *meshes, const gocfd_params *params,
                          gocfd_ensemble **out) {
  GUARD_BEGIN
  if (ncases <= 0 || !meshes || !params || !out) return fail(GOCFD_ERR_ARG, "bad argument");
  *out = nullptr;
  auto h = std::make_unique<gocfd_ensemble>();
  h->device = params[0].device;
  h->prec = params[0].precision;
  h->flux = params[0].flux_type;
  h->ssize = h->prec == 0 ? sizeof(double) : sizeof(float);
  if (profiler_attached()) h->graph_steps = 0;  // eager launches under rocprofv3
  for (int i = 0; i < ncases; ++i) {
    const gocfd_params &p = params[i];
    if (!meshes[i]) return fail(GOCFD_ERR_ARG, "null mesh");
    if (validate_params(&p) != GOCFD_OK) return GOCFD_ERR_ARG;
    if (p.precision != h->prec || p.flux_type != h->flux || p.device != h->device)
      return fail(GOCFD_ERR_ARG, "every case of an ensemble needs the same precision, flux type and device");
    auto c = std::make_unique<EnsCase>();
    c->prm = p;
    c->d = build_discretisation(meshes[i]->m, p.polynomial_order);
    init_case(c->d, p, c->fs, c->Q0);
    c->L = build_device_layout(c->d, dev::kTile, dev::kTP);
    h->cases.push_back(std::move(c));
  }
  // one group per order, limiter (with its Kappa) and time-stepping mode:
  // a group shares one kernel variant and one set of shock constants
  std::map<std::tuple<int, int, int, double>, std::vector<int>> by_kind;
  for (int i = 0; i < ncases; ++i) {
    const gocfd_params &p = h->cases[i]->prm;
    const bool shock = p.limiter == 1 && p.polynomial_order != 0;
    const double kappa = shock ? (p.kappa != 0. ? p.kappa : 5.) : 0.;
    by_kind[std::make_tuple(p.polynomial_order, shock ? 1 : 0, p.local_time_stepping ? 1 : 0, kappa)].push_back(i);
  }
  size_t lds = 0;
  for (auto &kv : by_kind) {
    auto g = std::make_unique<EnsGroup>();
    g->N = std::get<0>(kv.first);
    g->shock = std::get<1>(kv.first) != 0;
    g->lts = std::get<2>(kv.first) != 0;
    g->kappa = std::get<3>(kv.first);
    g->cases = kv.second;
    for (int ci : g->cases) {
      h->cases[ci]->group = int(h->groups.size());
      lds = std::max(lds, dev::stage_lds_bytes(g->N, h->prec, h->cases[ci]->L.Hmax));
    }
    h->groups.push_back(std::move(g));
  }
  check_device(h->device, lds + 4096);
  HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&h->ev0));
  HIPCHECK(hipEventCreate(&h->ev1));
  HIPCHECK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
  for (auto &g : h->groups) build_group(*h, *g);
  std::vector<dev::CaseInfo> ci(ncases);
  for (int i = 0; i < ncases; ++i) {
    const EnsCase &c = *h->cases[i];
    ci[i].ph = phys_of(c.fs, c.prm);
    ci[i].k_end = c.off + c.L.K;
    ci[i].tile0 = c.off / dev::kTile;
  }
  h->cinfo.upload(ci);
  HIPCHECK(hipMalloc(&h->ctl, sizeof(dev::Ctl) * ncases));
  {
    // the groups without shock capturing, for the multi-order launch
    std::vector<dev::StageArgs> ma;
    bool ok = true;
    for (size_t i = 0; i < h->groups.size(); ++i) {
      const EnsGroup &g = *h->groups[i];
      if (g.shock) continue;
      ok = ok && g.Fmax <= dev::kThreads && g.Wmax <= dev::kThreads && g.Xmax <= dev::kThreads;
      h->mgroups.push_back(int(i));
      ma.push_back(h->args(g));
    }
    // FP64 only: the FP32 multi-order kernel faulted the GPU in its first
    // run and stays off until that is explained
    h->multi_ok = ok && h->prec == 0 && !ma.empty() && ma.size() <= size_t(dev::kMultiMax);
    if (h->multi_ok) h->margs.upload(ma);
  }
  h->restart();
  HIPCHECK(hipStreamSynchronize(h->stream));
  *out = h.release();
  return GOCFD_OK;
  GUARD_END
}

void gocfd_ensemble_destroy(gocfd_ensemble *h) { delete h; }

int gocfd_ensemble_case_dims(gocfd_ensemble *h, int32_t c, int32_t *K, int32_t *NpInt) {
  if (!h || c < 0 || c >= int(h->cases.size())) return fail(GOCFD_ERR_ARG, "bad case index");
  if (K) *K = h->cases[c]->d.K;
  if (NpInt) *NpInt = h->cases[c]->d.ops.NpInt;
  return GOCFD_OK;
}

int gocfd_ensemble_step(gocfd_ensemble *h, int32_t nsteps) {
  GUARD_BEGIN
  if (!h || nsteps < 0) return fail(GOCFD_ERR_ARG, "bad argument");
  HIPCHECK(hipSetDevice(h->device));
  h->enqueue_steps(nsteps);
  return h->finish();
  GUARD_END
}

int gocfd_ensemble_get_state(gocfd_ensemble *h, int32_t c, double *Q) {
  GUARD_BEGIN
  if (!h || !Q || c < 0 || c >= int(h->cases.size())) return fail(GOCFD_ERR_ARG, "bad argument");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  const EnsCase &cs = *h->cases[c];
  const EnsGroup &g = *h->groups[cs.group];
  std::vector<double> buf(size_t(cs.L.Kp) * g.NV);
  HIPCHECK(hipMemcpy(buf.data(), static_cast<double *>(g.q[0]) + size_t(cs.off) * g.NV,
                     buf.size() * sizeof(double), hipMemcpyDeviceToHost));
  read_element_major(cs.L, cs.d.ops.NpInt, buf.data(), Q);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_ensemble_time(gocfd_ensemble *h, int32_t c, double *time, double *global_dt, int32_t *steps) {
  GUARD_BEGIN
  if (!h || c < 0 || c >= int(h->cases.size())) return fail(GOCFD_ERR_ARG, "bad argument");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  dev::Ctl k;
  HIPCHECK(hipMemcpy(&k, h->ctl + c, sizeof(k), hipMemcpyDeviceToHost));
  if (time) *time = k.time[k.steps & 1];
  if (global_dt) *global_dt = k.dt[4];
  if (steps) *steps = k.steps;
  return GOCFD_OK;
  GUARD_END
}

int gocfd_ensemble_set_launch(gocfd_ensemble *h, int32_t parallel, int32_t graph_steps, int32_t xcd_tiles) {
  GUARD_BEGIN
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  HIPCHECK(hipSetDevice(h->device));
  return h->set_launch(parallel, graph_steps, xcd_tiles);
  GUARD_END
}

int gocfd_ensemble_get_launch(gocfd_ensemble *h, int32_t *parallel, int32_t *graph_steps, int32_t *xcd_tiles) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (parallel) *parallel = h->parallel;
  if (graph_steps) *graph_steps = h->graph_steps;
  if (xcd_tiles) *xcd_tiles = h->xcd_tiles;
  return GOCFD_OK;
}

int gocfd_ensemble_tune(gocfd_ensemble *h, int32_t trial_blocks) {
  GUARD_BEGIN
  if (!h || trial_blocks <= 0) return fail(GOCFD_ERR_ARG, "bad argument");
  if (h->step_index != 0) return fail(GOCFD_ERR_ARG, "tune before the first step (tuning restarts from the initial state)");
  HIPCHECK(hipSetDevice(h->device));
  return h->tune(trial_blocks);
  GUARD_END
}

int gocfd_ensemble_kernel_time(gocfd_ensemble *h, double *ms, int32_t *steps) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (ms) *ms = h->last_ms;
  if (steps) *steps = h->last_steps;
  return GOCFD_OK;
}

}  // extern "C"
