// C ABI of the 2D Euler DFR solver (include/gocfd_euler2d.h): host
// orchestration of the device-resident solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "euler2d_handle.hpp"

extern "C" {

const char *gocfd_last_error(void) { return last_error(); }

int gocfd_mesh_read(const char *path, gocfd_mesh **out) {
  GUARD_BEGIN
  if (!path || !out) return fail(GOCFD_ERR_ARG, "null argument");
  auto m = std::make_unique<gocfd_mesh>();
  m->m = read_mesh(path);
  *out = m.release();
  return GOCFD_OK;
  GUARD_END
}

int gocfd_mesh_from_arrays(int32_t K, int32_t Nv, const double *VX, const double *VY, const int32_t *EToV,
                           int32_t ntags, const char *const *tags, const int32_t *edge_offset,
                           const int32_t *edge_count, const int32_t *edge_verts, gocfd_mesh **out) {
  GUARD_BEGIN
  if (K <= 0 || Nv <= 0 || !VX || !VY || !EToV || !out) return fail(GOCFD_ERR_ARG, "bad mesh arrays");
  auto m = std::make_unique<gocfd_mesh>();
  m->m.K = K;
  m->m.Nv = Nv;
  m->m.VX.assign(VX, VX + Nv);
  m->m.VY.assign(VY, VY + Nv);
  m->m.EToV.assign(EToV, EToV + size_t(3) * K);
  for (int t = 0; t < ntags; ++t) {
    std::vector<std::pair<int, int>> e(edge_count[t]);
    for (int j = 0; j < edge_count[t]; ++j)
      e[j] = {edge_verts[2 * (size_t(edge_offset[t]) + j)], edge_verts[2 * (size_t(edge_offset[t]) + j) + 1]};
    std::string tag(tags[t]);
    for (auto &c : tag) c = char(std::tolower((unsigned char)c));
    m->m.add_tag(tag, std::move(e));
  }
  *out = m.release();
  return GOCFD_OK;
  GUARD_END
}

int gocfd_mesh_dims(const gocfd_mesh *m, int32_t *K, int32_t *Nv, int32_t *ntags) {
  if (!m) return fail(GOCFD_ERR_ARG, "null mesh");
  if (K) *K = m->m.K;
  if (Nv) *Nv = m->m.Nv;
  if (ntags) *ntags = int32_t(m->m.tags.size());
  return GOCFD_OK;
}

int gocfd_mesh_arrays(const gocfd_mesh *m, double *VX, double *VY, int32_t *EToV) {
  if (!m) return fail(GOCFD_ERR_ARG, "null mesh");
  if (VX) std::memcpy(VX, m->m.VX.data(), sizeof(double) * m->m.Nv);
  if (VY) std::memcpy(VY, m->m.VY.data(), sizeof(double) * m->m.Nv);
  if (EToV) for (size_t i = 0; i < m->m.EToV.size(); ++i) EToV[i] = m->m.EToV[i];
  return GOCFD_OK;
}

void gocfd_mesh_free(gocfd_mesh *m) { delete m; }

int gocfd_euler2d_create(const gocfd_mesh *mesh, const gocfd_params *p, gocfd_euler2d **out) {
  GUARD_BEGIN
  if (!mesh || !p || !out) return fail(GOCFD_ERR_ARG, "null argument");
  *out = nullptr;
  if (validate_params(p) != GOCFD_OK) return GOCFD_ERR_ARG;
  auto h = std::make_unique<gocfd_euler2d>();
  h->prm = *p;
  h->lts = p->local_time_stepping != 0;
  // NewScalarDissipation (dissipation.go:111-165) for order >= 1; Kappa
  // replaces the default 5 in the sensor ramp, the viscosity scale stays 5/1.5
  h->shock = p->limiter == 1 && p->polynomial_order != 0;
  h->kappa = p->kappa != 0. ? p->kappa : 5.;
  h->eps0 = 5. / 1.5;
  h->prec = p->precision;
  h->ssize = h->prec == 0 ? sizeof(double) : sizeof(float);
  h->d = build_discretisation(mesh->m, p->polynomial_order);
  std::vector<double> Q0;
  init_case(h->d, *p, h->fs, Q0);
  h->L = build_device_layout(h->d, dev::kTile, dev::kTP);
  h->NV = 4 * h->d.ops.NpInt;
  if (p->device < 0) {  // host-only handle: setup and introspection, no device work
    h->device = -1;
    *out = h.release();
    return GOCFD_OK;
  }
  // Device
  check_device(p->device, dev::stage_lds_bytes(p->polynomial_order, h->prec, h->L.Hmax) + 4096);
  h->device = p->device;
  HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&h->ev0));
  HIPCHECK(hipEventCreate(&h->ev1));
  const size_t nval = size_t(h->L.Kp) * h->NV;
  HIPCHECK(hipMalloc(&h->q[0], nval * sizeof(double)));
  for (int s = 1; s < 5; ++s) {
    HIPCHECK(hipMalloc(&h->q[s], nval * h->ssize));
    HIPCHECK(hipMemset(h->q[s], 0, nval * h->ssize));
  }
  HIPCHECK(hipMalloc(&h->R, nval * sizeof(double)));
  HIPCHECK(hipMemset(h->R, 0, nval * sizeof(double)));
  HIPCHECK(hipMalloc(&h->ctl, sizeof(dev::Ctl)));
  auto &L = h->L;
  h->Jdet.upload(L.Jdet); h->Jinv.upload(L.Jinv); h->IInII.upload(L.IInII);
  h->Nx.upload(L.Nx); h->Ny.upload(L.Ny); h->oohK.upload(L.oohK);
  h->bxy.upload(L.bpoint_xy);
  h->code.upload(L.face_code); h->nbr_slot.upload(L.nbr_slot);
  h->halo_elem.upload(L.halo_elem); h->halo_face.upload(L.halo_face);
  h->wave_list.upload(L.wave_list); h->wave_cnt.upload(L.wave_cnt); h->wave_code.upload(L.wave_code);
  h->flux_list.upload(L.flux_list); h->flux_cnt.upload(L.flux_cnt); h->bnd_cnt.upload(L.bnd_cnt);
  h->flux_geo.upload(L.flux_geo, h->prec, p->polynomial_order);
  h->perm.upload(L.perm);
  if (!h->shock) {
    h->xp_list.upload(L.xp_list); h->xp_dst.upload(L.xp_dst); h->xp_cnt.upload(L.xp_cnt);
    h->ht.alloc(size_t(2) * L.ntiles * L.Fmax * (h->d.N + 2) * 4);  // FP64 size, enough for either arithmetic
  }
  noalloc_setup(h->device);
  {
    dev::CaseInfo ci{};
    ci.ph = phys_of(h->fs, h->prm);
    ci.k_end = L.K;
    ci.tile0 = 0;
    h->cases.upload(std::vector<dev::CaseInfo>{ci});
    h->tile_case.upload(std::vector<int>(L.ntiles, 0));
  }
  if (h->lts) {
    h->fw.alloc(size_t(2) * 3 * L.Kp);
    h->fwv.alloc(size_t(2) * 3 * L.Kp);
    h->DTk.alloc(L.Kp);
    h->DTv.alloc(L.Kp);
    h->wsrc.upload(L.wsrc);
  }
  if (h->shock) {
    const int NR = h->d.ops.NpInt + 3 * h->d.ops.NpEdge;
    const std::vector<double> sops = build_shock_operators(h->d.ops);
    h->sh_ops.upload(sops);
    h->sh_opsF.upload(std::vector<float>(sops.begin(), sops.end()));
    h->evert.upload(L.evert); h->hK.upload(L.hK);
    h->nbr_elem.upload(L.nbr_elem); h->nbr_face.upload(L.nbr_face);
    h->Nxo.upload(L.Nxo); h->Nyo.upload(L.Nyo); h->vooL.upload(L.vooL);
    h->sigV.alloc(size_t(2) * std::max(h->d.Nv, 1));
    h->epsV.alloc(size_t(2) * std::max(h->d.Nv, 1));
    h->Se.alloc(L.Kp); h->sigma.alloc(L.Kp); h->eps.alloc(L.Kp);
    h->epsF.alloc(size_t(L.Kp) * NR);
    h->diss.alloc(size_t(L.Kp) * 4 * 2 * NR);  // FP64 or FP32 values (the larger size)
    h->active.alloc(L.Kp);
    h->act_list.alloc(L.Kp);
    h->act_cnt.alloc(1);
  }
  std::vector<double> divT(h->d.ops.DivInt.size());  // [NpFlux][NpInt] for contiguous per-DOF rows
  for (int i = 0; i < h->d.ops.NpInt; ++i)
    for (int j = 0; j < h->d.ops.NpFlux; ++j)
      divT[size_t(j) * h->d.ops.NpInt + i] = h->d.ops.DivInt[size_t(i) * h->d.ops.NpFlux + j];
  HIPCHECK(dev::upload_operators(p->polynomial_order, divT.data(), h->d.ops.Interp.data()));
  h->Interp64.upload(h->d.ops.Interp);
  h->to_device(Q0.data(), h->q[0]);
  h->restart(0.0);
  HIPCHECK(hipStreamSynchronize(h->stream));
  *out = h.release();
  return GOCFD_OK;
  GUARD_END
}

void gocfd_euler2d_destroy(gocfd_euler2d *h) { delete h; }

int gocfd_euler2d_dims(const gocfd_euler2d *h, int32_t *K, int32_t *NpInt, int32_t *NpEdge, int32_t *NpFlux,
                       int32_t *nedges) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (K) *K = h->d.K;
  if (NpInt) *NpInt = h->d.ops.NpInt;
  if (NpEdge) *NpEdge = h->d.ops.NpEdge;
  if (NpFlux) *NpFlux = h->d.ops.NpFlux;
  if (nedges) *nedges = int32_t(h->d.edges.size());
  return GOCFD_OK;
}

int gocfd_euler2d_set_state(gocfd_euler2d *h, const double *Q) {
  GUARD_BEGIN
  if (!h || !Q) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  dev::Ctl c = h->read_ctl();
  h->steps_base += c.steps;
  h->to_device(Q, h->q[0]);
  h->restart(c.time[c.steps & 1]);
  HIPCHECK(hipStreamSynchronize(h->stream));
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_get_state(gocfd_euler2d *h, double *Q) {
  GUARD_BEGIN
  if (!h || !Q) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->from_device(h->q[0], Q);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_get_residual(gocfd_euler2d *h, double *R) {
  GUARD_BEGIN
  if (!h || !R) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->from_device(h->R, R);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_rhs(gocfd_euler2d *h, const double *Q, double t, double *RHS) {
  GUARD_BEGIN
  if (!h || !Q || !RHS) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  const size_t tot = size_t(h->L.Kp) * h->NV;
  if (!h->scratch_in) HIPCHECK(hipMalloc(&h->scratch_in, tot * sizeof(double)));
  if (h->rhs_out.n < tot) h->rhs_out.alloc(tot);
  h->to_device(Q, h->scratch_in);
  dev::StageArgs a = h->rhs_args(h->scratch_in, t, h->rhs_out.p, h->stream);
  h->launch_one_stage(a);
  if (h->shock) {  // the vertex buffer this evaluation used must be clear for the next stage
    HIPCHECK(dev::launch_fill_u64(h->sigV.p, h->sigV.n, 0ull, h->stream));
    HIPCHECK(dev::launch_fill_u64(h->epsV.p, h->epsV.n, 0ull, h->stream));
  }
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->from_device(h->rhs_out.p, RHS);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_step_async(gocfd_euler2d *h, int32_t nsteps) {
  GUARD_BEGIN
  if (!h || nsteps < 0) return fail(GOCFD_ERR_ARG, "bad argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  h->enqueue_steps(nsteps);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_sync(gocfd_euler2d *h) {
  GUARD_BEGIN
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  return h->finish_batch();
  GUARD_END
}

int gocfd_euler2d_step(gocfd_euler2d *h, int32_t nsteps) {
  GUARD_BEGIN
  if (!h || nsteps < 0) return fail(GOCFD_ERR_ARG, "bad argument");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  h->enqueue_steps(nsteps);
  return h->finish_batch();
  GUARD_END
}

int gocfd_euler2d_time(gocfd_euler2d *h, double *time, double *global_dt, int32_t *steps) {
  GUARD_BEGIN
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  dev::Ctl c = h->read_ctl();
  const int par = c.steps & 1;
  if (time) *time = c.time[par];
  if (global_dt) *global_dt = c.dt[4];
  if (steps) *steps = h->steps_base + c.steps;
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_set_time(gocfd_euler2d *h, double time) {
  GUARD_BEGIN
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  NEED_DEVICE(h);
  HIPCHECK(hipSetDevice(h->device));
  dev::Ctl c = h->read_ctl();
  h->steps_base += c.steps;
  h->restart(time);
  HIPCHECK(hipStreamSynchronize(h->stream));
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_get_edges(gocfd_euler2d *h, int64_t *out) {
  if (!h || !out) return fail(GOCFD_ERR_ARG, "null argument");
  for (size_t i = 0; i < h->d.edges.size(); ++i) {
    const Edge &E = h->d.edges[i];
    int64_t *r = out + 10 * i;
    r[0] = int64_t(E.key & 0xffffffffu); r[1] = int64_t(E.key >> 32); r[2] = E.nconn;
    r[3] = E.tri[0]; r[4] = E.nconn > 1 ? E.tri[1] : -1;
    r[5] = E.enr[0]; r[6] = E.nconn > 1 ? E.enr[1] : -1;
    r[7] = E.dir[0]; r[8] = E.nconn > 1 ? E.dir[1] : -1; r[9] = E.bc;
  }
  return GOCFD_OK;
}

int gocfd_euler2d_get_connectivity(gocfd_euler2d *h, int32_t *EtoE, int32_t *FaceEdge) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (EtoE) for (size_t i = 0; i < h->d.EtoE.size(); ++i) EtoE[i] = h->d.EtoE[i];
  if (FaceEdge) for (size_t i = 0; i < h->d.FaceEdge.size(); ++i) FaceEdge[i] = h->d.FaceEdge[i];
  return GOCFD_OK;
}

int gocfd_euler2d_get_operators(gocfd_euler2d *h, double *DivInt, double *Div, double *Interp, double *R, double *S) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  const auto &o = h->d.ops;
  if (DivInt) std::memcpy(DivInt, o.DivInt.data(), o.DivInt.size() * sizeof(double));
  if (Div) std::memcpy(Div, o.Div.data(), o.Div.size() * sizeof(double));
  if (Interp) std::memcpy(Interp, o.Interp.data(), o.Interp.size() * sizeof(double));
  if (R) std::memcpy(R, o.R.data(), o.R.size() * sizeof(double));
  if (S) std::memcpy(S, o.S.data(), o.S.size() * sizeof(double));
  return GOCFD_OK;
}

int gocfd_euler2d_get_geometry(gocfd_euler2d *h, double *Jdet, double *Jinv, double *FaceNx, double *FaceNy,
                               double *IInII, double *EdgeLenMax) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  const auto &d = h->d;
  auto cp = [](double *dst, const std::vector<double> &v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  cp(Jdet, d.Jdet); cp(Jinv, d.Jinv); cp(FaceNx, d.FaceNx); cp(FaceNy, d.FaceNy);
  cp(IInII, d.IInII); cp(EdgeLenMax, d.EdgeLenMax);
  return GOCFD_OK;
}

int gocfd_euler2d_get_coords(gocfd_euler2d *h, double *SolX, double *SolY, double *FluxX, double *FluxY) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  const auto &d = h->d;
  auto cp = [](double *dst, const std::vector<double> &v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  cp(SolX, d.SolX); cp(SolY, d.SolY); cp(FluxX, d.FluxX); cp(FluxY, d.FluxY);
  return GOCFD_OK;
}

int gocfd_euler2d_interp_matrix(gocfd_euler2d *h, int32_t npts, const double *r, const double *s, double *out) {
  GUARD_BEGIN
  if (!h || npts < 0 || (npts > 0 && (!r || !s || !out))) return fail(GOCFD_ERR_ARG, "bad argument");
  const std::vector<double> M = interp_matrix(h->d.ops, npts, r, s);
  std::memcpy(out, M.data(), M.size() * sizeof(double));
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_get_shock(gocfd_euler2d *h, double *Se, double *Sigma, double *Eps) {
  GUARD_BEGIN
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  NEED_DEVICE(h);
  if (!h->shock) return fail(GOCFD_ERR_ARG, "the solver was created without the PerssonC0 limiter");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  const double *src[3] = {h->Se.p, h->sigma.p, h->eps.p};
  double *dst[3] = {Se, Sigma, Eps};
  std::vector<double> tmp(h->L.Kp);
  for (int i = 0; i < 3; ++i) {
    if (!dst[i]) continue;
    HIPCHECK(hipMemcpy(tmp.data(), src[i], tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int kk = 0; kk < h->L.K; ++kk) dst[i][h->L.perm[kk]] = tmp[kk];
  }
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_get_element_dt(gocfd_euler2d *h, double *DT) {
  GUARD_BEGIN
  if (!h || !DT) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  if (!h->lts) return fail(GOCFD_ERR_ARG, "the solver was created without LocalTimeStepping");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<double> tmp(h->L.Kp);
  HIPCHECK(hipMemcpy(tmp.data(), h->DTk.p, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int kk = 0; kk < h->L.K; ++kk) DT[h->L.perm[kk]] = tmp[kk];
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_vjp(gocfd_euler2d *h, int32_t nsteps, int32_t checkpoint_every, const double *dQn, double *dQ0,
                      double *dQinf) {
  GUARD_BEGIN
  if (!h || !dQn || !dQ0) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  if (nsteps < 1 || checkpoint_every < 0) return fail(GOCFD_ERR_ARG, "nsteps must be positive, checkpoint_every >= 0");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  capi::AdjointRunner &r = h->adjoint();
  const double t0 = h->current_time();
  const size_t tot = size_t(h->L.Kp) * h->NV;
  DevBuf<double> lamN, lam0;
  lamN.alloc(tot);
  lam0.alloc(tot);
  h->to_device_zero_pad(dQn, lamN.p);
  double far[4];
  r.vjp(static_cast<const double *>(h->q[0]), t0, nsteps, checkpoint_every, lamN.p, lam0.p, far);
  h->from_device(lam0.p, dQ0);
  if (dQinf) for (int c = 0; c < 4; ++c) dQinf[c] = far[c];
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_jvp(gocfd_euler2d *h, int32_t nsteps, const double *dQ0, const double *dQinf, double *Qn,
                      double *dQn) {
  GUARD_BEGIN
  if (!h || !dQ0 || !dQn) return fail(GOCFD_ERR_ARG, "null argument");
  NEED_DEVICE(h);
  if (nsteps < 1) return fail(GOCFD_ERR_ARG, "nsteps must be positive");
  HIPCHECK(hipSetDevice(h->device));
  HIPCHECK(hipStreamSynchronize(h->stream));
  capi::AdjointRunner &r = h->adjoint();
  const double t0 = h->current_time();
  const size_t tot = size_t(h->L.Kp) * h->NV;
  DevBuf<double> dq0, qn, dqn;
  dq0.alloc(tot);
  qn.alloc(tot);
  dqn.alloc(tot);
  h->to_device_zero_pad(dQ0, dq0.p);
  r.jvp(static_cast<const double *>(h->q[0]), dq0.p, dQinf, t0, nsteps, qn.p, dqn.p);
  if (Qn) h->from_device(qn.p, Qn);
  h->from_device(dqn.p, dQn);
  return GOCFD_OK;
  GUARD_END
}

int gocfd_euler2d_kernel_time(gocfd_euler2d *h, double *avg_stage_ms, int32_t *launches) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (avg_stage_ms)
    *avg_stage_ms = h->last_batch_launches ? h->last_batch_ms / h->last_batch_launches : 0.0;
  if (launches) *launches = h->last_batch_launches;
  return GOCFD_OK;
}

int gocfd_euler2d_tiling(gocfd_euler2d *h, int32_t *tile, int32_t *ntiles, int32_t *max_halo) {
  if (!h) return fail(GOCFD_ERR_ARG, "null handle");
  if (tile) *tile = h->L.E;
  if (ntiles) *ntiles = h->L.ntiles;
  if (max_halo) *max_halo = h->L.Hmax;
  return GOCFD_OK;
}

}  // extern "C"
