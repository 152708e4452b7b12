// Shared C-ABI helpers (capi_common.hpp).
#include "capi_common.hpp"

#include <algorithm>

namespace gocfd {
namespace capi {

namespace {
thread_local std::string g_last_error;

// IVortex.GetStateC (isentropic_vortex/analytic_vortex.go:31-82), host copy for initialisation
void ivortex_host(double t, double x, double y, double q[4]) {
  const double beta = 5., X0 = 5., Y0 = 0., G = 1.4;
  const double oo2pi = 0.5 * (1. / M_PI), GM1 = G - 1, OOGM1 = 1. / GM1, fac = 16 * G * (M_PI * M_PI);
  double u = 1.0, v = 0.;
  const double xm = x - u * t, ym = y - v * t;
  const double r2 = (xm - X0) * (xm - X0) + (ym - Y0) * (ym - Y0);
  const double ex1r = std::exp(1 - r2);
  const double tv1 = 1.0 - (GM1 * beta * beta * std::exp(2.0 * (1.0 - r2)) / fac);
  u -= beta * ex1r * (ym - Y0) * oo2pi;
  v += beta * ex1r * (xm - X0) * oo2pi;
  const double rho = std::pow(tv1, OOGM1), p = std::pow(rho, G);
  q[0] = rho; q[1] = rho * u; q[2] = rho * v; q[3] = p * (1. / (G - 1.)) + 0.5 * rho * (u * u + v * v);
}
}  // namespace

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}
const char *last_error() { return g_last_error.c_str(); }

int validate_params(const gocfd_params *p) {
  if (p->polynomial_order < 0 || p->polynomial_order > 4)
    return fail(GOCFD_ERR_ARG, "polynomial order must be 0..4 on the GPU path");
  if (p->flux_type < 0 || p->flux_type > 3) return fail(GOCFD_ERR_ARG, "unable to use flux type");
  if (p->init_type != GOCFD_INIT_FREESTREAM && p->init_type != GOCFD_INIT_IVORTEX &&
      p->init_type != GOCFD_INIT_SHOCKTUBE)
    return fail(GOCFD_ERR_ARG, "unknown init type");
  if (p->limiter != 0 && p->limiter != 1) return fail(GOCFD_ERR_ARG, "unable to use limiter");
  if (p->precision != GOCFD_PREC_FP64 && p->precision != GOCFD_PREC_FP32)
    return fail(GOCFD_ERR_ARG, "unknown precision");
  return GOCFD_OK;
}

void init_case(Discretisation &d, const gocfd_params &p, CaseStreams &fs, std::vector<double> &Q0) {
  Q0.assign(size_t(4) * d.ops.NpInt * d.K, 0.0);
  const size_t npts = size_t(d.ops.NpInt) * d.K;
  if (p.init_type == GOCFD_INIT_IVORTEX) {
    const double qq[4] = {1, 1, 0, 3};
    fs.far.from_qinf(1.4, qq);
    fs.in = fs.out = fs.far;
    for (auto &E : d.edges)
      if (E.bc == BC_Wall) E.bc = BC_IVortex;
    for (size_t i = 0; i < npts; ++i) {
      double q[4];
      ivortex_host(0, d.SolX[i], d.SolY[i], q);
      for (int n = 0; n < 4; ++n) Q0[n * npts + i] = q[n];
    }
  } else if (p.init_type == GOCFD_INIT_SHOCKTUBE) {
    // Sod's states left and right of x = 0.5, which are also the In and Out
    // boundary states (euler.go:738-771)
    const double g = 1.4, qin[4] = {1, 0, 0, 1 / (g - 1)}, qout[4] = {0.125, 0, 0, 0.1 / (g - 1)};
    fs.far.from_mach(p.minf, p.gamma, p.alpha);
    fs.in.from_qinf(g, qin);
    fs.out.from_qinf(g, qout);
    for (size_t i = 0; i < npts; ++i) {
      const double *q = d.SolX[i] < 0.5 ? fs.in.q : fs.out.q;
      for (int n = 0; n < 4; ++n) Q0[n * npts + i] = q[n];
    }
  } else {
    fs.far.from_mach(p.minf, p.gamma, p.alpha);
    fs.in = fs.out = fs.far;
    for (int n = 0; n < 4; ++n)
      for (size_t i = 0; i < npts; ++i) Q0[n * npts + i] = 0.0 + fs.far.q[n];
  }
}

void check_device(int device, size_t lds) {
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) throw SetupError(GOCFD_ERR_DEVICE, "no such HIP device");
  HIPCHECK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw SetupError(GOCFD_ERR_DEVICE, std::string("device is not gfx950: ") + prop.gcnArchName);
  int lds_max = 0, lds_optin = 0;
  HIPCHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess)
    lds_max = std::max(lds_max, lds_optin);
  if (lds > size_t(lds_max))
    throw SetupError(GOCFD_ERR_SETUP, "stage kernel needs " + std::to_string(lds) + " B of LDS, device allows " +
                                          std::to_string(lds_max));
}

dev::Phys phys_of(const CaseStreams &fs, const gocfd_params &p) {
  dev::Phys ph{};
  ph.gamma = fs.far.gamma; ph.minf = fs.far.minf; ph.pinf = fs.far.pinf; ph.cinf = fs.far.cinf;
  for (int i = 0; i < 4; ++i) ph.qinf[i] = fs.far.q[i];
  const FreeStream *io[2] = {&fs.in, &fs.out};
  for (int j = 0; j < 2; ++j) {
    dev::FarState &s = ph.io[j];
    for (int i = 0; i < 4; ++i) s.q[i] = io[j]->q[i];
    s.gamma = io[j]->gamma; s.minf = io[j]->minf; s.pinf = io[j]->pinf; s.cinf = io[j]->cinf;
  }
  ph.cfl = p.cfl; ph.final_time = p.final_time; ph.flux = p.flux_type;
  return ph;
}

void fill_element_major(const DeviceLayout &L, int NpInt, const double *Q, double *h) {
  const double pad[4] = {1.0, 0.0, 0.0, 2.5};
  const int NV = 4 * NpInt, K = L.K;
  for (int kk = 0; kk < L.Kp; ++kk) {
    const int k = L.perm[kk];
    for (int n = 0; n < 4; ++n)
      for (int i = 0; i < NpInt; ++i)
        h[size_t(kk) * NV + n * NpInt + i] = k < 0 ? pad[n] : Q[size_t(n) * NpInt * K + size_t(i) * K + k];
  }
}

void read_element_major(const DeviceLayout &L, int NpInt, const double *h, double *Q) {
  const int NV = 4 * NpInt, K = L.K;
  for (int kk = 0; kk < L.K; ++kk) {
    const int k = L.perm[kk];
    for (int n = 0; n < 4; ++n)
      for (int i = 0; i < NpInt; ++i) Q[size_t(n) * NpInt * K + size_t(i) * K + k] = h[size_t(kk) * NV + n * NpInt + i];
  }
}

}  // namespace capi
}  // namespace gocfd
