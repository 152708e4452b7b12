// Internal helpers shared by the C-ABI translation units (single solver and
// batched ensemble): error state, HIP checks, device buffers, freestream and
// initial states.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dfr_setup.hpp"
#include "gocfd_euler2d.h"
#include "kernels.hpp"

namespace gocfd {
namespace capi {

int fail(int code, const std::string &msg);
const char *last_error();

#define HIPCHECK(expr)                                                                     \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      throw ::gocfd::SetupError(GOCFD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GUARD_BEGIN try {
#define GUARD_END                                                                        \
  }                                                                                      \
  catch (const ::gocfd::SetupError &e) { return ::gocfd::capi::fail(e.code, e.what()); } \
  catch (const std::exception &e) { return ::gocfd::capi::fail(GOCFD_ERR_ARG, e.what()); }

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  void alloc(size_t count) {
    free();
    if (count == 0) count = 1;
    HIPCHECK(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  void upload(const std::vector<T> &v) {
    alloc(v.size());
    if (!v.empty()) HIPCHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { free(); }
};

// A host array of doubles on the device in the stage kernel's arithmetic
// type: FP32 for the FP32-bulk precision from N=2 on, else FP64
struct StageTypeBuf {
  DevBuf<double> d;
  DevBuf<float> f;
  void upload(const std::vector<double> &v, int prec, int N) {
    if (prec != 0 && N >= 2) f.upload(std::vector<float>(v.begin(), v.end()));
    else d.upload(v);
  }
  const void *ptr() const { return f.p ? static_cast<const void *>(f.p) : static_cast<const void *>(d.p); }
};

// Freestream (model_problems/Euler2D/fluids.go:245-273)
struct FreeStream {
  double gamma = 1.4, minf = 0, alpha = 0, q[4] = {1, 0, 0, 0}, pinf = 0, cinf = 0;
  void from_qinf(double g, const double qq[4]) {
    gamma = g;
    for (int i = 0; i < 4; ++i) q[i] = qq[i];
    const double oorho = 1. / q[0], u = q[1] * oorho, v = q[2] * oorho;
    pinf = (g - 1.) * (q[3] - 0.5 * q[0] * (u * u + v * v));
    cinf = std::sqrt(std::fabs(g * pinf * oorho));
  }
  void from_mach(double M, double g, double a) {
    const double ooggm1 = 1. / (g * (g - 1.));
    const double qq[4] = {1, M * std::cos(a * M_PI / 180.), M * std::sin(a * M_PI / 180.), ooggm1 + 0.5 * M * M};
    from_qinf(g, qq);
    minf = M;
    alpha = a;
  }
};

// The far-field states of a case: FSFar (NewEuler, euler.go:70) and the
// FSIn / FSOut of the shock tube (euler.go:742-743), which equal FSFar in
// the other cases
struct CaseStreams {
  FreeStream far, in, out;
};

// Parameter checks shared by every solver kind; returns GOCFD_OK or an error
int validate_params(const gocfd_params *p);

// InitializeSolution (euler.go:728-794): the freestreams, the initial state
// Q0 [4][NpInt][K] in the reference layout, and the IVortex boundary
// conversion of Wall edges (initialization.go:61-83).
void init_case(Discretisation &d, const gocfd_params &p, CaseStreams &fs, std::vector<double> &Q0);

// Device checks: gfx950 and enough LDS for the stage kernel
void check_device(int device, size_t lds_bytes);

// Physics constants of a case for the kernels
dev::Phys phys_of(const CaseStreams &fs, const gocfd_params &p);

// Element-major FP64 device layout of a reference-layout state, padding
// elements set to the inert state
void fill_element_major(const DeviceLayout &L, int NpInt, const double *Q, double *h);
void read_element_major(const DeviceLayout &L, int NpInt, const double *h, double *Q);

}  // namespace capi
}  // namespace gocfd

// The opaque mesh handle of the C ABI
struct gocfd_mesh {
  gocfd::Mesh m;
};
