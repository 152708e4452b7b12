// Forward-mode dual numbers (value + one tangent) for the tangent-linear and
// adjoint kernels. The pointwise physics (gas relations, numerical fluxes,
// the Riemann-invariant boundary state) is templated on its arithmetic type;
// instantiated with Dual it returns the directional derivative beside the
// value, which gives the tangent-linear step directly and, one input
// direction at a time, the pointwise Jacobians that the adjoint kernels
// apply transposed.
//
// The type and its functions live in their own namespace, so that ordinary
// name lookup in the solver's namespace keeps finding the double and float
// math functions; calls with Dual arguments find these by argument-dependent
// lookup.
#pragma once

#include <hip/hip_runtime.h>

namespace gocfd {
namespace ad {

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.), d(0.) {}
  __host__ __device__ Dual(double x) : v(x), d(0.) {}
  __host__ __device__ Dual(double x, double t) : v(x), d(t) {}
};

__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { a = a - b; return a; }
__device__ __forceinline__ Dual &operator*=(Dual &a, Dual b) { a = a * b; return a; }

// Comparisons take the values (branches follow the primal computation)
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__device__ __forceinline__ bool operator<(Dual a, double b) { return a.v < b; }
__device__ __forceinline__ bool operator<=(Dual a, double b) { return a.v <= b; }
__device__ __forceinline__ bool operator>(Dual a, double b) { return a.v > b; }
__device__ __forceinline__ bool operator>=(Dual a, double b) { return a.v >= b; }

// Piecewise functions take the derivative of the branch the value takes
// (fabs at 0 and ties of fmax/fmin: the first argument's branch).
__device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0. ? -a : a; }
__device__ __forceinline__ Dual fmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
__device__ __forceinline__ Dual fmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
__device__ __forceinline__ Dual copysign(Dual a, Dual b) {
  const double r = ::copysign(a.v, b.v);
  return Dual(r, (signbit(a.v) == signbit(b.v)) ? a.d : -a.d);
}
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2. * s));
}
__device__ __forceinline__ Dual exp(Dual a) {
  const double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
__device__ __forceinline__ Dual log10(Dual a) { return Dual(::log10(a.v), a.d / (a.v * 2.302585092994045684)); }
__device__ __forceinline__ Dual sin(Dual a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
__device__ __forceinline__ Dual cos(Dual a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
// x^p for a constant exponent: p x^(p-1) dx, written p (x^p / x) dx
__device__ __forceinline__ Dual pow(Dual a, double p) {
  const double r = ::pow(a.v, p);
  return Dual(r, p * (r / a.v) * a.d);
}
// The solver's division and square root helpers, in dual arithmetic
__device__ __forceinline__ Dual fdiv(Dual a, Dual b) { return a / b; }
__device__ __forceinline__ Dual fsqrt(Dual a) { return sqrt(a); }

// Value and tangent access that also works on plain doubles
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(Dual a) { return a.v; }
__device__ __forceinline__ double tan_of(double) { return 0.; }
__device__ __forceinline__ double tan_of(Dual a) { return a.d; }

}  // namespace ad
}  // namespace gocfd
