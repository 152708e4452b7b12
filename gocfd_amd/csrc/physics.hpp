// Pointwise Euler physics shared by the device kernels: gas relations,
// physical flux, numerical fluxes, Riemann-invariant boundary state and the
// analytic isentropic vortex. Templated on the arithmetic type C (double for
// the FP64 path, float for the bulk of the FP32 path).
#pragma once

#include <hip/hip_runtime.h>

namespace gocfd {
namespace dev {

// Division and square root in the bulk arithmetic type: IEEE in FP64 (the
// drop-in path); in FP32 (the FP32-bulk path) the hardware reciprocal and
// square root (v_rcp_f32, v_sqrt_f32, about 1 ulp) instead of the
// correctly rounded sequences.
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// Reciprocal for the reciprocal form of the Roe flux: in FP64 the hardware
// estimate (v_rcp_f64) refined by two Newton steps, within an ulp of 1/x for
// the normal-range operands it gets (densities, sums of their square roots,
// sound speeds), in a third of the IEEE division's instructions; in FP32 the
// hardware reciprocal.
__device__ __forceinline__ double frcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// GetFlowFunctionBase StaticPressure (model_problems/Euler2D/fluids.go:289-336)
template <class C>
__device__ __forceinline__ C pressure(C g, C rho, C rhoU, C rhoV, C E) {
  const C oorho = fdiv(C(1), rho);
  const C u = rhoU * oorho, v = rhoV * oorho;
  const C q = C(0.5) * rho * (u * u + v * v);
  return (g - C(1)) * (E - q);
}

// FluxCalcBase (model_problems/Euler2D/fluxes.go:76-87)
template <class C>
__device__ __forceinline__ void flux_xy(C g, const C q[4], C Fx[4], C Fy[4]) {
  const C oorho = fdiv(C(1), q[0]);
  const C u = q[1] * oorho, v = q[2] * oorho;
  const C p = pressure(g, q[0], q[1], q[2], q[3]);
  Fx[0] = q[1]; Fx[1] = q[1] * u + p; Fx[2] = q[1] * v; Fx[3] = u * (q[3] + p);
  Fy[0] = q[2]; Fy[1] = q[2] * u; Fy[2] = q[2] * v + p; Fy[3] = v * (q[3] + p);
}

// One component n of the physical flux (n is wave-uniform in the kernels)
template <class C>
__device__ __forceinline__ void flux_component(int n, C g, C rho, C rhoU, C rhoV, C E, C &Fx, C &Fy) {
  const C oorho = fdiv(C(1), rho);
  const C u = rhoU * oorho, v = rhoV * oorho;
  const C p = (g - C(1)) * (E - C(0.5) * rho * (u * u + v * v));
  switch (n) {
    case 0: Fx = rhoU; Fy = rhoV; break;
    case 1: Fx = rhoU * u + p; Fy = rhoV * u; break;
    case 2: Fx = rhoU * v; Fy = rhoV * v + p; break;
    default: Fx = u * (E + p); Fy = v * (E + p); break;
  }
}

template <class C>
__device__ __forceinline__ C sound_speed(C g, const C q[4]) {
  const C p = pressure(g, q[0], q[1], q[2], q[3]);
  return fsqrt(fabs(g * p * (fdiv(C(1), q[0]))));
}

// AvgFlux (fluxes.go:135-159)
template <class C>
__device__ __forceinline__ void avg_flux(C g, const C L[4], const C R[4], C nx, C ny, C F[4]) {
  C FxL[4], FyL[4], FxR[4], FyR[4];
  flux_xy(g, L, FxL, FyL);
  flux_xy(g, R, FxR, FyR);
#pragma unroll
  for (int c = 0; c < 4; ++c) F[c] = nx * (C(0.5) * (FxL[c] + FxR[c])) + ny * (C(0.5) * (FyL[c] + FyR[c]));
}

// LaxFlux (fluxes.go:161-190): dissipation uses max(|V|+c), not the normal speed
template <class C>
__device__ __forceinline__ void lax_flux(C g, const C L[4], const C R[4], C nx, C ny, C F[4]) {
  const C uL = fdiv(L[1], L[0]), vL = fdiv(L[2], L[0]), uR = fdiv(R[1], R[0]), vR = fdiv(R[2], R[0]);
  const C pL = pressure(g, L[0], L[1], L[2], L[3]), pR = pressure(g, R[0], R[1], R[2], R[3]);
  const C CL = sound_speed(g, L), CR = sound_speed(g, R);
  const C maxV = fmax(fsqrt(uL * uL + vL * vL) + CL, fsqrt(uR * uR + vR * vR) + CR);
  F[0] = C(0.5) * (nx * (L[1] + R[1]) + ny * (L[2] + R[2]));
  F[1] = C(0.5) * (nx * (L[1] * uL + R[1] * uR + pL + pR) + ny * (L[1] * vL + R[1] * vR));
  F[2] = C(0.5) * (nx * (L[2] * uL + R[2] * uR) + ny * (L[2] * vL + R[2] * vR + pL + pR));
  F[3] = C(0.5) * (nx * ((pL + L[3]) * uL + (pR + R[3]) * uR) + ny * ((pL + L[3]) * vL + (pR + R[3]) * vR));
#pragma unroll
  for (int c = 0; c < 4; ++c) F[c] += C(0.5) * maxV * (L[c] - R[c]);
}

// RoeFlux (fluxes.go:284-413): normal frame, Roe averages, no entropy fix.
// With RCP (the stage kernel without shock capturing) each division by rho,
// by the sum of the square roots, by c and by c^2 is a multiplication by one
// reciprocal (frcp), and the static pressures come from the rotated
// velocities: within about an ulp per operation of the divisions, with no
// IEEE division sequence left.
template <class C, bool RCP = false>
__device__ __forceinline__ void roe_flux(C g, const C L[4], const C R[4], C nx, C ny, C F[4]) {
  const C GM1 = g - C(1);
  const C rULr = L[1] * nx + L[2] * ny, rVLr = L[1] * (-ny) + L[2] * nx;
  const C rURr = R[1] * nx + R[2] * ny, rVRr = R[1] * (-ny) + R[2] * nx;
  const C rhoL = L[0], rhoR = R[0];
  const C sL = fsqrt(rhoL), sR = fsqrt(rhoR), sLR = sL + sR;
  const C rho = sL * sR;
  C uL, vL, uR, vR, pL, pR, u, v, h, c2, cc, dW1, dW2, dW4;
  if constexpr (RCP) {
    const C ooL = frcp(rhoL), ooR = frcp(rhoR), ooS = frcp(sLR);
    uL = rULr * ooL; vL = rVLr * ooL; uR = rURr * ooR; vR = rVRr * ooR;
    // the static pressure from the rotated velocities (|u| is invariant)
    pL = GM1 * (L[3] - C(0.5) * rhoL * (uL * uL + vL * vL));
    pR = GM1 * (R[3] - C(0.5) * rhoR * (uR * uR + vR * vR));
    const C hL = (L[3] + pL) * ooL, hR = (R[3] + pR) * ooR;
    u = (sL * uL + sR * uR) * ooS;
    v = (sL * vL + sR * vR) * ooS;
    h = (sL * hL + sR * hR) * ooS;
    c2 = GM1 * (h - C(0.5) * (u * u + v * v));
    cc = fsqrt(c2);
    const C ooc = frcp(cc), ooc2 = ooc * ooc;
    dW1 = -C(0.5) * (rho * (uR - uL)) * ooc + C(0.5) * (pR - pL) * ooc2;
    dW2 = (rhoR - rhoL) - (pR - pL) * ooc2;
    dW4 = C(0.5) * (rho * (uR - uL)) * ooc + C(0.5) * (pR - pL) * ooc2;
  } else {
    pL = pressure(g, L[0], L[1], L[2], L[3]);
    pR = pressure(g, R[0], R[1], R[2], R[3]);
    uL = fdiv(rULr, L[0]); vL = fdiv(rVLr, L[0]); uR = fdiv(rURr, R[0]); vR = fdiv(rVRr, R[0]);
    const C hL = fdiv(L[3] + pL, rhoL), hR = fdiv(R[3] + pR, rhoR);
    u = fdiv(sL * uL + sR * uR, sLR);
    v = fdiv(sL * vL + sR * vR, sLR);
    h = fdiv(sL * hL + sR * hR, sLR);
    c2 = GM1 * (h - C(0.5) * (u * u + v * v));
    cc = fsqrt(c2);
    dW1 = fdiv(-C(0.5) * (rho * (uR - uL)), cc) + fdiv(C(0.5) * (pR - pL), c2);
    dW2 = (rhoR - rhoL) - fdiv(pR - pL, c2);
    dW4 = fdiv(C(0.5) * (rho * (uR - uL)), cc) + fdiv(C(0.5) * (pR - pL), c2);
  }
  C dW3 = rho * (vR - vL);
  dW1 = fabs(u - cc) * dW1;
  dW2 = fabs(u) * dW2;
  dW3 = fabs(u) * dW3;
  dW4 = fabs(u + cc) * dW4;
  C f0 = C(0.5) * (rULr + rURr);
  C f1 = C(0.5) * (rULr * uL + rURr * uR + pL + pR);
  C f2 = C(0.5) * (rVLr * uL + rVRr * uR);
  C f3 = C(0.5) * ((pL + L[3]) * uL + (pR + R[3]) * uR);
  f0 -= C(0.5) * (dW1 + dW2 + dW4);
  f1 -= C(0.5) * (dW1 * (u - cc) + dW2 * u + dW4 * (u + cc));
  f2 -= C(0.5) * (dW1 * v + dW2 * v + dW3 + dW4 * v);
  f3 -= C(0.5) * (dW1 * (h - u * cc) + C(0.5) * dW2 * (u * u + v * v) + dW3 * v + dW4 * (h + u * cc));
  F[0] = f0;
  F[1] = nx * f1 - ny * f2;
  F[2] = ny * f1 + nx * f2;
  F[3] = f3;
}

// RoeERFlux (fluxes.go:415-503), energy dissipation term as in the reference
template <class C>
__device__ __forceinline__ void roeer_flux(C g, const C L[4], const C R[4], C nx, C ny, C F[4]) {
  const C GM1 = g - C(1);
  const C rhoL = L[0], rhoR = R[0];
  const C ooL = fdiv(C(1), rhoL), ooR = fdiv(C(1), rhoR);
  const C sL = fsqrt(rhoL), sR = fsqrt(rhoR);
  const C uL = L[1] * ooL, vL = L[2] * ooL, uR = R[1] * ooR, vR = R[2] * ooR;
  const C EL = L[3], ER = R[3];
  const C UL = nx * uL + ny * vL, UR = nx * uR + ny * vR;
  const C pL = pressure(g, L[0], L[1], L[2], L[3]), pR = pressure(g, R[0], R[1], R[2], R[3]);
  const C HL = EL + pL, HR = ER + pR;
  const C ooRs = fdiv(C(1), sL + sR);
  const C u = (sL * uL + sR * uR) * ooRs, v = (sL * vL + sR * vR) * ooRs, h = (HL + HR) * ooRs;
  const C rho = sL * sR, H = h * rho, U = nx * u + ny * v;
  const C C2 = GM1 * (h - C(0.5) * (u * u + v * v));
  const C cc = fsqrt(C2), ooC = fdiv(C(1), cc), Uabs = fabs(U);
  F[0] = C(0.5) * (UL * rhoL + UR * rhoR);
  F[1] = C(0.5) * (UL * rhoL * uL + pL * nx + UR * rhoR * uR + pR * nx);
  F[2] = C(0.5) * (UL * rhoL * vL + pL * ny + UR * rhoR * vR + pR * ny);
  F[3] = C(0.5) * (UL * HL + UR * HR);
  const C Uef = C(0.05) * cc;
  const C du = uR - uL, dv = vR - vL, dV2 = du * du + dv * dv;
  const C ooVmag = fdiv(C(1), fsqrt(u * u + v * v));
  C n1x, n1y;
  if (dV2 < C(0.01) * C2) { n1x = nx; n1y = ny; } else { n1x = ooVmag * du; n1y = ooVmag * dv; }
  const C n2x = n1y * (nx * n1y - n1x * ny), n2y = -n1x * (nx * n1y - n1x * ny);
  const C alp1 = nx * n1x + ny * n1y, alp2 = nx * n2x + ny * n2y;
  const C U1x = n1x * u, U1y = n1y * v, U2x = n2x * u, U2y = n2y * v;
  const C Urot = fsqrt(alp1 * alp1 * (U1x * U1x + U1y * U1y)) + fsqrt(alp2 * alp2 * (U2x * U2x + U2y * U2y));
  const C sigma = fmax(Uabs, fmin(Uef, Urot));
  // The reference's sign = math.Copysign: sign(x, 1) has the magnitude of x
  // and a positive sign, i.e. |x| (fluxes.go:426,491,494)
  const C UabsP = Uabs - C(0.25) * fmax(C(0), UR - UL) * (fabs(U + cc) - fabs(U - cc));
  const C dU = UR - UL, dP = pR - pL, dRho = rhoR - rhoL, dRU = rhoR * uR - rhoL * uL;
  const C dRV = rhoR * vR - rhoL * vL, dE = ER - EL;
  const C dPu = rho * dU * fmax(C(0), cc - UabsP);
  const C swt = Uabs * fmin(UabsP, cc);
  const C dPp = swt * dP * ooC, dUu = swt * dU * ooC;
  F[0] -= C(0.5) * (sigma * dRho + (dPu + dPp) * C(0) + dUu * rho);
  F[1] -= C(0.5) * (sigma * dRU + (dPu + dPp) * nx + dUu * rho * u);
  F[2] -= C(0.5) * (sigma * dRV + (dPu + dPp) * ny + dUu * rho * v);
  F[3] -= C(0.5) * (sigma * dE + (dPu + dPp) * ny + dUu * H);
}

template <class C, bool RCP = false>
__device__ __forceinline__ void numflux(int kind, C g, const C L[4], const C R[4], C nx, C ny, C F[4]) {
  switch (kind) {
    case 0: avg_flux(g, L, R, nx, ny, F); break;
    case 1: lax_flux(g, L, R, nx, ny, F); break;
    case 2: roe_flux<C, RCP>(g, L, R, nx, ny, F); break;
    default: roeer_flux(g, L, R, nx, ny, F); break;
  }
}

// RiemannBC (model_problems/Euler2D/bcs.go:70-133), in FP64 (T = double) or
// in FP64 dual arithmetic for the tangent-linear and adjoint kernels.
// pinf/cinf come from the freestream even for the analytic vortex state.
template <class T, class P>
__device__ __forceinline__ void riemann_bc(double g, double minf, P pinf, P cinf, const T qi[4],
                                           const T qinf[4], double nx, double ny, T qb[4]) {
  const T rhoI = qi[0], uI = qi[1] / qi[0], vI = qi[2] / qi[0];
  const T pI = pressure(T(g), qi[0], qi[1], qi[2], qi[3]);
  const T cI = sound_speed(T(g), qi);
  const double GM1 = g - 1., OOGM1 = 1. / GM1;
  const T rhoF = qinf[0], uF = qinf[1] / qinf[0], vF = qinf[2] / qinf[0];
  const double tx = -ny, ty = nx;
  const T VnI = nx * uI + ny * vI;
  if (minf <= 1.) {
    const T VnF = nx * uF + ny * vF;
    const T Rinf = VnF - 2. * cinf * OOGM1;
    const T Rint = VnI + 2. * cI * OOGM1;
    const T Vn = 0.5 * (Rint + Rinf);
    const T c = 0.25 * GM1 * (Rint - Rinf);
    T Vt, beta;
    if (VnI < 0) { Vt = tx * uF + ty * vF; beta = pinf / pow(rhoF, g); }
    else { Vt = tx * uI + ty * vI; beta = pI / pow(rhoI, g); }
    const T u = Vn * nx + Vt * tx, v = Vn * ny + Vt * ty;
    const T rho = pow(c * c / (g * beta), OOGM1);
    const T p = beta * pow(rho, g);
    qb[0] = rho; qb[1] = rho * u; qb[2] = rho * v; qb[3] = p * OOGM1 + 0.5 * rho * (u * u + v * v);
  } else if (VnI < 0) {
    for (int c = 0; c < 4; ++c) qb[c] = qinf[c];
  } else {
    for (int c = 0; c < 4; ++c) qb[c] = qi[c];
  }
}

// IVortex.GetStateC with beta=5, centre (5,0), gamma=1.4, Ufs=1
// (isentropic_vortex/analytic_vortex.go:31-82, initialization.go:61-71)
__device__ __forceinline__ void ivortex_state(double t, double x, double y, double q[4]) {
  const double beta = 5., X0 = 5., Y0 = 0., G = 1.4;
  const double oo2pi = 0.5 * (1. / M_PI), GM1 = G - 1, OOGM1 = 1. / GM1;
  const double fac = 16 * G * (M_PI * M_PI);
  double u = 1.0, v = 0.;
  const double xm = x - u * t, ym = y - v * t;
  const double r2 = (xm - X0) * (xm - X0) + (ym - Y0) * (ym - Y0);
  const double ex1r = exp(1 - r2);
  const double tv1 = 1.0 - (GM1 * beta * beta * exp(2.0 * (1.0 - r2)) / fac);
  u -= beta * ex1r * (ym - Y0) * oo2pi;
  v += beta * ex1r * (xm - X0) * oo2pi;
  const double rho = pow(tv1, OOGM1);
  const double p = pow(rho, G);
  q[0] = rho; q[1] = rho * u; q[2] = rho * v;
  q[3] = p * (1. / (G - 1.)) + 0.5 * rho * (u * u + v * v);
}

}  // namespace dev
}  // namespace gocfd
