/*
 * CPU oracle: plain-C FP64 restatement of the reference 2D Euler DFR solver.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Not part of the product.
 *
 * Follows, function by function, the reference files
 *   readfiles/readSU2Grid.go, readfiles/readGambitGrid.go     (mesh ingest)
 *   types/elemental.go, types/cfd.go, DG2D/triangulation.go    (edge maps)
 *   DG1D/elements.go, DG1D/utils.go, DG2D/basis_polynomials.go (Jacobi bases)
 *   DG2D/lagrange_element.go, DG2D/raviart_thomas_element.go,
 *   DG2D/rt_basis_simplex.go, DG2D/basis_polynomial_construction.go (RT element)
 *   DG2D/dfr_startup.go, DG2D/element_utils.go                 (geometry)
 *   model_problems/Euler2D/{euler.go, edges.go, fluxes.go, bcs.go, fluids.go,
 *     initialization.go, dissipation.go}, isentropic_vortex/analytic_vortex.go
 *   DG2D/dfr_shock_capturing.go                                 (shock sensor)
 *
 * Where the reference is nondeterministic (Go map iteration order), the
 * resolution taken is stated at the function.
 *
 * Parity unpinned: the last bits of the operator matrices (the reference uses
 * gonum LU and Dgemm; this file uses Gauss-Jordan and plain loops) and the bit
 * pattern of multi-step trajectories. Those are checked to tolerance only.
 */
#include "oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wsj_table.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads the parallel loops of the oracle use */
int or_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define MAXTAGS 64

static void seterr(char *err, int errlen, const char *fmt, ...) {
  if (!err || errlen <= 0) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err, (size_t)errlen, fmt, ap);
  va_end(ap);
}

/* ------------------------------------------------------------------------ */
/* Dense linear algebra (row-major)                                          */
/* ------------------------------------------------------------------------ */

/* C[m x n] = A[m x k] * B[k x n]  (utils/matrix_extended.go:244-270 Mul) */
static void mm(const double *A, const double *B, double *C, int m, int k, int n) {
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) {
      double acc = 0;
      for (int l = 0; l < k; l++) acc += A[i * k + l] * B[l * n + j];
      C[i * n + j] = acc;
    }
}

/* Inverse by Gauss-Jordan with partial pivoting. Stands in for gonum
 * Getrf/Getri (utils/matrix_extended.go:953-996, InverseWithCheck :1323-1346).
 * Returns 0 on success, -1 if singular. */
static int minv(const double *A, double *Ainv, int n) {
  double *a = (double *)malloc(sizeof(double) * n * n);
  memcpy(a, A, sizeof(double) * n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Ainv[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int c = 0; c < n; c++) {
    int p = c;
    double best = fabs(a[c * n + c]);
    for (int r = c + 1; r < n; r++)
      if (fabs(a[r * n + c]) > best) { best = fabs(a[r * n + c]); p = r; }
    if (best < 1e-300) { free(a); return -1; }
    if (p != c)
      for (int j = 0; j < n; j++) {
        double t = a[c * n + j]; a[c * n + j] = a[p * n + j]; a[p * n + j] = t;
        t = Ainv[c * n + j]; Ainv[c * n + j] = Ainv[p * n + j]; Ainv[p * n + j] = t;
      }
    double d = a[c * n + c];
    for (int j = 0; j < n; j++) { a[c * n + j] /= d; Ainv[c * n + j] /= d; }
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      double f = a[r * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        a[r * n + j] -= f * a[c * n + j];
        Ainv[r * n + j] -= f * Ainv[c * n + j];
      }
    }
  }
  free(a);
  return 0;
}

/* utils.POW: integer power by repeated multiplication */
static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

/* ------------------------------------------------------------------------ */
/* 1D Jacobi polynomials: DG1D/elements.go:257-328, DG1D/utils.go:35-46     */
/* ------------------------------------------------------------------------ */

static double gamma0(double alpha, double beta) {
  double ab1 = alpha + beta + 1., a1 = alpha + 1., b1 = beta + 1.;
  return tgamma(a1) * tgamma(b1) * pow(2, ab1) / ab1 / tgamma(ab1);
}
static double gamma1(double alpha, double beta) {
  double ab = alpha + beta, a1 = alpha + 1., b1 = beta + 1.;
  return a1 * b1 * gamma0(alpha, beta) / (ab + 3.0);
}

/* Orthonormal Jacobi polynomial P_N^(alpha,beta)(x) by the three-term
 * recurrence of DG1D/elements.go:257-314 */
static double jacobiP(double x, double alpha, double beta, int N) {
  double rg = 1. / sqrt(gamma0(alpha, beta));
  if (N == 0) return rg;
  double ab = alpha + beta;
  double rg1 = 1. / sqrt(gamma1(alpha, beta));
  double pm1 = rg;
  double p = rg1 * ((ab + 2.0) * x / 2.0 + (alpha - beta) / 2.0);
  if (N == 1) return p;
  double a1 = alpha + 1., b1 = beta + 1., ab1 = ab + 1.;
  double aold = 2.0 * sqrt(a1 * b1 / (ab + 3.0)) / (ab + 2.0);
  for (int i = 0; i < N - 1; i++) {
    double ip1 = (double)(i + 1), ip2 = ip1 + 1;
    double h1 = 2.0 * ip1 + ab;
    double anew = 2.0 / (h1 + 2.0) *
                  sqrt(ip2 * (ip1 + ab1) * (ip1 + a1) * (ip1 + b1) / (h1 + 1.0) / (h1 + 3.0));
    double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0);
    double xb = -bnew + x;
    double pn = (-aold * pm1 + xb * p) / anew;
    pm1 = p;
    p = pn;
    aold = anew;
  }
  return p;
}

/* DG1D/elements.go:316-328 */
static double gradJacobiP(double x, double alpha, double beta, int N) {
  if (N == 0) return 0.0;
  double fN = (double)N;
  return jacobiP(x, alpha + 1, beta + 1, N - 1) * sqrt(fN * (fN + alpha + beta + 1));
}

/* ------------------------------------------------------------------------ */
/* 2D orthonormal simplex basis: DG2D/basis_polynomials.go:55-153,          */
/* DG2D/element_utils.go:132-140                                             */
/* ------------------------------------------------------------------------ */

static void rs_to_ab(double r, double s, double *a, double *b) {
  *a = (s != 1) ? 2 * (1 + r) / (1 - s) - 1 : -1;
  *b = s;
}

static double simplex2DP(double r, double s, int i, int j) {
  double a, b;
  rs_to_ab(r, s, &a, &b);
  double h1 = jacobiP(a, 0, 0, i);
  double h2 = jacobiP(b, (double)(2 * i + 1), 0, j);
  return sqrt(2.0) * h1 * h2 * ipow(1 - b, i);
}

static void gradSimplex2DP(double r, double s, int id, int jd, double *ddr, double *dds) {
  double a, b;
  rs_to_ab(r, s, &a, &b);
  double fa = jacobiP(a, 0, 0, id), dfa = gradJacobiP(a, 0, 0, id);
  double gb = jacobiP(b, 2 * (double)id + 1, 0, jd), dgb = gradJacobiP(b, 2 * (double)id + 1, 0, jd);
  double dr = dfa * gb;
  if (id > 0) dr *= ipow(0.5 * (1 - b), id - 1);
  dr *= pow(2, (double)id + 0.5);
  double ds = 0.5 * dfa * gb * (1 + a);
  if (id > 0) ds *= ipow(0.5 * (1 - b), id - 1);
  double tmp = dgb * ipow(0.5 * (1 - b), id);
  if (id > 0) tmp -= 0.5 * (double)id * gb * ipow(0.5 * (1 - b), id - 1);
  ds += fa * tmp;
  ds *= pow(2, (double)id + 0.5);
  *ddr = dr;
  *dds = ds;
}

/* Vandermonde2D: rows = points, columns ordered (i, j<=N-i) */
static void vandermonde2D(int N, const double *R, const double *S, int np, double *V) {
  int Np = (N + 1) * (N + 2) / 2;
  for (int p = 0; p < np; p++) {
    int sk = 0;
    for (int i = 0; i <= N; i++)
      for (int j = 0; j <= N - i; j++) V[p * Np + sk++] = simplex2DP(R[p], S[p], i, j);
  }
}

/* Lagrange basis value / gradient on a 2D node set through Vinv
 * (GetOrthogonalPolynomialAtJ, basis_polynomials.go:192-214) */
static double lag2d(int N, const double *Vinv, double r, double s, int jcol, int deriv) {
  int Np = (N + 1) * (N + 2) / 2;
  double phi = 0;
  int sk = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N - i; j++) {
      double t;
      if (deriv == 0) t = simplex2DP(r, s, i, j);
      else {
        double dr, ds;
        gradSimplex2DP(r, s, i, j, &dr, &ds);
        t = (deriv == 1) ? dr : ds;
      }
      phi += t * Vinv[sk * Np + jcol];
      sk++;
    }
  return phi;
}

/* 1D Lagrange basis on the edge nodes (JacobiBasis1D.GetOrthogonalPolynomialAtJ,
 * basis_polynomials.go:304-312) */
static double lag1d(int P, const double *Vinv1, double t, int jcol, int deriv) {
  int Np = P + 1;
  double phi = 0;
  for (int i = 0; i < Np; i++)
    phi += (deriv ? gradJacobiP(t, 0, 0, i) : jacobiP(t, 0, 0, i)) * Vinv1[i * Np + jcol];
  return phi;
}

/* ------------------------------------------------------------------------ */
/* WSJ nodes: nodes2d_williams_shunn_jameson.go:146-190                     */
/* ------------------------------------------------------------------------ */
static void wsj_rs(int N, double *R, double *S, double *W) {
  int Np = (N + 1) * (N + 2) / 2;
  for (int i = 0; i < Np; i++) {
    const double *b = kWSJ[N][i];
    R[i] = -b[0] + b[1] - b[2];
    S[i] = -b[0] - b[1] + b[2];
    if (W) W[i] = b[3];
  }
}

/* GetOptimizedEdgePoints, raviart_thomas_element.go:377-400 */
static int rt_edge_points(int NRT, double *g) {
  static const double e1[] = {-0.38490018, 0.38490018};
  static const double e2[] = {-0.028, 0.000, 0.028};
  static const double e3[] = {-0.482, -0.161, 0.161, 0.482};
  static const double e4[] = {-0.586, -0.293, 0.000, 0.293, 0.586};
  static const double e5[] = {-0.634, -0.381, -0.127, 0.127, 0.381, 0.634};
  static const double e6[] = {-0.671, -0.452, -0.219, 0.000, 0.219, 0.452, 0.671};
  static const double e7[] = {-0.808, -0.578, -0.343, -0.113, 0.113, 0.343, 0.578, 0.808};
  static const double e8[] = {-0.692, -0.523, -0.350, -0.176, 0.000, 0.176, 0.350, 0.523, 0.692};
  const double *t[] = {NULL, e1, e2, e3, e4, e5, e6, e7, e8};
  if (NRT < 1 || NRT > 8) return -1;
  for (int i = 0; i <= NRT; i++) g[i] = t[NRT][i];
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Operators                                                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  int N, P, NpInt, NpEdge, NpFlux;
  double *R, *S;           /* RT points [NpFlux] */
  double *Div, *DivInt;    /* [NpFlux x NpFlux], [NpInt x NpFlux] */
  double *Interp;          /* FluxEdgeInterp [3NpEdge x NpInt] */
  double *solR, *solS, *solW; /* solution nodes [NpInt] */
  double *V, *Vinv, *Mass; /* solution element [NpInt x NpInt] */
} ops_t;

/* Value and divergence of RT basis function j at (r,s): rt_basis_simplex.go:39-194,
 * basis_polynomial_construction.go:331-359 */
typedef struct {
  int P, NpInt, NpEdge;
  const double *Vinv2; /* interior Lagrange (order P-1) */
  const double *Vinv1; /* edge Lagrange (order P) */
} rtb_t;

static void rt_phi(const rtb_t *b, int j, double r, double s, double v[2], double *div) {
  double rp = 0.5 * (r + 1.), rm = 0.5 * (r - 1.), sp = 0.5 * (s + 1.), sm = 0.5 * (s - 1.);
  double vec[2], vdiv, poly, g[2];
  int NpInt = b->NpInt, NpEdge = b->NpEdge;
  if (j < 2 * NpInt) {
    int jj = (j < NpInt) ? j : j - NpInt;
    if (j < NpInt) { vec[0] = sp * rp; vec[1] = sp * sm; vdiv = (3. * s + 1.) / 4.; }
    else { vec[0] = rp * rm; vec[1] = rp * sp; vdiv = (3. * r + 1.) / 4.; }
    poly = lag2d(b->P - 1, b->Vinv2, r, s, jj, 0);
    g[0] = lag2d(b->P - 1, b->Vinv2, r, s, jj, 1);
    g[1] = lag2d(b->P - 1, b->Vinv2, r, s, jj, 2);
  } else {
    int jj = j - 2 * NpInt;
    if (jj < NpEdge) { /* E1 bottom: parameter r */
      poly = lag1d(b->P, b->Vinv1, r, jj, 0);
      g[0] = lag1d(b->P, b->Vinv1, r, jj, 1); g[1] = 0;
      vec[0] = rp; vec[1] = sm; vdiv = 1.;
    } else if (jj < 2 * NpEdge) { /* E2 hypotenuse: parameter s */
      jj -= NpEdge;
      double sr2 = sqrt(2.0);
      poly = lag1d(b->P, b->Vinv1, s, jj, 0);
      g[0] = 0; g[1] = lag1d(b->P, b->Vinv1, s, jj, 1);
      vec[0] = sr2 * rp; vec[1] = sr2 * sp; vdiv = 1. * sr2;
    } else { /* E3 left: parameter -s */
      jj -= 2 * NpEdge;
      poly = lag1d(b->P, b->Vinv1, -s, jj, 0);
      g[0] = 0; g[1] = -lag1d(b->P, b->Vinv1, -s, jj, 1);
      vec[0] = rm; vec[1] = sp; vdiv = 1.;
    }
  }
  v[0] = vec[0] * poly;
  v[1] = vec[1] * poly;
  *div = poly * vdiv + (vec[0] * g[0] + vec[1] * g[1]);
}

/* NewRTElement(N+1) (raviart_thomas_element.go:192-249) plus the solution
 * element (lagrange_element.go:54-92) and FluxEdgeInterp (dfr_startup.go:49-64) */
static int build_ops(int N, ops_t *o, char *err, int errlen) {
  if (N < 0 || N > 7) { seterr(err, errlen, "Polynomial order must be 0..7, have %d", N); return -1; }
  int P = N + 1;
  int NpFlux = (P + 1) * (P + 3), NpInt = P * (P + 1) / 2, NpEdge = P + 1;
  o->N = N; o->P = P; o->NpInt = NpInt; o->NpEdge = NpEdge; o->NpFlux = NpFlux;
  o->R = calloc(NpFlux, sizeof(double)); o->S = calloc(NpFlux, sizeof(double));
  o->solR = calloc(NpInt, sizeof(double)); o->solS = calloc(NpInt, sizeof(double));
  o->solW = calloc(NpInt, sizeof(double));
  wsj_rs(N, o->solR, o->solS, o->solW);
  /* ExtendGeomToRT :429-496 */
  double g[16];
  if (rt_edge_points(P, g)) { seterr(err, errlen, "no RT edge points for order %d", P); return -1; }
  for (int i = 0; i < NpInt; i++) {
    o->R[i] = o->R[i + NpInt] = o->solR[i];
    o->S[i] = o->S[i + NpInt] = o->solS[i];
  }
  for (int i = 0; i < NpEdge; i++) {
    double gp = g[i];
    o->R[2 * NpInt + i] = gp; o->S[2 * NpInt + i] = -1;
    double gpT = 0.5 * (gp + 1);
    o->R[2 * NpInt + NpEdge + i] = 1 - 2 * gpT; o->S[2 * NpInt + NpEdge + i] = -1 + 2 * gpT;
    o->R[2 * NpInt + 2 * NpEdge + i] = -1; o->S[2 * NpInt + 2 * NpEdge + i] = -gp;
  }
  /* Interior Lagrange basis: NewJacobiBasis2D(P-1, RInt, SInt) */
  double *V2 = calloc(NpInt * NpInt, sizeof(double)), *V2inv = calloc(NpInt * NpInt, sizeof(double));
  vandermonde2D(P - 1, o->solR, o->solS, NpInt, V2);
  if (minv(V2, V2inv, NpInt)) { seterr(err, errlen, "singular interior Vandermonde"); return -1; }
  /* Edge Lagrange basis: NewJacobiBasis1D(P, REdge) */
  double *V1 = calloc(NpEdge * NpEdge, sizeof(double)), *V1inv = calloc(NpEdge * NpEdge, sizeof(double));
  for (int i = 0; i < NpEdge; i++)
    for (int j = 0; j < NpEdge; j++) V1[i * NpEdge + j] = jacobiP(o->R[2 * NpInt + i], 0, 0, j);
  if (minv(V1, V1inv, NpEdge)) { seterr(err, errlen, "singular edge Vandermonde"); return -1; }
  rtb_t b = {P, NpInt, NpEdge, V2inv, V1inv};
  /* DOF unit vectors :224-234 */
  double oosr2 = 0.5 * sqrt(2.0);
  double *V = calloc(NpFlux * NpFlux, sizeof(double)), *VInv = calloc(NpFlux * NpFlux, sizeof(double));
  double *DivB = calloc(NpFlux * NpFlux, sizeof(double));
  for (int i = 0; i < NpFlux; i++) {
    double bv[2];
    if (i < NpInt) { bv[0] = 1; bv[1] = 0; }
    else if (i < 2 * NpInt) { bv[0] = 0; bv[1] = 1; }
    else if (i < 2 * NpInt + NpEdge) { bv[0] = 0; bv[1] = -1; }
    else if (i < 2 * NpInt + 2 * NpEdge) { bv[0] = oosr2; bv[1] = oosr2; }
    else { bv[0] = -1; bv[1] = 0; }
    for (int j = 0; j < NpFlux; j++) {
      double v[2], d;
      rt_phi(&b, j, o->R[i], o->S[i], v, &d);
      V[i * NpFlux + j] = bv[0] * v[0] + bv[1] * v[1]; /* ComposeV :265-282 */
      DivB[i * NpFlux + j] = d;                        /* ComputeDivergenceMatrix :284-297 */
    }
  }
  if (minv(V, VInv, NpFlux)) { seterr(err, errlen, "singular RT Vandermonde"); return -1; }
  o->Div = calloc(NpFlux * NpFlux, sizeof(double));
  mm(DivB, VInv, o->Div, NpFlux, NpFlux, NpFlux);
  o->DivInt = calloc(NpInt * NpFlux, sizeof(double));
  memcpy(o->DivInt, o->Div, sizeof(double) * NpInt * NpFlux);
  /* Solution element: V, Vinv, Mass = V^T W V (lagrange_element.go:79-82) */
  o->V = calloc(NpInt * NpInt, sizeof(double)); o->Vinv = calloc(NpInt * NpInt, sizeof(double));
  o->Mass = calloc(NpInt * NpInt, sizeof(double));
  vandermonde2D(N, o->solR, o->solS, NpInt, o->V);
  if (minv(o->V, o->Vinv, NpInt)) { seterr(err, errlen, "singular solution Vandermonde"); return -1; }
  for (int i = 0; i < NpInt; i++)
    for (int j = 0; j < NpInt; j++) {
      double acc = 0;
      for (int l = 0; l < NpInt; l++) acc += o->V[l * NpInt + i] * o->solW[l] * o->V[l * NpInt + j];
      o->Mass[i * NpInt + j] = acc;
    }
  /* FluxEdgeInterp = V2D(edge points) * Vinv (basis_polynomials.go:49-53) */
  int ne = 3 * NpEdge;
  double *Ve = calloc(ne * NpInt, sizeof(double));
  vandermonde2D(N, o->R + 2 * NpInt, o->S + 2 * NpInt, ne, Ve);
  o->Interp = calloc(ne * NpInt, sizeof(double));
  mm(Ve, o->Vinv, o->Interp, ne, NpInt, NpInt);
  free(V2); free(V2inv); free(V1); free(V1inv); free(V); free(VInv); free(DivB); free(Ve);
  return 0;
}

static void free_ops(ops_t *o) {
  free(o->R); free(o->S); free(o->Div); free(o->DivInt); free(o->Interp);
  free(o->solR); free(o->solS); free(o->solW); free(o->V); free(o->Vinv); free(o->Mass);
}

/* ------------------------------------------------------------------------ */
/* Mesh ingest                                                               */
/* ------------------------------------------------------------------------ */
typedef struct { int v0, v1; } edgeint_t; /* ordered vertex pair (types.EdgeInt) */
typedef struct {
  char tag[64];
  int n;
  edgeint_t *e;
} bctag_t;

struct or_mesh {
  int K, Nv;
  double *VX, *VY;
  int *EToV;
  int ntags;
  bctag_t tags[MAXTAGS];
};

static void lower_trim(char *s) {
  char *p = s;
  while (*p == ' ' || *p == '\t') p++;
  memmove(s, p, strlen(p) + 1);
  size_t l = strlen(s);
  while (l > 0 && (s[l - 1] == ' ' || s[l - 1] == '\t' || s[l - 1] == '\r' || s[l - 1] == '\n')) s[--l] = 0;
  for (p = s; *p; p++) *p = (char)tolower((unsigned char)*p);
}

/* BCMAP.AddEdges: a repeated tag appends (types/cfd.go:86-104) */
static int mesh_add_edges(or_mesh *m, const char *tag, edgeint_t *e, int n) {
  for (int t = 0; t < m->ntags; t++)
    if (!strcmp(m->tags[t].tag, tag)) {
      m->tags[t].e = realloc(m->tags[t].e, sizeof(edgeint_t) * (m->tags[t].n + n));
      memcpy(m->tags[t].e + m->tags[t].n, e, sizeof(edgeint_t) * n);
      m->tags[t].n += n;
      free(e);
      return 0;
    }
  if (m->ntags >= MAXTAGS) return -1;
  bctag_t *b = &m->tags[m->ntags++];
  snprintf(b->tag, sizeof(b->tag), "%s", tag);
  b->n = n;
  b->e = e;
  return 0;
}

static int getline_(FILE *f, char *buf, int len) {
  if (!fgets(buf, len, f)) return -1;
  size_t l = strlen(buf);
  while (l > 0 && (buf[l - 1] == '\n' || buf[l - 1] == '\r')) buf[--l] = 0;
  return 0;
}

/* getToken / readNumber / readLabel (readSU2Grid.go:104-154) */
static int su2_token(FILE *f, char *tok, int len) {
  char line[1024];
  for (;;) {
    if (getline_(f, line, sizeof line)) return -1;
    char *p = line;
    while (*p == ' ') p++;
    if (*p == '%') continue;
    char *eq = strchr(p, '=');
    if (!eq) return -2;
    snprintf(tok, len, "%s", eq + 1);
    return 0;
  }
}

static or_mesh *read_su2(FILE *f, char *err, int errlen) {
  char tok[1024], line[1024];
  or_mesh *m = calloc(1, sizeof(or_mesh));
  int ndim;
  if (su2_token(f, tok, sizeof tok) || sscanf(tok, "%d", &ndim) != 1) goto bad;
  if (su2_token(f, tok, sizeof tok) || sscanf(tok, "%d", &m->K) != 1) goto bad;
  m->EToV = malloc(sizeof(int) * 3 * m->K);
  for (int k = 0; k < m->K; k++) {
    int t, a, b, c;
    if (getline_(f, line, sizeof line) || sscanf(line, "%d %d %d %d", &t, &a, &b, &c) != 4) goto bad;
    if (t != 5) { seterr(err, errlen, "unable to deal with non-triangular elements right now"); goto fail; }
    m->EToV[3 * k] = a; m->EToV[3 * k + 1] = b; m->EToV[3 * k + 2] = c;
  }
  if (su2_token(f, tok, sizeof tok) || sscanf(tok, "%d", &m->Nv) != 1) goto bad;
  m->VX = malloc(sizeof(double) * m->Nv); m->VY = malloc(sizeof(double) * m->Nv);
  for (int i = 0; i < m->Nv; i++) {
    if (getline_(f, line, sizeof line) || sscanf(line, "%lf %lf", &m->VX[i], &m->VY[i]) != 2) goto bad;
  }
  int nmark = 0;
  if (su2_token(f, tok, sizeof tok) == 0) sscanf(tok, "%d", &nmark);
  for (int n = 0; n < nmark; n++) {
    char label[256];
    int ne;
    if (su2_token(f, tok, sizeof tok) || sscanf(tok, "%255s", label) != 1) goto bad;
    lower_trim(label);
    if (su2_token(f, tok, sizeof tok) || sscanf(tok, "%d", &ne) != 1) goto bad;
    edgeint_t *e = malloc(sizeof(edgeint_t) * (ne > 0 ? ne : 1));
    for (int i = 0; i < ne; i++) {
      int t, a, b;
      if (getline_(f, line, sizeof line) || sscanf(line, "%d %d %d", &t, &a, &b) != 3) { free(e); goto bad; }
      if (t != 3) { free(e); seterr(err, errlen, "BCs should only contain line elements in 2D"); goto fail; }
      e[i].v0 = a; e[i].v1 = b;
    }
    if (mesh_add_edges(m, label, e, ne)) goto bad;
  }
  return m;
bad:
  seterr(err, errlen, "malformed SU2 file");
fail:
  or_mesh_free(m);
  return NULL;
}

static int skiplines(FILE *f, int n) {
  char line[1024];
  for (int i = 0; i < n; i++)
    if (getline_(f, line, sizeof line)) return -1;
  return 0;
}

/* ReadGambit2d (readGambitGrid.go:21-165, 200-272, 334-363) */
static or_mesh *read_gambit(FILE *f, char *err, int errlen) {
  char line[1024];
  or_mesh *m = calloc(1, sizeof(or_mesh));
  int Nmats, Nbcs, Nsd, dum;
  if (skiplines(f, 6)) goto bad;
  if (getline_(f, line, sizeof line) ||
      sscanf(line, "%d %d %d %d %d %d", &m->Nv, &m->K, &Nmats, &Nbcs, &Nsd, &dum) != 6) goto bad;
  if (Nsd != 2) { seterr(err, errlen, "only 2D Gambit meshes are supported"); goto fail; }
  if (skiplines(f, 2)) goto bad;
  m->VX = calloc(m->Nv, sizeof(double)); m->VY = calloc(m->Nv, sizeof(double));
  for (int i = 0; i < m->Nv; i++) {
    int ind; double x, y;
    if (getline_(f, line, sizeof line) || sscanf(line, "%d %lf %lf", &ind, &x, &y) != 3) goto bad;
    if (ind < 1 || ind > m->Nv) goto bad;
    m->VX[ind - 1] = x; m->VY[ind - 1] = y;
  }
  if (skiplines(f, 2)) goto bad;
  m->EToV = calloc(3 * (size_t)m->K, sizeof(int));
  for (int i = 0; i < m->K; i++) {
    int ind, typ, nf, a, b, c;
    if (getline_(f, line, sizeof line) || sscanf(line, "%d %d %d %d %d %d", &ind, &typ, &nf, &a, &b, &c) != 6) goto bad;
    if (ind < 1 || ind > m->K) goto bad;
    m->EToV[3 * (ind - 1)] = a - 1; m->EToV[3 * (ind - 1) + 1] = b - 1; m->EToV[3 * (ind - 1) + 2] = c - 1;
  }
  if (skiplines(f, 2)) goto bad;
  for (int g = 0; g < Nmats; g++) {
    int gn, elnum; double matval;
    if (getline_(f, line, sizeof line) ||
        sscanf(line, "GROUP: %d ELEMENTS: %d MATERIAL: %lf", &gn, &elnum, &matval) != 3) goto bad;
    if (skiplines(f, 2)) goto bad; /* title + flags line */
    int nl = elnum / 10 + ((elnum % 10) ? 1 : 0);
    if (skiplines(f, nl)) goto bad;
    if (skiplines(f, 2)) goto bad;
  }
  for (int i = 0; i < Nbcs; i++) {
    if (i != 0 && skiplines(f, 1)) goto bad;
    char name[64]; double pf; int bcid, nfaces;
    if (getline_(f, line, sizeof line) || sscanf(line, "%63s", name) != 1) goto bad;
    lower_trim(name);
    if (!strcmp(name, "cyl")) {
      if (sscanf(line, "%*s %lf %d", &pf, &nfaces) != 2) goto bad;
    } else {
      if (sscanf(line, "%*s %d %d", &bcid, &nfaces) != 2) goto bad;
    }
    edgeint_t *e = malloc(sizeof(edgeint_t) * (nfaces > 0 ? nfaces : 1));
    for (int j = 0; j < nfaces; j++) {
      int kp1, n2, fp1;
      if (getline_(f, line, sizeof line) || sscanf(line, "%d %d %d", &kp1, &n2, &fp1) != 3) { free(e); goto bad; }
      int *v = &m->EToV[3 * (kp1 - 1)];
      if (fp1 == 1) { e[j].v0 = v[0]; e[j].v1 = v[1]; }
      else if (fp1 == 2) { e[j].v0 = v[1]; e[j].v1 = v[2]; }
      else { e[j].v0 = v[2]; e[j].v1 = v[0]; }
    }
    if (mesh_add_edges(m, name, e, nfaces)) goto bad;
    if (skiplines(f, 1)) goto bad;
  }
  return m;
bad:
  seterr(err, errlen, "malformed Gambit neutral file");
fail:
  or_mesh_free(m);
  return NULL;
}

or_mesh *or_mesh_read(const char *path, char *err, int errlen) {
  size_t l = strlen(path);
  if (l < 4 || path[l - 4] != '.') { seterr(err, errlen, "unable to determine file type from name: %s", path); return NULL; }
  const char *ext = path + l - 3;
  FILE *f = fopen(path, "r");
  if (!f) { seterr(err, errlen, "unable to open file %s", path); return NULL; }
  or_mesh *m = NULL;
  if (!strcmp(ext, "su2")) m = read_su2(f, err, errlen);
  else if (!strcmp(ext, "neu")) m = read_gambit(f, err, errlen);
  else seterr(err, errlen, "unsupported file type: %s", path);
  fclose(f);
  return m;
}

void or_mesh_free(or_mesh *m) {
  if (!m) return;
  free(m->VX); free(m->VY); free(m->EToV);
  for (int t = 0; t < m->ntags; t++) free(m->tags[t].e);
  free(m);
}
int or_mesh_K(const or_mesh *m) { return m->K; }
int or_mesh_Nv(const or_mesh *m) { return m->Nv; }
void or_mesh_get(const or_mesh *m, double *VX, double *VY, int *EToV) {
  memcpy(VX, m->VX, sizeof(double) * m->Nv);
  memcpy(VY, m->VY, sizeof(double) * m->Nv);
  memcpy(EToV, m->EToV, sizeof(int) * 3 * m->K);
}

/* BCTAG.GetFLAG (types/cfd.go:28-73) */
static int bc_flag(const char *tag) {
  char base[64];
  snprintf(base, sizeof base, "%s", tag);
  char *d = strchr(base, '-');
  if (d && d != base) *d = 0;
  static const struct { const char *n; int f; } map[] = {
      {"inflow", OR_BC_IN}, {"in", OR_BC_IN}, {"out", OR_BC_OUT}, {"outflow", OR_BC_OUT},
      {"wall", OR_BC_WALL}, {"far", OR_BC_FAR}, {"cyl", OR_BC_CYL}, {"dirichlet", OR_BC_DIRICHLET},
      {"neuman", OR_BC_NEUMAN}, {"slip", OR_BC_SLIP}, {"periodic", OR_BC_PERIODIC},
      {"periodicreversed", OR_BC_PERIODIC_REVERSED}};
  for (size_t i = 0; i < sizeof map / sizeof map[0]; i++)
    if (!strcmp(map[i].n, base)) return map[i].f;
  return -1;
}

/* ------------------------------------------------------------------------ */
/* Triangulation (DG2D/triangulation.go:19-171)                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t key;
  int nconn;
  int tri[2], enr[2], dir[2];
  int bc;
  double IInII[2];
} edge_t;

typedef struct { uint64_t key; int seq; int k, en, dir; } face_rec;

static int cmp_face(const void *a, const void *b) {
  const face_rec *x = a, *y = b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->seq - y->seq;
}

static uint64_t edge_key(int a, int b) {
  uint64_t lo = (uint64_t)(a < b ? a : b), hi = (uint64_t)(a < b ? b : a);
  return lo + (hi << 32);
}

static int find_edge(const edge_t *E, int ne, uint64_t key) {
  int lo = 0, hi = ne - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (E[mid].key == key) return mid;
    if (E[mid].key < key) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* AddTri's ||n|| factor (triangulation.go:149-171) with GetEdgeCoordinates (:230-235) */
static double edge_iinii(uint64_t key, int rev, int enr, const double *VX, const double *VY,
                         double *nx, double *ny) {
  int lo = (int)(key & 0xffffffffu), hi = (int)(key >> 32);
  /* GetVertices(!rev): [lo,hi], swapped when !rev */
  int a = rev ? lo : hi, b = rev ? hi : lo;
  double dx = VX[b] - VX[a], dy = VY[b] - VY[a];
  double len = sqrt(dy * dy + dx * dx);
  if (nx) {
    double oonorm = 1. / sqrt(dx * dx + dy * dy); /* CalculateFaceNorms dfr_startup.go:304-307 */
    *nx = -dy * oonorm;
    *ny = dx * oonorm;
  }
  return (enr == 1) ? len / (2. * sqrt(2)) : len / 2.;
}

/* Curve.ReOrder + AssembleCurve (types/elemental.go:141-241). When the list
 * does not start at a curve end the reference starts from a map-order
 * dependent end; this restatement takes the end with the smaller vertex id. */
static int curve_reorder(edgeint_t *c, int l, int reverse) {
  /* collect vertices */
  int nv = 0, cap = 2 * l;
  int *vid = malloc(sizeof(int) * cap), *cnt = calloc(cap, sizeof(int));
  int (*ve)[2] = malloc(sizeof(int[2]) * cap);
  for (int i = 0; i < l; i++) {
    int vv[2] = {c[i].v0, c[i].v1};
    for (int t = 0; t < 2; t++) {
      int f = -1;
      for (int q = 0; q < nv; q++) if (vid[q] == vv[t]) { f = q; break; }
      if (f < 0) { f = nv++; vid[f] = vv[t]; }
      if (cnt[f] >= 2) { free(vid); free(cnt); free(ve); return -1; }
      ve[f][cnt[f]++] = i;
    }
  }
  int ends[2], ne = 0;
  for (int q = 0; q < nv; q++) if (cnt[q] == 1) { if (ne == 2) { ne = 3; break; } ends[ne++] = q; }
  if (ne != 2) { free(vid); free(cnt); free(ve); return -1; }
  int start = reverse ? l - 1 : 0;
  int ok = (ve[ends[0]][0] == start) || (ve[ends[1]][0] == start);
  if (!ok) {
    int e0 = (vid[ends[0]] < vid[ends[1]]) ? ends[0] : ends[1];
    start = ve[e0][0];
  }
  edgeint_t *out = malloc(sizeof(edgeint_t) * l);
  int ii = 0;
  out[ii++] = c[start];
  int endv = c[start].v0;
  int q;
  for (q = 0; q < nv && vid[q] != endv; q++) ;
  if (cnt[q] != 2) endv = c[start].v1;
  int cur = start;
  for (;;) {
    for (q = 0; q < nv && vid[q] != endv; q++) ;
    if (cnt[q] != 2) break;
    int nxt = (ve[q][0] != cur) ? ve[q][0] : ve[q][1];
    if (nxt == cur || ii >= l) break;
    out[ii++] = c[nxt];
    endv = (c[nxt].v0 == endv) ? c[nxt].v1 : c[nxt].v0;
    cur = nxt;
  }
  int rc = (ii == l) ? 0 : -1;
  if (!rc) memcpy(c, out, sizeof(edgeint_t) * l);
  free(out); free(vid); free(cnt); free(ve);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Solver state                                                              */
/* ------------------------------------------------------------------------ */
typedef struct { double gamma, Qinf[4], Pinf, QQinf, Cinf, Alpha, Minf; } fs_t;

struct or_solver {
  or_params p;
  ops_t o;
  int K, Nv, NpInt, NpEdge, NpFlux;
  double *VX, *VY;
  int *EToV;
  int ne;
  edge_t *E;
  int *EtoE;      /* [K][3] */
  int *EtoEdge;   /* [K][3]: own edge (EtoEdges) */
  int *FaceEdge;  /* [3][K]: EdgeNumber as edge index */
  double *Jdet, *Jinv, *FNx, *FNy, *IInII, *EdgeLenMax;
  double *SolX, *SolY, *FluxX, *FluxY;
  fs_t fs;          /* FSFar */
  fs_t fsin, fsout; /* FSIn, FSOut: set by the shock-tube case, else FSFar */
  int ivortex;
  /* state */
  double *Q, *Q1, *Q2, *Q3, *Q4, *RHSQ, *Resid, *DT, *DTVisc, *QFace, *FRT;
  double *EdgeFlux, *EdgeQ, *EdgeAgg, *EdgeAggVisc, *EdgeVisc;
  double time, gdt;
  int steps;
  int nan_scan;   /* 0 none, 1 stage input + q0 + DT, 2 the reference's eleven arrays */
  /* shock capturing */
  int shock;
  double *Se, *Sigma, *Eps, *SigmaV, *EpsV, *EpsF, *SigmaF, *DissX, *DissY, *Clip, *Dmat, *Pmat;
  double mf[64];
  double kappa_sd, eps0;
  double *Bary; /* [NpFlux][3] */
};

/* FreeStream (fluids.go:245-273) and GetFlowFunctionBase (:289-336) */
static double fs_pressure(double g, double rho, double rhoU, double rhoV, double E) {
  double oorho = 1. / rho, u = rhoU * oorho, v = rhoV * oorho;
  double U2 = u * u + v * v, q = 0.5 * rho * U2;
  return (g - 1.) * (E - q);
}
static double fs_sound(double g, double rho, double rhoU, double rhoV, double E) {
  double oorho = 1. / rho;
  double p = fs_pressure(g, rho, rhoU, rhoV, E);
  return sqrt(fabs(g * p * oorho));
}
static double fs_velocity(double rho, double rhoU, double rhoV) {
  double oorho = 1. / rho, u = rhoU * oorho, v = rhoV * oorho;
  return sqrt(u * u + v * v);
}
static void fs_from_qinf(fs_t *fs, double gamma, const double q[4]) {
  fs->gamma = gamma;
  memcpy(fs->Qinf, q, sizeof(double) * 4);
  fs->Pinf = fs_pressure(gamma, q[0], q[1], q[2], q[3]);
  double oorho = 1. / q[0], u = q[1] * oorho, v = q[2] * oorho;
  fs->QQinf = 0.5 * q[0] * (u * u + v * v);
  fs->Cinf = fs_sound(gamma, q[0], q[1], q[2], q[3]);
}
static void fs_new(fs_t *fs, double Minf, double Gamma, double Alpha) {
  double ooggm1 = 1. / (Gamma * (Gamma - 1.));
  double uinf = Minf * cos(Alpha * M_PI / 180.), vinf = Minf * sin(Alpha * M_PI / 180.);
  double q[4] = {1, uinf, vinf, ooggm1 + 0.5 * Minf * Minf};
  fs_from_qinf(fs, Gamma, q);
  fs->Alpha = Alpha;
  fs->Minf = Minf;
}

void or_freestream(double Minf, double Gamma, double Alpha, double *q4, double *pinf, double *cinf) {
  fs_t fs;
  fs_new(&fs, Minf, Gamma, Alpha);
  memcpy(q4, fs.Qinf, sizeof(double) * 4);
  *pinf = fs.Pinf;
  *cinf = fs.Cinf;
}

/* IVortex.GetState / GetStateC (isentropic_vortex/analytic_vortex.go:31-82),
 * Beta=5, centre (5,0), gamma=1.4, Ufs=1 (initialization.go:61-71) */
static void ivortex_state(double t, double x, double y, double q[4]) {
  const double beta = 5., X0 = 5., Y0 = 0., Gamma = 1.4;
  double oo2pi = 0.5 * (1. / M_PI), GM1 = Gamma - 1, OOGM1 = 1. / GM1;
  double pi2 = M_PI * M_PI, beta2 = beta * beta, fac = 16 * Gamma * pi2;
  double u = 1.0, v = 0.;
  double xmut = x - u * t, ymvt = y - v * t;
  double r2 = ipow(xmut - X0, 2) + ipow(ymvt - Y0, 2);
  double ex1r = exp(1 - r2);
  double tv1 = 1.0 - (GM1 * beta2 * exp(2.0 * (1.0 - r2)) / fac);
  u -= beta * ex1r * (ymvt - Y0) * oo2pi;
  v += beta * ex1r * (xmut - X0) * oo2pi;
  double rho = pow(tv1, OOGM1);
  double p = pow(rho, Gamma);
  double ooGM1 = 1. / (Gamma - 1.);
  double qq = 0.5 * rho * (u * u + v * v);
  q[0] = rho; q[1] = rho * u; q[2] = rho * v; q[3] = p * ooGM1 + qq;
}
void or_ivortex_state(double t, double x, double y, double *q4) { ivortex_state(t, x, y, q4); }

/* FluxCalcBase (fluxes.go:76-87) */
static void flux_xy(double g, const double q[4], double Fx[4], double Fy[4]) {
  double rho = q[0], rhoU = q[1], rhoV = q[2], E = q[3];
  double oorho = 1. / rho, u = rhoU * oorho, v = rhoV * oorho;
  double p = fs_pressure(g, rho, rhoU, rhoV, E);
  Fx[0] = rhoU; Fx[1] = rhoU * u + p; Fx[2] = rhoU * v; Fx[3] = u * (E + p);
  Fy[0] = rhoV; Fy[1] = rhoV * u; Fy[2] = rhoV * v + p; Fy[3] = v * (E + p);
}

/* AvgFlux (fluxes.go:135-159) */
static void avg_flux(double g, const double QL[4], const double QR[4], const double n[2], double F[4]) {
  double FxL[4], FyL[4], FxR[4], FyR[4];
  flux_xy(g, QL, FxL, FyL);
  flux_xy(g, QR, FxR, FyR);
  for (int c = 0; c < 4; c++) {
    double fx = 0.5 * (FxL[c] + FxR[c]), fy = 0.5 * (FyL[c] + FyR[c]);
    F[c] = n[0] * fx + n[1] * fy;
  }
}

/* LaxFlux (fluxes.go:161-190): speed magnitude, not normal speed */
static void lax_flux(double g, const double QL[4], const double QR[4], const double n[2], double F[4]) {
  double nx = n[0], ny = n[1];
  double rhoL = QL[0], rhoR = QR[0], rhoUL = QL[1], rhoVL = QL[2], rhoUR = QR[1], rhoVR = QR[2];
  double EL = QL[3], ER = QR[3];
  double uL = rhoUL / rhoL, vL = rhoVL / rhoL, uR = rhoUR / rhoR, vR = rhoVR / rhoR;
  double pL = fs_pressure(g, QL[0], QL[1], QL[2], QL[3]), pR = fs_pressure(g, QR[0], QR[1], QR[2], QR[3]);
  double CL = fs_sound(g, QL[0], QL[1], QL[2], QL[3]), CR = fs_sound(g, QR[0], QR[1], QR[2], QR[3]);
  double a = sqrt(uL * uL + vL * vL) + CL, b = sqrt(uR * uR + vR * vR) + CR;
  double maxV = a > b ? a : b;
  F[0] = 0.5 * (nx * (rhoUL + rhoUR) + ny * (rhoVL + rhoVR));
  F[1] = 0.5 * (nx * (rhoUL * uL + rhoUR * uR + pL + pR) + ny * (rhoUL * vL + rhoUR * vR));
  F[2] = 0.5 * (nx * (rhoVL * uL + rhoVR * uR) + ny * (rhoVL * vL + rhoVR * vR + pL + pR));
  F[3] = 0.5 * (nx * ((pL + EL) * uL + (pR + ER) * uR) + ny * ((pL + EL) * vL + (pR + ER) * vR));
  for (int c = 0; c < 4; c++) F[c] += 0.5 * maxV * (QL[c] - QR[c]);
}

/* RoeFlux (fluxes.go:284-413): no entropy fix */
static void roe_flux(double g, const double QL[4], const double QR[4], const double n[2], double F[4]) {
  double GM1 = g - 1;
  double rhoULr = QL[1] * n[0] + QL[2] * n[1], rhoVLr = QL[1] * (-n[1]) + QL[2] * n[0];
  double rhoURr = QR[1] * n[0] + QR[2] * n[1], rhoVRr = QR[1] * (-n[1]) + QR[2] * n[0];
  double rhoL = QL[0], uL = rhoULr / QL[0], vL = rhoVLr / QL[0];
  double rhoR = QR[0], uR = rhoURr / QR[0], vR = rhoVRr / QR[0];
  double pL = fs_pressure(g, QL[0], QL[1], QL[2], QL[3]);
  double pR = fs_pressure(g, QR[0], QR[1], QR[2], QR[3]);
  double hL = (QL[3] + pL) / rhoL, hR = (QR[3] + pR) / rhoR;
  double rhoLs = sqrt(rhoL), rhoRs = sqrt(rhoR), rhoLsRs = rhoLs + rhoRs;
  double rho = rhoLs * rhoRs;
  double u = (rhoLs * uL + rhoRs * uR) / rhoLsRs;
  double v = (rhoLs * vL + rhoRs * vR) / rhoLsRs;
  double h = (rhoLs * hL + rhoRs * hR) / rhoLsRs;
  double c2 = GM1 * (h - 0.5 * (u * u + v * v));
  double C = sqrt(c2);
  double dW1 = -0.5 * (rho * (uR - uL)) / C + 0.5 * (pR - pL) / c2;
  double dW2 = (rhoR - rhoL) - (pR - pL) / c2;
  double dW3 = rho * (vR - vL);
  double dW4 = 0.5 * (rho * (uR - uL)) / C + 0.5 * (pR - pL) / c2;
  dW1 = fabs(u - C) * dW1;
  dW2 = fabs(u) * dW2;
  dW3 = fabs(u) * dW3;
  dW4 = fabs(u + C) * dW4;
  double f0 = 0.5 * (rhoULr + rhoURr);
  double f1 = 0.5 * (rhoULr * uL + rhoURr * uR + pL + pR);
  double f2 = 0.5 * (rhoVLr * uL + rhoVRr * uR);
  double f3 = 0.5 * ((pL + QL[3]) * uL + (pR + QR[3]) * uR);
  f0 -= 0.5 * (dW1 + dW2 + dW4);
  f1 -= 0.5 * (dW1 * (u - C) + dW2 * u + dW4 * (u + C));
  f2 -= 0.5 * (dW1 * v + dW2 * v + dW3 + dW4 * v);
  f3 -= 0.5 * (dW1 * (h - u * C) + 0.5 * dW2 * (u * u + v * v) + dW3 * v + dW4 * (h + u * C));
  F[0] = f0;
  F[1] = n[0] * f1 - n[1] * f2;
  F[2] = n[1] * f1 + n[0] * f2;
  F[3] = f3;
}

/* RoeERFlux (fluxes.go:415-503), including the (dPu+dPp)*ny energy term */
static void roeer_flux(double g, const double QL[4], const double QR[4], const double n[2], double F[4]) {
  double GM1 = g - 1, nx = n[0], ny = n[1];
  double rhoL = QL[0], rhoR = QR[0];
  double ooRhoL = 1. / rhoL, ooRhoR = 1. / rhoR;
  double rhoLs = sqrt(rhoL), rhoRs = sqrt(rhoR);
  double uL = QL[1] * ooRhoL, vL = QL[2] * ooRhoL, uR = QR[1] * ooRhoR, vR = QR[2] * ooRhoR;
  double EL = QL[3], ER = QR[3];
  double UL = nx * uL + ny * vL, UR = nx * uR + ny * vR;
  double pL = fs_pressure(g, QL[0], QL[1], QL[2], QL[3]), pR = fs_pressure(g, QR[0], QR[1], QR[2], QR[3]);
  double HL = EL + pL, HR = ER + pR;
  double ooRs = 1. / (rhoLs + rhoRs);
  double u = (rhoLs * uL + rhoRs * uR) * ooRs, v = (rhoLs * vL + rhoRs * vR) * ooRs, h = (HL + HR) * ooRs;
  double rho = rhoLs * rhoRs, H = h * rho, U = nx * u + ny * v;
  double C2 = GM1 * (h - 0.5 * (u * u + v * v));
  double C = sqrt(C2), ooC = 1. / C, Uabs = fabs(U);
  F[0] = 0.5 * (UL * rhoL + UR * rhoR);
  F[1] = 0.5 * (UL * rhoL * uL + pL * nx + UR * rhoR * uR + pR * nx);
  F[2] = 0.5 * (UL * rhoL * vL + pL * ny + UR * rhoR * vR + pR * ny);
  F[3] = 0.5 * (UL * HL + UR * HR);
  double Uef = 0.05 * C;
  double du = uR - uL, dv = vR - vL, deltaV2 = du * du + dv * dv;
  double ooVmag = 1. / sqrt(u * u + v * v);
  double n1x, n1y;
  if (deltaV2 < 0.01 * C2) { n1x = nx; n1y = ny; } else { n1x = ooVmag * du; n1y = ooVmag * dv; }
  double n2x = n1y * (nx * n1y - n1x * ny), n2y = -n1x * (nx * n1y - n1x * ny);
  double alp1 = nx * n1x + ny * n1y, alp2 = nx * n2x + ny * n2y;
  double U1x = n1x * u, U1y = n1y * v, U2x = n2x * u, U2y = n2y * v;
  double Urot = sqrt(alp1 * alp1 * (U1x * U1x + U1y * U1y)) + sqrt(alp2 * alp2 * (U2x * U2x + U2y * U2y));
  double sigma = fmax(Uabs, fmin(Uef, Urot));
  /* sign = math.Copysign (fluxes.go:426): sign(x, 1) is |x|, not the sign of x */
  double sgnp = fabs(U + C), sgnm = fabs(U - C);
  double UabsPrime = Uabs - 0.25 * fmax(0, UR - UL) * (sgnp - sgnm);
  double dU = UR - UL, dP = pR - pL, dRho = rhoR - rhoL, dRhoU = rhoR * uR - rhoL * uL;
  double dRhoV = rhoR * vR - rhoL * vL, dE = ER - EL;
  double dPu = rho * dU * fmax(0, C - UabsPrime);
  double swt = fabs(U) * fmin(UabsPrime, C);
  double dPp = swt * dP * ooC, dUu = swt * dU * ooC;
  F[0] -= 0.5 * (sigma * dRho + (dPu + dPp) * 0 + dUu * rho);
  F[1] -= 0.5 * (sigma * dRhoU + (dPu + dPp) * nx + dUu * rho * u);
  F[2] -= 0.5 * (sigma * dRhoV + (dPu + dPp) * ny + dUu * rho * v);
  F[3] -= 0.5 * (sigma * dE + (dPu + dPp) * ny + dUu * H);
}

void or_numerical_flux(int flux, double g, const double *QL, const double *QR, const double *n, double *F) {
  switch (flux) {
  case OR_FLUX_AVERAGE: avg_flux(g, QL, QR, n, F); break;
  case OR_FLUX_LAX: lax_flux(g, QL, QR, n, F); break;
  case OR_FLUX_ROE: roe_flux(g, QL, QR, n, F); break;
  default: roeer_flux(g, QL, QR, n, F); break;
  }
}

/* RiemannBC (bcs.go:70-133). pInf and CInf come from the freestream object
 * even when QInf is the analytic vortex state (bcs.go:85-86). */
static void riemann_bc(const fs_t *FS, const double Qi[4], const double QInf[4], const double n[2],
                       double Q[4]) {
  double rhoInt = Qi[0], uInt = Qi[1] / Qi[0], vInt = Qi[2] / Qi[0];
  double pInt = fs_pressure(FS->gamma, Qi[0], Qi[1], Qi[2], Qi[3]);
  double CInt = fs_sound(FS->gamma, Qi[0], Qi[1], Qi[2], Qi[3]);
  double pInf = FS->Pinf, CInf = FS->Cinf, Gamma = FS->gamma, GM1 = Gamma - 1., OOGM1 = 1. / GM1;
  double rhoInf = QInf[0], uInf = QInf[1] / QInf[0], vInf = QInf[2] / QInf[0];
  double tangent[2] = {-n[1], n[0]};
  double VnormInt = n[0] * uInt + n[1] * vInt;
  if (FS->Minf <= 1.) {
    double VnormInf = n[0] * uInf + n[1] * vInf;
    double Rinf = VnormInf - 2. * CInf * OOGM1;
    double Rint = VnormInt + 2. * CInt * OOGM1;
    double Vnorm = 0.5 * (Rint + Rinf);
    double C = 0.25 * GM1 * (Rint - Rinf);
    double Vtang, Beta;
    if (VnormInt < 0) {
      Vtang = tangent[0] * uInf + tangent[1] * vInf;
      Beta = pInf / pow(rhoInf, Gamma);
    } else {
      Vtang = tangent[0] * uInt + tangent[1] * vInt;
      Beta = pInt / pow(rhoInt, Gamma);
    }
    double u = Vnorm * n[0] + Vtang * tangent[0];
    double v = Vnorm * n[1] + Vtang * tangent[1];
    double rho = pow(C * C / (Gamma * Beta), OOGM1);
    double p = Beta * pow(rho, Gamma);
    Q[0] = rho; Q[1] = rho * u; Q[2] = rho * v; Q[3] = p * OOGM1 + 0.5 * rho * (u * u + v * v);
  } else {
    if (VnormInt < 0) memcpy(Q, QInf, sizeof(double) * 4);
    else memcpy(Q, Qi, sizeof(double) * 4);
  }
}

void or_riemann_bc(double gamma, double minf, double pinf, double cinf, const double *Qint,
                   const double *Qinf, const double *n, double *Qbc) {
  fs_t fs;
  memset(&fs, 0, sizeof fs);
  fs.gamma = gamma; fs.Minf = minf; fs.Pinf = pinf; fs.Cinf = cinf;
  memcpy(fs.Qinf, Qinf, sizeof(double) * 4);
  riemann_bc(&fs, Qint, Qinf, n, Qbc);
}

/* ------------------------------------------------------------------------ */
/* Setup: NewEuler (euler.go:61-136)                                         */
/* ------------------------------------------------------------------------ */

static int setup_topology(or_solver *s, const or_mesh *m, char *err, int errlen) {
  int K = s->K;
  face_rec *fr = malloc(sizeof(face_rec) * 3 * (size_t)K);
  /* NewTriangulation :30-38; NewEdge :114-147 */
  for (int k = 0; k < K; k++) {
    int v[3] = {s->EToV[3 * k], s->EToV[3 * k + 1], s->EToV[3 * k + 2]};
    for (int e = 0; e < 3; e++) {
      int a = v[e], b = v[(e + 1) % 3];
      face_rec *f = &fr[3 * k + e];
      f->key = edge_key(a, b);
      f->seq = 3 * k + e;
      f->k = k; f->en = e; f->dir = (a > b);
    }
  }
  qsort(fr, 3 * (size_t)K, sizeof(face_rec), cmp_face);
  s->E = calloc(3 * (size_t)K, sizeof(edge_t));
  s->EtoEdge = malloc(sizeof(int) * 3 * K);
  int ne = 0;
  for (int i = 0; i < 3 * K; i++) {
    if (i == 0 || fr[i].key != fr[i - 1].key) {
      edge_t *E = &s->E[ne++];
      E->key = fr[i].key;
      E->nconn = 0;
      E->bc = OR_BC_NONE;
    }
    edge_t *E = &s->E[ne - 1];
    if (E->nconn > 1) { seterr(err, errlen, "incorrect edge construction, more than two connected triangles"); free(fr); return -1; }
    int c = E->nconn++;
    E->tri[c] = fr[i].k; E->enr[c] = fr[i].en; E->dir[c] = fr[i].dir;
    E->IInII[c] = edge_iinii(E->key, fr[i].dir, fr[i].en, s->VX, s->VY, NULL, NULL);
    s->EtoEdge[3 * fr[i].k + fr[i].en] = ne - 1;
  }
  free(fr);
  s->ne = ne;
  /* EtoE :39-51 */
  s->EtoE = malloc(sizeof(int) * 3 * K);
  for (int i = 0; i < 3 * K; i++) s->EtoE[i] = -1;
  for (int i = 0; i < ne; i++)
    if (s->E[i].nconn == 2) {
      s->EtoE[3 * s->E[i].tri[0] + s->E[i].enr[0]] = s->E[i].tri[1];
      s->EtoE[3 * s->E[i].tri[1] + s->E[i].enr[1]] = s->E[i].tri[0];
    }
  /* BCs :52-104, tags in file order */
  int reversedLast = 0;
  for (int t = 0; t < m->ntags; t++) {
    const bctag_t *b = &m->tags[t];
    int flag = bc_flag(b->tag);
    if (flag < 0) { seterr(err, errlen, "unable to find BC with base name: [%s]", b->tag); return -1; }
    switch (flag) {
    case OR_BC_FAR: case OR_BC_IVORTEX: case OR_BC_WALL: case OR_BC_IN: case OR_BC_OUT: case OR_BC_CYL:
      for (int i = 0; i < b->n; i++) {
        int ei = find_edge(s->E, ne, edge_key(b->e[i].v0, b->e[i].v1));
        if (ei < 0) { seterr(err, errlen, "BC edge not in mesh"); return -1; }
        s->E[ei].bc = flag;
      }
      break;
    case OR_BC_PERIODIC_REVERSED:
      reversedLast = 1;
      /* fallthrough */
    case OR_BC_PERIODIC: {
      int l = b->n;
      if (l % 2) { seterr(err, errlen, "periodic boundaries must be present in pairs, have dimension %d, not even", l); return -1; }
      int l2 = l / 2;
      edgeint_t *c0 = malloc(sizeof(edgeint_t) * l2), *c1 = malloc(sizeof(edgeint_t) * l2);
      memcpy(c0, b->e, sizeof(edgeint_t) * l2);
      memcpy(c1, b->e + l2, sizeof(edgeint_t) * l2);
      if (curve_reorder(c0, l2, 0) || curve_reorder(c1, l2, reversedLast)) {
        seterr(err, errlen, "unable to construct contiguous curve from periodic edges");
        free(c0); free(c1); return -1;
      }
      for (int i = 0; i < l2; i++) {
        int e1 = find_edge(s->E, ne, edge_key(c0[i].v0, c0[i].v1));
        int e2 = find_edge(s->E, ne, edge_key(c1[i].v0, c1[i].v1));
        if (e1 < 0 || e2 < 0) { seterr(err, errlen, "periodic edge not in mesh"); free(c0); free(c1); return -1; }
        edge_t *E1 = &s->E[e1], *E2 = &s->E[e2];
        E1->bc = flag; E2->bc = flag;
        if (E1->nconn != 1) { seterr(err, errlen, "periodic edge already connected"); free(c0); free(c1); return -1; }
        /* ee1.AddTri(e2.key, k2, 1, en2, dir) :95-98 */
        E1->tri[1] = E2->tri[0]; E1->enr[1] = E2->enr[0]; E1->dir[1] = E2->dir[0];
        E1->IInII[1] = edge_iinii(E2->key, E2->dir[0], E2->enr[0], s->VX, s->VY, NULL, NULL);
        E1->nconn = 2;
      }
      free(c0); free(c1);
      break;
    }
    default:
      seterr(err, errlen, "BC type for tag %s not implemented yet", b->tag);
      return -1;
    }
  }
  return 0;
}

/* CalculateJacobian (dfr_startup.go:256-287), CalculateFaceNorms (:289-314),
 * GetEdgeLengths (:125-147), CalculateElementLocalGeometry (element_utils.go:163-175).
 * Periodic slot resolution (the reference's last writer is map-order dependent,
 * dfr_startup.go:296-309): EdgeNumber(k2,e2) = the coupled edge, FaceNorm(k2,e2)
 * = k2's own outward normal, IInII identical either way. */
static void setup_geometry(or_solver *s) {
  int K = s->K, NpInt = s->NpInt, NpFlux = s->NpFlux;
  s->Jdet = malloc(sizeof(double) * K); s->Jinv = malloc(sizeof(double) * 4 * K);
  for (int k = 0; k < K; k++) {
    int *v = &s->EToV[3 * k];
    double v1x = s->VX[v[0]], v2x = s->VX[v[1]], v3x = s->VX[v[2]];
    double v1y = s->VY[v[0]], v2y = s->VY[v[1]], v3y = s->VY[v[2]];
    double xr = 0.5 * (v2x - v1x), yr = 0.5 * (v2y - v1y);
    double xs = 0.5 * (v3x - v1x), ys = 0.5 * (v3y - v1y);
    double det = xr * ys - xs * yr;
    s->Jdet[k] = det;
    s->Jinv[4 * k] = ys / det; s->Jinv[4 * k + 1] = -xs / det;
    s->Jinv[4 * k + 2] = -yr / det; s->Jinv[4 * k + 3] = xr / det;
  }
  s->FNx = calloc(3 * K, sizeof(double)); s->FNy = calloc(3 * K, sizeof(double));
  s->IInII = calloc(3 * K, sizeof(double));
  s->FaceEdge = malloc(sizeof(int) * 3 * K);
  for (int i = 0; i < 3 * K; i++) s->FaceEdge[i] = -1;
  /* pass 1: every edge's own connections (conn 0 always; conn 1 of regular shared edges) */
  for (int ei = 0; ei < s->ne; ei++) {
    edge_t *E = &s->E[ei];
    for (int c = 0; c < E->nconn; c++) {
      int periodic_partner = (c == 1) && (E->bc == OR_BC_PERIODIC || E->bc == OR_BC_PERIODIC_REVERSED);
      if (periodic_partner) continue;
      int k = E->tri[c], en = E->enr[c], ind = k + K * en;
      s->IInII[ind] = E->IInII[c];
      s->FaceEdge[ind] = ei;
      edge_iinii(E->key, E->dir[c], en, s->VX, s->VY, &s->FNx[ind], &s->FNy[ind]);
    }
  }
  /* pass 2: periodic coupling, EdgeNumber of the partner face -> coupled edge */
  for (int ei = 0; ei < s->ne; ei++) {
    edge_t *E = &s->E[ei];
    if (E->nconn == 2 && (E->bc == OR_BC_PERIODIC || E->bc == OR_BC_PERIODIC_REVERSED)) {
      int ind = E->tri[1] + K * E->enr[1];
      s->FaceEdge[ind] = ei;
      s->IInII[ind] = E->IInII[1];
    }
  }
  s->EdgeLenMax = malloc(sizeof(double) * K);
  for (int k = 0; k < K; k++) {
    double mx = -1;
    for (int e = 0; e < 3; e++) if (s->IInII[k + K * e] > mx) mx = s->IInII[k + K * e];
    s->EdgeLenMax[k] = s->Jdet[k] / mx;
  }
  s->SolX = malloc(sizeof(double) * NpInt * K); s->SolY = malloc(sizeof(double) * NpInt * K);
  s->FluxX = malloc(sizeof(double) * NpFlux * K); s->FluxY = malloc(sizeof(double) * NpFlux * K);
  for (int k = 0; k < K; k++) {
    int *v = &s->EToV[3 * k];
    for (int i = 0; i < NpFlux; i++) {
      double r = s->o.R[i], sv = s->o.S[i];
      double x = (-(r + sv) * s->VX[v[0]] + (r + 1) * s->VX[v[1]] + (sv + 1) * s->VX[v[2]]) * 0.5;
      double y = (-(r + sv) * s->VY[v[0]] + (r + 1) * s->VY[v[1]] + (sv + 1) * s->VY[v[2]]) * 0.5;
      s->FluxX[k + i * K] = x; s->FluxY[k + i * K] = y;
    }
    for (int i = 0; i < NpInt; i++) {
      double r = s->o.solR[i], sv = s->o.solS[i];
      s->SolX[k + i * K] = (-(r + sv) * s->VX[v[0]] + (r + 1) * s->VX[v[1]] + (sv + 1) * s->VX[v[2]]) * 0.5;
      s->SolY[k + i * K] = (-(r + sv) * s->VY[v[0]] + (r + 1) * s->VY[v[1]] + (sv + 1) * s->VY[v[2]]) * 0.5;
    }
  }
}

static void setup_shock(or_solver *s);

or_solver *or_create(const char *meshfile, const or_params *p, char *err, int errlen) {
  or_mesh *m = or_mesh_read(meshfile, err, errlen);
  if (!m) return NULL;
  or_solver *s = calloc(1, sizeof(or_solver));
  s->nan_scan = 1;
  s->p = *p;
  if (p->flux < 0 || p->flux > 3) { seterr(err, errlen, "unable to use flux type %d", p->flux); goto fail; }
  if (build_ops(p->N, &s->o, err, errlen)) goto fail;
  s->K = m->K; s->Nv = m->Nv;
  s->NpInt = s->o.NpInt; s->NpEdge = s->o.NpEdge; s->NpFlux = s->o.NpFlux;
  s->VX = malloc(sizeof(double) * m->Nv); s->VY = malloc(sizeof(double) * m->Nv);
  s->EToV = malloc(sizeof(int) * 3 * m->K);
  or_mesh_get(m, s->VX, s->VY, s->EToV);
  if (setup_topology(s, m, err, errlen)) goto fail;
  setup_geometry(s);
  int K = s->K, NpInt = s->NpInt, NpEdge = s->NpEdge, NpFlux = s->NpFlux;
  size_t nq = 4 * (size_t)NpInt * K;
  s->Q = calloc(nq, sizeof(double)); s->Q1 = calloc(nq, sizeof(double)); s->Q2 = calloc(nq, sizeof(double));
  s->Q3 = calloc(nq, sizeof(double)); s->Q4 = calloc(nq, sizeof(double)); s->RHSQ = calloc(nq, sizeof(double));
  s->Resid = calloc(nq, sizeof(double));
  s->DT = calloc(K, sizeof(double)); s->DTVisc = calloc(K, sizeof(double));
  s->QFace = calloc(4 * (size_t)3 * NpEdge * K, sizeof(double));
  s->FRT = calloc(4 * (size_t)NpFlux * K, sizeof(double));
  s->EdgeFlux = calloc(4 * (size_t)NpEdge * s->ne, sizeof(double));
  s->EdgeQ = calloc(4 * (size_t)NpEdge * s->ne, sizeof(double));
  s->EdgeVisc = calloc(4 * (size_t)NpEdge * s->ne, sizeof(double));
  s->EdgeAgg = calloc(s->ne, sizeof(double)); s->EdgeAggVisc = calloc(s->ne, sizeof(double));
  /* InitializeSolution (euler.go:728-794) */
  fs_new(&s->fs, p->Minf, p->Gamma, p->Alpha);
  s->fsin = s->fs;
  s->fsout = s->fs;
  if (p->init == OR_INIT_FREESTREAM) {
    for (int n = 0; n < 4; n++)
      for (size_t i = 0; i < (size_t)NpInt * K; i++) s->Q[n * (size_t)NpInt * K + i] = 0.0 + s->fs.Qinf[n];
  } else if (p->init == OR_INIT_IVORTEX) {
    double q[4] = {1, 1, 0, 3};
    fs_from_qinf(&s->fs, 1.4, q);
    s->fs.Minf = 0; s->fs.Alpha = 0;
    s->ivortex = 1;
    for (size_t i = 0; i < (size_t)NpInt * K; i++) {
      double qq[4];
      ivortex_state(0, s->SolX[i], s->SolY[i], qq);
      for (int n = 0; n < 4; n++) s->Q[n * (size_t)NpInt * K + i] = qq[n];
    }
    for (int e = 0; e < s->ne; e++) if (s->E[e].bc == OR_BC_WALL) s->E[e].bc = OR_BC_IVORTEX;
  } else if (p->init == OR_INIT_SHOCKTUBE) {
    /* InitializeSolution SHOCKTUBE (euler.go:738-771): Sod states left and
     * right of x = 0.5, which are also the In and Out boundary states */
    double gam = 1.4, qin[4] = {1, 0, 0, 1 / (gam - 1)}, qout[4] = {0.125, 0, 0, 0.1 / (gam - 1)};
    fs_from_qinf(&s->fsin, gam, qin);
    s->fsin.Minf = 0; s->fsin.Alpha = 0;
    fs_from_qinf(&s->fsout, gam, qout);
    s->fsout.Minf = 0; s->fsout.Alpha = 0;
    for (size_t i = 0; i < (size_t)NpInt * K; i++) {
      const double *qq = s->SolX[i] < 0.5 ? s->fsin.Qinf : s->fsout.Qinf;
      for (int n = 0; n < 4; n++) s->Q[n * (size_t)NpInt * K + i] = qq[n];
    }
  } else {
    seterr(err, errlen, "init type %d not supported by the oracle", p->init);
    goto fail;
  }
  if (p->limiter == 1 && p->N != 0) setup_shock(s);
  or_mesh_free(m);
  return s;
fail:
  or_mesh_free(m);
  or_destroy(s);
  return NULL;
}

void or_destroy(or_solver *s) {
  if (!s) return;
  free_ops(&s->o);
  void *ptrs[] = {s->VX, s->VY, s->EToV, s->E, s->EtoE, s->EtoEdge, s->FaceEdge, s->Jdet, s->Jinv,
                  s->FNx, s->FNy, s->IInII, s->EdgeLenMax, s->SolX, s->SolY, s->FluxX, s->FluxY,
                  s->Q, s->Q1, s->Q2, s->Q3, s->Q4, s->RHSQ, s->Resid, s->DT, s->DTVisc, s->QFace,
                  s->FRT, s->EdgeFlux, s->EdgeQ, s->EdgeAgg, s->EdgeAggVisc, s->EdgeVisc, s->Se,
                  s->Sigma, s->Eps, s->SigmaV, s->EpsV, s->EpsF, s->SigmaF, s->DissX, s->DissY,
                  s->Clip, s->Dmat, s->Pmat, s->Bary};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
  free(s);
}

int or_K(const or_solver *s) { return s->K; }
int or_NpInt(const or_solver *s) { return s->NpInt; }
int or_NpEdge(const or_solver *s) { return s->NpEdge; }
int or_NpFlux(const or_solver *s) { return s->NpFlux; }
int or_num_edges(const or_solver *s) { return s->ne; }
void or_get_state(const or_solver *s, double *Q) { memcpy(Q, s->Q, sizeof(double) * 4 * s->NpInt * (size_t)s->K); }
void or_set_state(or_solver *s, const double *Q) { memcpy(s->Q, Q, sizeof(double) * 4 * s->NpInt * (size_t)s->K); }
/* NaN scan before each stage: 0 none, 1 stage input, q0 and DT (the
 * default), 2 the reference's per-stage list (euler.go:471-484) */
void or_set_nan_scan(or_solver *s, int mode) { s->nan_scan = mode; }
double or_time(const or_solver *s) { return s->time; }
double or_dt(const or_solver *s) { return s->gdt; }
int or_step_count(const or_solver *s) { return s->steps; }
void or_get_residual(const or_solver *s, double *R) { memcpy(R, s->Resid, sizeof(double) * 4 * s->NpInt * (size_t)s->K); }

void or_get_flux_dofs(const or_solver *s, double *F) {
  memcpy(F, s->FRT, sizeof(double) * 4 * (size_t)s->NpFlux * s->K);
}

void or_get_operators(const or_solver *s, double *DivInt, double *Div, double *Interp, double *R, double *S) {
  int NpInt = s->NpInt, NpFlux = s->NpFlux, NpEdge = s->NpEdge;
  if (DivInt) memcpy(DivInt, s->o.DivInt, sizeof(double) * NpInt * NpFlux);
  if (Div) memcpy(Div, s->o.Div, sizeof(double) * NpFlux * NpFlux);
  if (Interp) memcpy(Interp, s->o.Interp, sizeof(double) * 3 * NpEdge * NpInt);
  if (R) memcpy(R, s->o.R, sizeof(double) * NpFlux);
  if (S) memcpy(S, s->o.S, sizeof(double) * NpFlux);
}
void or_get_solution_element(const or_solver *s, double *V, double *Vinv, double *Mass, double *R, double *S, double *W) {
  int n = s->NpInt;
  if (V) memcpy(V, s->o.V, sizeof(double) * n * n);
  if (Vinv) memcpy(Vinv, s->o.Vinv, sizeof(double) * n * n);
  if (Mass) memcpy(Mass, s->o.Mass, sizeof(double) * n * n);
  if (R) memcpy(R, s->o.solR, sizeof(double) * n);
  if (S) memcpy(S, s->o.solS, sizeof(double) * n);
  if (W) memcpy(W, s->o.solW, sizeof(double) * n);
}
void or_get_geometry(const or_solver *s, double *Jdet, double *Jinv, double *FaceNx, double *FaceNy, double *IInII, double *EdgeLenMax) {
  int K = s->K;
  if (Jdet) memcpy(Jdet, s->Jdet, sizeof(double) * K);
  if (Jinv) memcpy(Jinv, s->Jinv, sizeof(double) * 4 * K);
  if (FaceNx) memcpy(FaceNx, s->FNx, sizeof(double) * 3 * K);
  if (FaceNy) memcpy(FaceNy, s->FNy, sizeof(double) * 3 * K);
  if (IInII) memcpy(IInII, s->IInII, sizeof(double) * 3 * K);
  if (EdgeLenMax) memcpy(EdgeLenMax, s->EdgeLenMax, sizeof(double) * K);
}
void or_get_coords(const or_solver *s, double *SolX, double *SolY, double *FluxX, double *FluxY) {
  size_t a = (size_t)s->NpInt * s->K, b = (size_t)s->NpFlux * s->K;
  if (SolX) memcpy(SolX, s->SolX, sizeof(double) * a);
  if (SolY) memcpy(SolY, s->SolY, sizeof(double) * a);
  if (FluxX) memcpy(FluxX, s->FluxX, sizeof(double) * b);
  if (FluxY) memcpy(FluxY, s->FluxY, sizeof(double) * b);
}
void or_get_edges(const or_solver *s, long long *o) {
  for (int i = 0; i < s->ne; i++) {
    const edge_t *E = &s->E[i];
    long long *r = o + 10 * i;
    r[0] = (long long)(E->key & 0xffffffffu); r[1] = (long long)(E->key >> 32); r[2] = E->nconn;
    r[3] = E->tri[0]; r[4] = E->nconn > 1 ? E->tri[1] : -1;
    r[5] = E->enr[0]; r[6] = E->nconn > 1 ? E->enr[1] : -1;
    r[7] = E->dir[0]; r[8] = E->nconn > 1 ? E->dir[1] : -1; r[9] = E->bc;
  }
}
void or_get_connectivity(const or_solver *s, int *EtoE, int *FaceEdge) {
  if (EtoE) memcpy(EtoE, s->EtoE, sizeof(int) * 3 * s->K);
  if (FaceEdge) memcpy(FaceEdge, s->FaceEdge, sizeof(int) * 3 * s->K);
}

/* ------------------------------------------------------------------------ */
/* Right-hand side pieces                                                    */
/* ------------------------------------------------------------------------ */

/* InterpolateSolutionToEdges (edges.go:485-491) */
static void interp_to_edges(or_solver *s, const double *Q) {
  int K = s->K, NpInt = s->NpInt, ne = 3 * s->NpEdge;
  for (int n = 0; n < 4; n++) {
    const double *q = Q + (size_t)n * NpInt * K;
    double *qf = s->QFace + (size_t)n * ne * K;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < ne; p++)
      for (int k = 0; k < K; k++) {
        double acc = 0;
        for (int i = 0; i < NpInt; i++) acc += s->o.Interp[p * NpInt + i] * q[k + (size_t)i * K];
        qf[k + (size_t)p * K] = acc;
      }
  }
}

static inline double *QF(or_solver *s, int n, int p, int k) {
  return &s->QFace[(size_t)n * 3 * s->NpEdge * s->K + (size_t)p * s->K + k];
}

/* CalculateEdgeEulerFlux (edges.go:325-377) with calculateSharedEdgeFlux
 * (:379-411), calculateNonSharedEdgeFlux (:413-452), WallBC/IVortexBC/FarBC
 * (bcs.go:11-68). Boundary conditions overwrite the owner's face trace. */
static void edge_fluxes(or_solver *s, double t) {
  int K = s->K, NpEdge = s->NpEdge, NpInt = s->NpInt;
  double g = s->fs.gamma;
#pragma omp parallel for schedule(static)
  for (int ei = 0; ei < s->ne; ei++) {
    edge_t *E = &s->E[ei];
    int kL = E->tri[0], enL = E->enr[0], shiftL = enL * NpEdge;
    double nrm[2] = {s->FNx[kL + K * enL], s->FNy[kL + K * enL]};
    double *EF = s->EdgeFlux + (size_t)ei * NpEdge * 4;
    double *EQ = s->EdgeQ + (size_t)ei * NpEdge * 4;
    for (int i = 0; i < NpEdge; i++)
      for (int n = 0; n < 4; n++) EQ[4 * i + n] = *QF(s, n, i + shiftL, kL);
    for (int i = 0; i < 4 * NpEdge; i++) EF[i] = 0;
    if (E->nconn == 2) {
      int kR = E->tri[1], shiftR = E->enr[1] * NpEdge;
      for (int i = 0; i < NpEdge; i++) {
        double QL[4], QR[4];
        for (int n = 0; n < 4; n++) {
          QL[n] = *QF(s, n, i + shiftL, kL);
          QR[n] = *QF(s, n, NpEdge - 1 - i + shiftR, kR);
        }
        or_numerical_flux(s->p.flux, g, QL, QR, nrm, EF + 4 * i);
      }
      continue;
    }
    int normal_flux = 1;
    switch (E->bc) {
    case OR_BC_NONE: break;
    case OR_BC_FAR: case OR_BC_IN: case OR_BC_OUT: {
      /* FarBC with FSFar, FSIn or FSOut (edges.go:425-431); FSIn/FSOut only
       * exist for the shocktube case, elsewhere In/Out use FSFar where the
       * reference would dereference nil */
      const fs_t *FS = E->bc == OR_BC_IN ? &s->fsin : E->bc == OR_BC_OUT ? &s->fsout : &s->fs;
      for (int i = 0; i < NpEdge; i++) {
        double Qi[4], Qb[4];
        for (int n = 0; n < 4; n++) Qi[n] = *QF(s, n, i + shiftL, kL);
        riemann_bc(FS, Qi, FS->Qinf, nrm, Qb);
        for (int n = 0; n < 4; n++) *QF(s, n, i + shiftL, kL) = Qb[n];
      }
      break;
    }
    case OR_BC_IVORTEX:
      for (int i = 0; i < NpEdge; i++) {
        int indFlux = kL + (2 * NpInt + i + shiftL) * K;
        double qa[4], Qi[4], Qb[4];
        ivortex_state(t, s->FluxX[indFlux], s->FluxY[indFlux], qa);
        for (int n = 0; n < 4; n++) Qi[n] = *QF(s, n, i + shiftL, kL);
        riemann_bc(&s->fs, Qi, qa, nrm, Qb);
        for (int n = 0; n < 4; n++) *QF(s, n, i + shiftL, kL) = Qb[n];
      }
      break;
    case OR_BC_WALL: case OR_BC_CYL:
      normal_flux = 0;
      for (int i = 0; i < NpEdge; i++) {
        double p = fs_pressure(g, *QF(s, 0, i + shiftL, kL), *QF(s, 1, i + shiftL, kL),
                               *QF(s, 2, i + shiftL, kL), *QF(s, 3, i + shiftL, kL));
        EF[4 * i] = 0; EF[4 * i + 1] = nrm[0] * p; EF[4 * i + 2] = nrm[1] * p; EF[4 * i + 3] = 0;
      }
      break;
    case OR_BC_PERIODIC: case OR_BC_PERIODIC_REVERSED:
      normal_flux = 0;
      break;
    default: break;
    }
    if (normal_flux)
      for (int i = 0; i < NpEdge; i++) {
        double q[4], Fx[4], Fy[4];
        for (int n = 0; n < 4; n++) q[n] = *QF(s, n, i + shiftL, kL);
        flux_xy(g, q, Fx, Fy);
        for (int n = 0; n < 4; n++) EF[4 * i + n] = nrm[0] * Fx[n] + nrm[1] * Fy[n];
      }
  }
}

/* SetRTFluxInternal (euler.go:701-726) with CalculateFluxTransformed (fluxes.go:53-69) */
static void rt_flux_internal(or_solver *s, const double *Q) {
  int K = s->K, NpInt = s->NpInt, NpFlux = s->NpFlux;
  double g = s->fs.gamma;
  memset(s->FRT, 0, sizeof(double) * 4 * (size_t)NpFlux * K);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++) {
    double Jd = s->Jdet[k], *Ji = &s->Jinv[4 * k];
    for (int i = 0; i < NpInt; i++) {
      double q[4], Fx[4], Fy[4];
      for (int n = 0; n < 4; n++) q[n] = Q[(size_t)n * NpInt * K + k + (size_t)i * K];
      flux_xy(g, q, Fx, Fy);
      for (int n = 0; n < 4; n++) {
        double *F = s->FRT + (size_t)n * NpFlux * K;
        F[k + (size_t)i * K] = Jd * (Ji[0] * Fx[n] + Ji[1] * Fy[n]);
        F[k + (size_t)(i + NpInt) * K] = Jd * (Ji[2] * Fx[n] + Ji[3] * Fy[n]);
      }
    }
  }
}

/* SetRTFluxOnEdges (edges.go:454-483) with GetEdgeValues (:93-113) */
static void rt_flux_edges(or_solver *s) {
  int K = s->K, NpInt = s->NpInt, NpEdge = s->NpEdge, NpFlux = s->NpFlux;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++)
    for (int e = 0; e < 3; e++) {
      int ind2 = k + K * e;
      double IInII = s->IInII[ind2];
      int ei = s->FaceEdge[ind2];
      const double *EF = s->EdgeFlux + (size_t)ei * NpEdge * 4;
      int sign = (s->E[ei].tri[0] == k) ? 1 : -1;
      for (int n = 0; n < 4; n++) {
        double *F = s->FRT + (size_t)n * NpFlux * K;
        for (int i = 0; i < NpEdge; i++) {
          size_t ind = k + (size_t)(2 * NpInt + i + e * NpEdge) * K;
          if (sign > 0) F[ind] = EF[4 * i + n] * IInII;
          else F[ind] = -EF[4 * (NpEdge - i - 1) + n] * IInII;
        }
      }
    }
}

/* RHSInternalPoints (euler.go:665-699) */
static void rhs_internal(or_solver *s, double *RHS) {
  int K = s->K, NpInt = s->NpInt, NpFlux = s->NpFlux;
  for (int n = 0; n < 4; n++) {
    const double *F = s->FRT + (size_t)n * NpFlux * K;
    double *R = RHS + (size_t)n * NpInt * K;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < NpInt; i++)
      for (int k = 0; k < K; k++) {
        double acc = 0;
        for (int j = 0; j < NpFlux; j++) acc += s->o.DivInt[i * NpFlux + j] * F[k + (size_t)j * K];
        R[k + (size_t)i * K] = acc;
      }
    for (int k = 0; k < K; k++) {
      double oojd = 1. / s->Jdet[k];
      for (int i = 0; i < NpInt; i++) R[k + (size_t)i * K] *= -oojd;
    }
  }
}

/* StoreEdgeAggregates (edges.go:246-289) + GetHk (dfr_startup.go:89-98) */
static void edge_aggregates(or_solver *s) {
  int K = s->K, NpEdge = s->NpEdge, NpInt = s->NpInt;
  double Np12 = (double)((s->p.N + 1) * (s->p.N + 1));
#pragma omp parallel for schedule(static)
  for (int ei = 0; ei < s->ne; ei++) {
    edge_t *E = &s->E[ei];
    int k = E->tri[0], shift = E->enr[0] * NpEdge;
    double hK = s->EdgeLenMax[k] / Np12, oohK = 1. / hK;
    double mx = -DBL_MAX;
    for (int i = shift; i < shift + NpEdge; i++) {
      double q0 = *QF(s, 0, i, k), q1 = *QF(s, 1, i, k), q2 = *QF(s, 2, i, k), q3 = *QF(s, 3, i, k);
      double C = fs_sound(s->fs.gamma, q0, q1, q2, q3);
      double U = fs_velocity(q0, q1, q2);
      double w = oohK * (U + C);
      if (w > mx) mx = w;
    }
    s->EdgeAgg[ei] = mx;
    if (s->shock) {
      double mv = -DBL_MAX;
      int off = 2 * NpInt + shift;
      for (int i = off; i < off + NpEdge; i++) {
        double v = oohK * oohK * s->EpsF[k + (size_t)i * K];
        if (v > mv) mv = v;
      }
      s->EdgeAggVisc[ei] = mv;
    }
  }
}

static void shock_stage_pre(or_solver *s, double *QQQ, int rkStep);
static void shock_gradient_and_viscous_edges(or_solver *s, const double *QQQ);
static void shock_add_dissipation(or_solver *s, double *RHS);
static void limit_filter(or_solver *s, double *U, int nvars);

/* RHS of a state at time t. With the shock path on, this is the stage RHS
 * of StepWorker/rkAdvance for a stage other than rkStep 2 (no filtering of
 * the state): sensor, vertex merge, viscosity field, Euler fluxes, viscous
 * fluxes, RHS + dissipation, then the modal filter of the RHS. */
int or_rhs(or_solver *s, const double *Q, double t, double *RHS) {
  if (s->shock) {
    size_t nq = 4 * (size_t)s->NpInt * s->K;
    double *Qc = malloc(sizeof(double) * nq);
    memcpy(Qc, Q, sizeof(double) * nq);
    shock_stage_pre(s, Qc, -1);
    interp_to_edges(s, Qc);
    edge_fluxes(s, t);
    shock_gradient_and_viscous_edges(s, Qc);
    rt_flux_internal(s, Qc);
    rt_flux_edges(s);
    rhs_internal(s, RHS);
    shock_add_dissipation(s, RHS);
    limit_filter(s, RHS, 4);
    free(Qc);
    return 0;
  }
  interp_to_edges(s, Q);
  edge_fluxes(s, t);
  rt_flux_internal(s, Q);
  rt_flux_edges(s);
  rhs_internal(s, RHS);
  return 0;
}

static int has_nan(const double *a, size_t n) {
  for (size_t i = 0; i < n; i++) if (a[i] != a[i]) return 1;
  return 0;
}

/* One SSP54 step: Step/StepWorker (euler.go:408-653), rkAdvance (:460-566),
 * InitializeDT (:655-663), calculateGlobalDT (:945-971), CalculateLocalDT
 * (:973-1002), CalcElementMaxWaveSpeed (edges.go:291-323) */
int or_step(or_solver *s) {
  int K = s->K, NpInt = s->NpInt;
  size_t nq = 4 * (size_t)NpInt * K, nv = (size_t)NpInt * K;
  double *stages[5] = {s->Q, s->Q1, s->Q2, s->Q3, s->Q4};
  double Np1 = (double)(s->p.N + 1), NP12 = Np1 * Np1;
  for (int rk = 0; rk < 5; rk++) {
    double *QQQ = stages[rk];
    if (s->shock) shock_stage_pre(s, QQQ, rk);
    interp_to_edges(s, QQQ);
    edge_fluxes(s, s->time);
    if (s->shock) shock_gradient_and_viscous_edges(s, QQQ);
    edge_aggregates(s);
    if (rk == 0)
      for (int k = 0; k < K; k++) s->DT[k] = -100;
    double gmax = -DBL_MAX, gmaxv = -DBL_MAX;
    for (int k = 0; k < K; k++)
      for (int e = 0; e < 3; e++) {
        int ei = s->EtoEdge[3 * k + e];
        double w = s->EdgeAgg[ei];
        if (w > gmax) gmax = w;
        s->DT[k] = fmax(s->DT[k], w);
        if (s->shock) {
          double wv = s->EdgeAggVisc[ei];
          if (wv > gmaxv) gmaxv = wv;
          s->DTVisc[k] = fmax(s->DTVisc[k], wv);
        }
      }
    if (!s->p.LocalTimeStepping) {
      s->gdt = s->p.CFL / gmax;
      if (s->shock) {
        double C_diff = 1. / NP12;
        s->gdt = fmin(s->gdt, C_diff / gmaxv);
      }
      if (s->time + s->gdt > s->p.FinalTime) s->gdt = s->p.FinalTime - s->time;
      for (int k = 0; k < K; k++) s->DT[k] = s->gdt;
    } else {
      for (int k = 0; k < K; k++) s->DT[k] = s->p.CFL / s->DT[k];
      if (s->shock) {
        double C_diff = 1. / NP12;
        for (int k = 0; k < K; k++)
          if (s->DTVisc[k] > 1.e-9) {
            s->DTVisc[k] = C_diff / s->DTVisc[k];
            s->DT[k] = fmin(s->DT[k], s->DTVisc[k]);
          }
      }
    }
    /* rkAdvance: the reference checks every array for NaN first (checkFunc,
     * euler.go:471-492: QQQ, F_RT_DOF, RHSQ, Jinv, Jdet, Q, Q1..Q4, DT) */
    if (s->nan_scan == 2) {
      size_t nf = 4 * (size_t)s->NpFlux * K;
      if (has_nan(QQQ, nq) || has_nan(s->FRT, nf) || has_nan(s->RHSQ, nq) || has_nan(s->Jinv, 4 * (size_t)K) ||
          has_nan(s->Jdet, K) || has_nan(s->Q, nq) || has_nan(s->Q1, nq) || has_nan(s->Q2, nq) ||
          has_nan(s->Q3, nq) || has_nan(s->Q4, nq) || has_nan(s->DT, K))
        return 1;
    } else if (s->nan_scan == 1) {
      if (has_nan(QQQ, nq) || has_nan(s->Q, nq) || has_nan(s->DT, K)) return 1;
    }
    rt_flux_internal(s, QQQ);
    rt_flux_edges(s);
    rhs_internal(s, s->RHSQ);
    if (s->shock) {
      shock_add_dissipation(s, s->RHSQ);
      limit_filter(s, s->RHSQ, 4);
    }
    for (int n = 0; n < 4; n++) {
      double *q0 = s->Q + n * nv, *q1 = s->Q1 + n * nv, *q2 = s->Q2 + n * nv, *q3 = s->Q3 + n * nv;
      double *q4 = s->Q4 + n * nv, *R = s->Resid + n * nv, *RHS = s->RHSQ + n * nv;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < NpInt; i++)
        for (int k = 0; k < K; k++) {
          size_t ind = k + (size_t)i * K;
          double dtRHS = s->DT[k] * RHS[ind];
          switch (rk) {
          case 0: q1[ind] = q0[ind] + 0.391752226571890 * dtRHS; break;
          case 1: q2[ind] = 0.444370493651235 * q0[ind] + 0.555629506348765 * q1[ind] + 0.368410593050371 * dtRHS; break;
          case 2: q3[ind] = 0.620101851488403 * q0[ind] + 0.379898148511597 * q2[ind] + 0.251891774271694 * dtRHS; break;
          case 3:
            R[ind] = RHS[ind];
            q4[ind] = 0.178079954393132 * q0[ind] + 0.821920045606868 * q3[ind] + 0.544974750228521 * dtRHS;
            break;
          case 4: {
            double dtR3 = s->DT[k] * R[ind];
            R[ind] = -q0[ind] + 0.517231671970585 * q2[ind] + 0.096059710526146 * q3[ind] +
                     0.386708617503269 * q4[ind] + 0.063692468666290 * dtR3 + 0.226007483236906 * dtRHS;
            q0[ind] += R[ind];
          } break;
          }
        }
    }
  }
  s->time += s->gdt;
  s->steps++;
  return has_nan(s->Q, nq) ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Shock capturing: Persson sensor + C0 artificial viscosity + modal        */
/* filter (dissipation.go, DG2D/dfr_shock_capturing.go)                      */
/* ------------------------------------------------------------------------ */

/* NewScalarDissipation (dissipation.go:111-165), NewAliasShockFinder
 * (dfr_shock_capturing.go:71-107), createInterpolationStencil (:414-449) */
static void setup_shock(or_solver *s) {
  int K = s->K, N = s->p.N, Np = s->NpInt, NpFlux = s->NpFlux, NpEdge = s->NpEdge;
  s->shock = 1;
  s->kappa_sd = 5.;
  s->eps0 = s->kappa_sd / 1.5;
  if (s->p.Kappa != 0.) s->kappa_sd = s->p.Kappa;
  s->Se = calloc(K, sizeof(double)); s->Sigma = calloc(K, sizeof(double)); s->Eps = calloc(K, sizeof(double));
  s->SigmaV = calloc(s->Nv, sizeof(double)); s->EpsV = calloc(s->Nv, sizeof(double));
  s->EpsF = calloc((size_t)NpFlux * K, sizeof(double)); s->SigmaF = calloc((size_t)NpFlux * K, sizeof(double));
  s->DissX = calloc(4 * (size_t)NpFlux * K, sizeof(double)); s->DissY = calloc(4 * (size_t)NpFlux * K, sizeof(double));
  (void)NpEdge;
  /* Clipper = V * diag(modes i+j < N) * Vinv; D = I - Clipper; P = M * D */
  double *diag = calloc(Np, sizeof(double));
  int ii = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N - i; j++) diag[ii++] = (i + j >= N) ? 0.0 : 1.0;
  double *tmp = calloc(Np * Np, sizeof(double));
  s->Clip = calloc(Np * Np, sizeof(double)); s->Dmat = calloc(Np * Np, sizeof(double));
  s->Pmat = calloc(Np * Np, sizeof(double));
  for (int i = 0; i < Np; i++)
    for (int j = 0; j < Np; j++) tmp[i * Np + j] = s->o.V[i * Np + j] * diag[j];
  mm(tmp, s->o.Vinv, s->Clip, Np, Np, Np);
  for (int i = 0; i < Np; i++)
    for (int j = 0; j < Np; j++) s->Dmat[i * Np + j] = ((i == j) ? 1.0 : 0.0) - s->Clip[i * Np + j];
  mm(s->o.Mass, s->Dmat, s->Pmat, Np, Np, Np);
  free(diag); free(tmp);
  /* ModalFilter with RecommendedFilterParameters (dfr_shock_capturing.go:43-69,109-140) */
  double alpha; int sp;
  if (N <= 2) { alpha = 4.0; sp = 4; }
  else if (N == 3) { alpha = 6.0; sp = 4; }
  else if (N == 4) { alpha = 8.0; sp = 5; }
  else if (N == 5) { alpha = 10.0; sp = 6; }
  else if (N == 6) { alpha = 14.0; sp = 6; }
  else { alpha = 18.0; sp = 7; }
  ii = 0;
  for (int i = 0; i <= N; i++)
    for (int j = 0; j <= N - i; j++) {
      int p = i + j;
      if (N <= 1 || p == 0) s->mf[ii] = 1.0;
      else s->mf[ii] = exp(-alpha * pow((double)p / (double)N, (double)sp));
      ii++;
    }
  /* barycentric coordinates of the RT points */
  s->Bary = calloc(3 * (size_t)NpFlux, sizeof(double));
  double RR[9] = {1, 1, 1, -1, 1, -1, -1, -1, 1}, RRinv[9];
  minv(RR, RRinv, 3);
  for (int i = 0; i < NpFlux; i++) {
    double C[3] = {1, s->o.R[i], s->o.S[i]};
    for (int a = 0; a < 3; a++) s->Bary[3 * i + a] = RRinv[3 * a] * C[0] + RRinv[3 * a + 1] * C[1] + RRinv[3 * a + 2] * C[2];
  }
}

/* UpdateSeMoment (dfr_shock_capturing.go:142-166) */
static void sensor_se(or_solver *s, const double *rho, double *Se) {
  int K = s->K, Np = s->NpInt;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++) {
    double num = 0, den = 0;
    for (int i = 0; i < Np; i++) {
      double X = 0, Y = 0, DQ = 0;
      for (int j = 0; j < Np; j++) {
        double q = rho[k + (size_t)j * K];
        X += s->Pmat[i * Np + j] * q;
        Y += s->o.Mass[i * Np + j] * q;
        DQ += s->Dmat[i * Np + j] * q;
      }
      num += DQ * X;
      den += rho[k + (size_t)i * K] * Y;
    }
    Se[k] = log10(num / den);
  }
}

/* The DG2D ShockIndicator (dfr_shock_capturing.go:189-235), S0 = 4/N^4 */
void or_shock_indicator(const or_solver *s, const double *rho, double kappa, double *sigma) {
  int K = s->K, Np = s->NpInt, N = s->p.N;
  double *U = malloc(sizeof(double) * Np), *Uc = malloc(sizeof(double) * Np);
  for (int k = 0; k < K; k++) {
    for (int i = 0; i < Np; i++) U[i] = rho[k + (size_t)i * K];
    for (int i = 0; i < Np; i++) {
      double acc = 0;
      for (int j = 0; j < Np; j++) acc += s->Clip[i * Np + j] * U[j];
      Uc[i] = acc;
    }
    double mn = 0, md = 0;
    for (int i = 0; i < Np; i++) {
      double mass = s->o.Mass[i + i * Np], t1 = U[i] - Uc[i];
      mn += mass * (t1 * t1);
      md += mass * (U[i] * U[i]);
    }
    double Se = log10(mn / md);
    double S0 = 4. / pow((double)N, 4.), left = S0 - kappa, right = S0 + kappa, ook = 0.5 / kappa;
    double sg;
    if (Se < left) sg = 0.;
    else if (Se <= right) sg = 0.5 * (1. + sin(M_PI * ook * (Se - S0)));
    else sg = 1.;
    sigma[k] = sg;
  }
  free(U); free(Uc);
}

/* MergeElementScalarToVertices with max (euler.go:1048-1065) */
static void vertex_max(const or_solver *s, const double *elem, double *vert) {
  char *seen = calloc(s->Nv, 1);
  for (int k = 0; k < s->K; k++)
    for (int a = 0; a < 3; a++) {
      int v = s->EToV[3 * k + a];
      if (!seen[v]) { vert[v] = elem[k]; seen[v] = 1; }
      else vert[v] = fmax(vert[v], elem[k]);
    }
  free(seen);
}
void or_vertex_max(const or_solver *s, const double *elem, double *vert) { vertex_max(s, elem, vert); }

/* limitAndFilterSolution (dissipation.go:606-622) applied to nvars variables */
static void limit_filter(or_solver *s, double *U, int nvars) {
  int K = s->K, Np = s->NpInt;
#pragma omp parallel
  {
  double *m = malloc(sizeof(double) * Np);
  for (int n = 0; n < nvars; n++) {
    double *u = U + (size_t)n * Np * K;
#pragma omp for schedule(static)
    for (int k = 0; k < K; k++) {
      double alpha = sin(0.5 * M_PI * s->Sigma[k]);
      for (int i = 0; i < Np; i++) {
        double acc = 0;
        for (int j = 0; j < Np; j++) acc += s->o.Vinv[i * Np + j] * u[k + (size_t)j * K];
        m[i] = acc;
      }
      for (int i = 1; i < Np; i++) m[i] *= s->mf[i] * (1. - alpha);
      for (int i = 0; i < Np; i++) {
        double acc = 0;
        for (int j = 0; j < Np; j++) acc += s->o.V[i * Np + j] * m[j];
        u[k + (size_t)i * K] = acc;
      }
    }
  }
  free(m);
  }
}

/* The sensor ramp of UpdateShockFinderSigma (dissipation.go:491-518):
 * 0 below S0 - kappa, 1 above S0 + kappa, 1/2 (1 + sin(pi (Se - S0) / (2 kappa)))
 * between, with the stepping path's S0 = 4/(N+1)^4 (dissipation.go:500) */
static double sigma_ramp(double se, double kappa, int N) {
  double S0 = 4. / pow((double)(N + 1), 4.);
  double left = S0 - kappa, right = S0 + kappa, ook = 0.5 / kappa, sg = 0.;
  if (se < left) sg = 0.;
  else if (se >= left && se <= right) sg = 0.5 * (1. + sin(M_PI * ook * (se - S0)));
  else if (se > right) sg = 1.;
  return sg;
}
double or_sigma_ramp(double se, double kappa, int N) { return sigma_ramp(se, kappa, N); }

/* Stage phases 1-3 of StepWorker (euler.go:574-616): sensor, sigma, eps,
 * vertex max-merge, vertex->element sigma mean, C0 interpolation of eps and
 * sigma, and the stage-state filter at rkStep == 2. */
static void shock_stage_pre(or_solver *s, double *QQQ, int rkStep) {
  int K = s->K, NpFlux = s->NpFlux, N = s->p.N;
  sensor_se(s, QQQ, s->Se);
  /* UpdateShockFinderSigma (dissipation.go:491-518) */
  double kappa = s->kappa_sd;
  double Np12 = (double)((N + 1) * (N + 1));
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++) {
    double sg = sigma_ramp(s->Se[k], kappa, N);
    s->Sigma[k] = sg;
    /* CalculateElementViscosity (:397-412) */
    s->Eps[k] = s->eps0 * (s->EdgeLenMax[k] / Np12) * sg;
  }
  vertex_max(s, s->Sigma, s->SigmaV);
  vertex_max(s, s->Eps, s->EpsV);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++) {
    double acc = 0;
    for (int a = 0; a < 3; a++) acc = acc + s->SigmaV[s->EToV[3 * k + a]];
    s->Sigma[k] = acc / 3.;
  }
  /* InterpolateEpsilonSigma (:219-242) */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < K; k++) {
    int *v = &s->EToV[3 * k];
    for (int i = 0; i < NpFlux; i++) {
      const double *b = &s->Bary[3 * i];
      s->EpsF[k + (size_t)i * K] = b[0] * s->EpsV[v[0]] + b[1] * s->EpsV[v[1]] + b[2] * s->EpsV[v[2]];
      s->SigmaF[k + (size_t)i * K] = b[0] * s->SigmaV[v[0]] + b[1] * s->SigmaV[v[1]] + b[2] * s->SigmaV[v[2]];
    }
  }
  if (rkStep == 2) limit_filter(s, QQQ, 4);
}

/* CalculateEpsilonGradient (dissipation.go:244-272) via
 * GetSolutionGradientUsingRTElement (euler.go:864-918) and DXMetric/DYMetric
 * (dfr_startup.go:213-254); then StoreEdgeViscousFlux (edges.go:151-244).
 * The viscous penalty jump uses the owner's stored trace for both sides
 * (edges.go:225-237). */
/* The RT-element gradient of variable n of QQQ at the NpFlux points, times
 * eps (eps == NULL: times 1), from GetSolutionGradientUsingRTElement
 * (euler.go:864-918): interior DOFs Jinv*U, edge DOFs n*IInII/Jdet times the
 * owner's stored trace (read reversed by the non-owner), then Div. */
static void rt_gradient(or_solver *s, const double *QQQ, int n, const double *eps, double *DX, double *DY) {
  int K = s->K, NpInt = s->NpInt, NpEdge = s->NpEdge, NpFlux = s->NpFlux;
#pragma omp parallel
  {
  double *DOFX = malloc(sizeof(double) * NpFlux), *DOFY = malloc(sizeof(double) * NpFlux);
#pragma omp for schedule(static)
  for (int k = 0; k < K; k++) {
    double *Ji = &s->Jinv[4 * k], ooJd = 1. / s->Jdet[k];
    for (int i = 0; i < NpInt; i++) {
      double Un = QQQ[(size_t)n * NpInt * K + k + (size_t)i * K];
      DOFX[i] = Ji[0] * Un; DOFY[i] = Ji[1] * Un;
      DOFX[i + NpInt] = Ji[2] * Un; DOFY[i + NpInt] = Ji[3] * Un;
    }
    for (int e = 0; e < 3; e++) {
      int faceInd = k + K * e, ei = s->FaceEdge[faceInd];
      int sign = (s->E[ei].tri[0] == k) ? 1 : -1;
      double nx = s->FNx[faceInd], ny = s->FNy[faceInd], IInII = s->IInII[faceInd];
      const double *EQ = s->EdgeQ + (size_t)ei * NpEdge * 4;
      for (int i = 0; i < NpEdge; i++) {
        int ii = (sign < 0) ? NpEdge - 1 - i : i;
        double Un = EQ[4 * ii + n];
        DOFX[2 * NpInt + e * NpEdge + i] = ooJd * nx * IInII * Un;
        DOFY[2 * NpInt + e * NpEdge + i] = ooJd * ny * IInII * Un;
      }
    }
    for (int i = 0; i < NpFlux; i++) {
      double ax = 0, ay = 0;
      for (int j = 0; j < NpFlux; j++) { ax += s->o.Div[i * NpFlux + j] * DOFX[j]; ay += s->o.Div[i * NpFlux + j] * DOFY[j]; }
      size_t ind = k + (size_t)i * K;
      double e = eps ? eps[ind] : 1.0;
      DX[ind] = ax * e;
      DY[ind] = ay * e;
    }
  }
  free(DOFX); free(DOFY);
  }
}

static void shock_gradient_and_viscous_edges(or_solver *s, const double *QQQ) {
  int K = s->K, NpInt = s->NpInt, NpEdge = s->NpEdge, NpFlux = s->NpFlux, N = s->p.N;
  for (int n = 0; n < 4; n++)
    rt_gradient(s, QQQ, n, s->EpsF, s->DissX + (size_t)n * NpFlux * K, s->DissY + (size_t)n * NpFlux * K);
  double Omega = 1. * (double)(N * N);
#pragma omp parallel for schedule(static)
  for (int ei = 0; ei < s->ne; ei++) {
    edge_t *E = &s->E[ei];
    int kL = E->tri[0], enL = E->enr[0], shiftL = NpEdge * enL;
    double nxL = s->FNx[kL + K * enL], nyL = s->FNy[kL + K * enL];
    double *VF = s->EdgeVisc + (size_t)ei * NpEdge * 4;
    double len = (enL == 1) ? (2. * sqrt(2)) * E->IInII[0] : 2. * E->IInII[0];
    double ooL = 1. / len;
    if (E->nconn == 1) {
      for (int n = 0; n < 4; n++)
        for (int i = 0; i < NpEdge; i++) {
          size_t indL = kL + (size_t)(2 * NpInt + shiftL + i) * K;
          VF[4 * i + n] = nxL * s->DissX[(size_t)n * NpFlux * K + indL] + nyL * s->DissY[(size_t)n * NpFlux * K + indL];
        }
    } else {
      int kR = E->tri[1], shiftR = NpEdge * E->enr[1];
      for (int n = 0; n < 4; n++)
        for (int i = 0; i < NpEdge; i++) {
          size_t indL = kL + (size_t)(2 * NpInt + shiftL + i) * K;
          size_t indR = kR + (size_t)(2 * NpInt + shiftR + NpEdge - 1 - i) * K;
          double vFL = nxL * s->DissX[(size_t)n * NpFlux * K + indL] + nyL * s->DissY[(size_t)n * NpFlux * K + indL];
          double vFR = nxL * s->DissX[(size_t)n * NpFlux * K + indR] + nyL * s->DissY[(size_t)n * NpFlux * K + indR];
          VF[4 * i + n] = 0.5 * (vFL + vFR);
        }
      const double *EQ = s->EdgeQ + (size_t)ei * NpEdge * 4;
      for (int n = 0; n < 4; n++)
        for (int i = 0; i < NpEdge; i++) {
          size_t indL = kL + (size_t)(2 * NpInt + shiftL + i) * K;
          size_t indR = kR + (size_t)(2 * NpInt + shiftR + NpEdge - 1 - i) * K;
          double LambdaAvg = 0.5 * (s->EpsF[indL] + s->EpsF[indR]);
          int ii = NpEdge - 1 - i;
          VF[4 * i + n] -= (Omega * LambdaAvg * ooL) * (EQ[4 * i + n] - EQ[4 * ii + n]);
        }
    }
  }
}

/* AddDissipation (dissipation.go:274-346) */
static void shock_add_dissipation(or_solver *s, double *RHS) {
  int K = s->K, NpInt = s->NpInt, NpEdge = s->NpEdge, NpFlux = s->NpFlux;
#pragma omp parallel
  {
  double *DOF = malloc(sizeof(double) * NpFlux);
  for (int n = 0; n < 4; n++) {
    const double *DX = s->DissX + (size_t)n * NpFlux * K, *DY = s->DissY + (size_t)n * NpFlux * K;
#pragma omp for schedule(static)
    for (int k = 0; k < K; k++) {
      double Jd = s->Jdet[k], *Ji = &s->Jinv[4 * k];
      for (int i = 0; i < NpInt; i++) {
        size_t ind = k + (size_t)i * K;
        DOF[i] = Jd * (Ji[0] * DX[ind] + Ji[1] * DY[ind]);
        DOF[i + NpInt] = Jd * (Ji[2] * DX[ind] + Ji[3] * DY[ind]);
      }
      for (int e = 0; e < 3; e++) {
        int fInd = k + K * e, ei = s->FaceEdge[fInd];
        int sign = (s->E[ei].tri[0] == k) ? 1 : -1;
        double IInII = s->IInII[fInd];
        const double *VF = s->EdgeVisc + (size_t)ei * NpEdge * 4;
        for (int i = 0; i < NpEdge; i++) {
          int ii = (sign == -1) ? NpEdge - 1 - i : i;
          DOF[2 * NpInt + e * NpEdge + i] = VF[4 * ii + n] * IInII * (double)sign;
        }
      }
      double oojd = 1. / Jd;
      for (int i = 0; i < NpInt; i++) {
        double acc = 0;
        for (int j = 0; j < NpFlux; j++) acc += s->o.DivInt[i * NpFlux + j] * DOF[j];
        RHS[(size_t)n * NpInt * K + k + (size_t)i * K] += oojd * acc;
      }
    }
  }
  free(DOF);
  }
}

void or_get_shock(const or_solver *s, double *Se, double *Sigma, double *Eps) {
  if (!s->shock) return;
  if (Se) memcpy(Se, s->Se, sizeof(double) * s->K);
  if (Sigma) memcpy(Sigma, s->Sigma, sizeof(double) * s->K);
  if (Eps) memcpy(Eps, s->Eps, sizeof(double) * s->K);
}

/* Gradient of the four variables of Q by the RT element with eps = 1
 * (GetSolutionGradient, euler.go:864-918): DX, DY [4][NpFlux][K]. */
void or_gradient(or_solver *s, const double *Q, double *DX, double *DY) {
  size_t nf = (size_t)s->NpFlux * s->K;
  interp_to_edges(s, Q);
  edge_fluxes(s, 0.0);
  for (int n = 0; n < 4; n++) rt_gradient(s, Q, n, NULL, DX + n * nf, DY + n * nf);
}

/* UpdateSeMoment sensor of a density field [NpInt][K]; needs the shock path. */
int or_sensor(or_solver *s, const double *rho, double *Se) {
  if (!s->shock) return 1;
  sensor_se(s, rho, Se);
  return 0;
}

/* limitAndFilterSolution of nvars fields [nvars][NpInt][K] with a given
 * per-element sigma [K]; needs the shock path. */
int or_filter(or_solver *s, double *U, const double *sigma, int nvars) {
  if (!s->shock) return 1;
  double *keep = malloc(sizeof(double) * s->K);
  memcpy(keep, s->Sigma, sizeof(double) * s->K);
  memcpy(s->Sigma, sigma, sizeof(double) * s->K);
  limit_filter(s, U, nvars);
  memcpy(s->Sigma, keep, sizeof(double) * s->K);
  free(keep);
  return 0;
}

/* Viscosity field and dissipative fluxes of the last shock-path evaluation:
 * EpsF [NpFlux][K], DissX/DissY [4][NpFlux][K], viscous edge fluxes
 * [nedges][NpEdge][4]. */
int or_get_dissipation(const or_solver *s, double *EpsF, double *DissX, double *DissY, double *EdgeVisc) {
  if (!s->shock) return 1;
  size_t nf = (size_t)s->NpFlux * s->K;
  if (EpsF) memcpy(EpsF, s->EpsF, sizeof(double) * nf);
  if (DissX) memcpy(DissX, s->DissX, sizeof(double) * 4 * nf);
  if (DissY) memcpy(DissY, s->DissY, sizeof(double) * 4 * nf);
  if (EdgeVisc) memcpy(EdgeVisc, s->EdgeVisc, sizeof(double) * 4 * (size_t)s->NpEdge * s->ne);
  return 0;
}

/* Per-element DT of the last stage (CalculateLocalDT) */
void or_get_dt(const or_solver *s, double *DT) { memcpy(DT, s->DT, sizeof(double) * s->K); }
