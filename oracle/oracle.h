/*
 * CPU oracle for the 2D Euler DFR right-hand side and SSP54 step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is a plain-C FP64 restatement of the
 * reference Go solver (Notargets/gocfd, model_problems/Euler2D and DG2D). It is
 * used by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg
 * as the checker; the product (gocfd_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows. Functions whose
 * results no golden vector, closed form or invariant pins are marked
 * "parity unpinned" in oracle.c.
 */
#ifndef GOCFD_ORACLE_H
#define GOCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Flux type codes: model_problems/Euler2D/fluxes.go:18-23 */
enum { OR_FLUX_AVERAGE = 0, OR_FLUX_LAX = 1, OR_FLUX_ROE = 2, OR_FLUX_ROE_ER = 3 };
/* Init type codes: model_problems/Euler2D/initialization.go:18-22 */
enum { OR_INIT_FREESTREAM = 0, OR_INIT_IVORTEX = 1, OR_INIT_SHOCKTUBE = 2 };
/* BC flags: types/cfd.go:13-26 */
enum {
  OR_BC_NONE = 0, OR_BC_IN, OR_BC_DIRICHLET, OR_BC_SLIP, OR_BC_FAR, OR_BC_WALL, OR_BC_CYL,
  OR_BC_NEUMAN, OR_BC_OUT, OR_BC_IVORTEX, OR_BC_PERIODIC, OR_BC_PERIODIC_REVERSED
};

typedef struct or_params {
  int N;               /* PolynomialOrder */
  int flux;            /* OR_FLUX_* */
  int init;            /* OR_INIT_* */
  double CFL, FinalTime;
  int MaxIterations;
  double Minf, Gamma, Alpha;
  int LocalTimeStepping;
  int limiter;         /* 0 none, 1 PerssonC0 */
  double Kappa;
} or_params;

typedef struct or_solver or_solver;

/* Mesh readers (readfiles/readSU2Grid.go:156, readfiles/readGambitGrid.go:21).
 * On success returns a mesh handle; *err receives a message on failure. */
typedef struct or_mesh or_mesh;
or_mesh *or_mesh_read(const char *path, char *err, int errlen);
void or_mesh_free(or_mesh *m);
int or_mesh_K(const or_mesh *m);
int or_mesh_Nv(const or_mesh *m);
void or_mesh_get(const or_mesh *m, double *VX, double *VY, int *EToV);

/* Solver (model_problems/Euler2D/euler.go:61 NewEuler) */
or_solver *or_create(const char *meshfile, const or_params *p, char *err, int errlen);
void or_destroy(or_solver *s);
int or_K(const or_solver *s);
int or_NpInt(const or_solver *s);
int or_NpEdge(const or_solver *s);
int or_NpFlux(const or_solver *s);
int or_num_edges(const or_solver *s);

/* State in the reference layout Q[n][i*K + k], n=0..3, i=0..NpInt-1 */
void or_get_state(const or_solver *s, double *Q);
void or_set_state(or_solver *s, const double *Q);
/* RHS of the given state at time t (euler.go:665-699 with edges.go:325-491) */
int or_rhs(or_solver *s, const double *Q, double t, double *RHS);
/* RT flux DOFs F_RT_DOF [4][NpFlux][K] of the last RHS evaluation (euler.go:701-726) */
void or_get_flux_dofs(const or_solver *s, double *F);
/* One SSP54 step (euler.go:408-653); returns 0, or 1 on NaN ("NAN found") */
int or_step(or_solver *s);
/* NaN scan before each stage: 0 none, 1 the stage input, q0 and DT (the
 * default), 2 the reference's per-stage list of eleven arrays (euler.go:471-484) */
void or_set_nan_scan(or_solver *s, int mode);
double or_time(const or_solver *s);
double or_dt(const or_solver *s);
int or_step_count(const or_solver *s);
void or_get_residual(const or_solver *s, double *R);

/* Operators and geometry for parity checks */
void or_get_operators(const or_solver *s, double *DivInt, double *Div, double *Interp,
                      double *R, double *S);
void or_get_solution_element(const or_solver *s, double *V, double *Vinv, double *Mass,
                             double *R, double *S, double *W);
void or_get_geometry(const or_solver *s, double *Jdet, double *Jinv, double *FaceNx,
                     double *FaceNy, double *IInII, double *EdgeLenMax);
void or_get_coords(const or_solver *s, double *SolX, double *SolY, double *FluxX,
                   double *FluxY);
/* Edge list sorted by key (min + max<<32): per edge
 * [key_lo, key_hi, nconn, tri0, tri1, edgeNum0, edgeNum1, dir0, dir1, bctype] */
void or_get_edges(const or_solver *s, long long *out10);
/* EtoE [K][3] and EdgeNumber face->edge index [3][K] (index into sorted edges) */
void or_get_connectivity(const or_solver *s, int *EtoE, int *FaceEdge);

/* Shock sensor quantities of the last stage (dissipation.go:491-518, 397-412) */
void or_get_shock(const or_solver *s, double *Se, double *Sigma, double *Eps);
/* Persson indicator of the DG2D shock finder on a density field (dfr_shock_capturing.go:189-209) */
void or_shock_indicator(const or_solver *s, const double *rho, double kappa, double *sigma);
/* The stepping path's sensor ramp sigma(Se) for Kappa and order N
 * (UpdateShockFinderSigma, dissipation.go:491-518) */
double or_sigma_ramp(double se, double kappa, int N);
/* Vertex max-merge of an element scalar (euler.go:1048-1065) */
void or_vertex_max(const or_solver *s, const double *elem, double *vert);
/* RT-element gradient with eps = 1 (GetSolutionGradient, euler.go:864-918) */
void or_gradient(or_solver *s, const double *Q, double *DX, double *DY);
/* Shock-path pieces; return 1 if the solver was made without PerssonC0 */
int or_sensor(or_solver *s, const double *rho, double *Se);
int or_filter(or_solver *s, double *U, const double *sigma, int nvars);
int or_get_dissipation(const or_solver *s, double *EpsF, double *DissX, double *DissY, double *EdgeVisc);
void or_get_dt(const or_solver *s, double *DT);

/* OpenMP threads used by the element and edge loops */
int or_threads(void);

/* Standalone pieces pinned by known-answer tests */
void or_ivortex_state(double t, double x, double y, double *q4);
void or_freestream(double Minf, double Gamma, double Alpha, double *q4, double *pinf,
                   double *cinf);
/* Numerical flux at one point: QL,QR [4], normal [2] -> F [4] */
void or_numerical_flux(int flux, double gamma, const double *QL, const double *QR,
                       const double *n, double *F);
void or_riemann_bc(double gamma, double minf, double pinf, double cinf, const double *Qint,
                   const double *Qinf, const double *n, double *Qbc);

#ifdef __cplusplus
}
#endif
#endif
