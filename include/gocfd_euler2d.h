/*
 * gocfd_euler2d.h — C ABI of the MI355X (gfx950) 2D Euler DFR solver.
 *
 * This is the drop-in boundary for the reference's Euler2D solver API
 * (Notargets/gocfd, model_problems/Euler2D). The reference has no FFI; each
 * entry point below states the reference interface it replaces. Plain
 * pointers, sizes and status codes only: a cgo binding
 * calls these directly.
 *
 * Layout conventions at this boundary are the reference's:
 *   state Q  : double[4][NpInt][K], index n*NpInt*K + i*K + k
 *              (c.Q[np][n] utils.Matrix NpInt x K, euler.go:47, with one shard)
 *   EToV     : int[K][3] zero-based vertex ids
 *   NpInt = (N+1)(N+2)/2, NpEdge = N+2, NpFlux = (N+2)(N+4)
 *
 * Every call returns 0 on success or a negative status; gocfd_last_error()
 * gives the message (the reference panics with the same text).
 * Calls on one handle are not re-entrant; different handles are independent.
 */
#ifndef GOCFD_EULER2D_H
#define GOCFD_EULER2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOCFD_OK 0
#define GOCFD_ERR_ARG -1      /* bad argument / unknown name (reference panics) */
#define GOCFD_ERR_MESH -2     /* mesh ingest or topology error */
#define GOCFD_ERR_SETUP -3    /* singular operator, unsupported order */
#define GOCFD_ERR_DEVICE -4   /* HIP runtime error or no usable gfx950 device */
#define GOCFD_ERR_NAN -5      /* "NAN found" (utils/system.go:20-24) */

/* Flux types: model_problems/Euler2D/fluxes.go:18-33 ("average","lax","roe","roe-er") */
#define GOCFD_FLUX_AVERAGE 0
#define GOCFD_FLUX_LAX 1
#define GOCFD_FLUX_ROE 2
#define GOCFD_FLUX_ROE_ER 3
/* Init types: model_problems/Euler2D/initialization.go:18-31 */
#define GOCFD_INIT_FREESTREAM 0
#define GOCFD_INIT_IVORTEX 1
#define GOCFD_INIT_SHOCKTUBE 2
/* Precision of the bulk arithmetic and state storage */
#define GOCFD_PREC_FP64 0    /* plain FP64 drop-in */
#define GOCFD_PREC_FP32 1    /* FP32 storage + flux/divergence; FP64 geometry, dt, SSP sums */

/* Solver parameters: the fields of InputParameters.InputParameters2D
 * (InputParameters/InputParameters.go:11-32) that reach the step. The host
 * language parses YAML (ghodss/yaml semantics) and fills this struct. */
typedef struct gocfd_params {
  int32_t polynomial_order;     /* PolynomialOrder (N), 0..4 on the GPU path */
  int32_t flux_type;            /* GOCFD_FLUX_* (FluxType) */
  int32_t init_type;            /* GOCFD_INIT_* (InitType) */
  int32_t max_iterations;       /* MaxIterations */
  double cfl;                   /* CFL */
  double final_time;            /* FinalTime */
  double minf, gamma, alpha;    /* Minf, Gamma, Alpha (degrees) */
  int32_t local_time_stepping;  /* LocalTimeStepping */
  int32_t limiter;              /* 0 none, 1 PerssonC0 (Limiter) */
  double kappa;                 /* Kappa */
  int32_t precision;            /* GOCFD_PREC_* */
  int32_t device;               /* HIP device ordinal; -1 = host-only handle (setup and
                                   introspection only, no device work) */
} gocfd_params;

typedef struct gocfd_mesh gocfd_mesh;
typedef struct gocfd_euler2d gocfd_euler2d;

/* Message of the last failed call on this thread. */
const char *gocfd_last_error(void);

/* ---- Mesh ingest: readfiles.ReadSU2 (readfiles/readSU2Grid.go:156) and
 *      readfiles.ReadGambit2d (readfiles/readGambitGrid.go:21); the type is
 *      chosen from the extension as in DG2D/dfr_startup.go:184-204. ---- */
int gocfd_mesh_read(const char *path, gocfd_mesh **out);
/* Build a mesh from arrays (what ReadSU2 returns): K, VX/VY[Nv], EToV[K*3],
 * and boundary tags: ntags tags, tag t named tags[t] with edge_count[t] edges
 * whose ordered vertex pairs are edge_verts[2*(edge_offset[t]+j)+{0,1}]. */
int gocfd_mesh_from_arrays(int32_t K, int32_t Nv, const double *VX, const double *VY,
                           const int32_t *EToV, int32_t ntags, const char *const *tags,
                           const int32_t *edge_offset, const int32_t *edge_count,
                           const int32_t *edge_verts, gocfd_mesh **out);
int gocfd_mesh_dims(const gocfd_mesh *m, int32_t *K, int32_t *Nv, int32_t *ntags);
int gocfd_mesh_arrays(const gocfd_mesh *m, double *VX, double *VY, int32_t *EToV);
void gocfd_mesh_free(gocfd_mesh *m);

/* ---- Solver: Euler2D.NewEuler(ip, meshFile, ...) (model_problems/Euler2D/euler.go:61).
 *      Builds DFR2D (DG2D/dfr_startup.go:43), the initial state
 *      (euler.go:728-794) and the device-resident solver. The mesh may be
 *      freed afterwards. ---- */
int gocfd_euler2d_create(const gocfd_mesh *mesh, const gocfd_params *params, gocfd_euler2d **out);
void gocfd_euler2d_destroy(gocfd_euler2d *h);
/* Sizes: K, NpInt, NpEdge, NpFlux, number of unique edges */
int gocfd_euler2d_dims(const gocfd_euler2d *h, int32_t *K, int32_t *NpInt, int32_t *NpEdge,
                       int32_t *NpFlux, int32_t *nedges);

/* ---- State transfer (c.Q, euler.go:47; RecombineShardsKBy4 euler.go:191) ---- */
int gocfd_euler2d_set_state(gocfd_euler2d *h, const double *Q);
int gocfd_euler2d_get_state(gocfd_euler2d *h, double *Q);

/* ---- RHS of a state at time t (parity hook for RHSInternalPoints
 *      euler.go:665-699 with CalculateEdgeEulerFlux edges.go:325-377 and
 *      SetRTFluxInternal/SetRTFluxOnEdges). Q and RHS in the reference layout. ---- */
int gocfd_euler2d_rhs(gocfd_euler2d *h, const double *Q, double t, double *RHS);

/* ---- RHS of a device-resident state with no device allocation. No
 *      counterpart in the reference: for callers that cannot allocate device
 *      memory (code under HIP graph capture, tight serving loops, harnesses
 *      that forbid hipMalloc). Same result as gocfd_euler2d_rhs, bit for
 *      bit. Q and RHS are DEVICE pointers on the handle's device, in the
 *      reference layout [4][NpInt][K]. stream: a hipStream_t to enqueue on
 *      (the call then returns without waiting), or NULL for the handle's own
 *      stream (the call returns when the RHS is written). The element-major
 *      staging lives in a fixed 32 MiB device array per device: meshes with
 *      2 * Kp * 4 * NpInt * 8 bytes above that (Kp = K rounded up to the
 *      64-element tile; about K > 34,900 at N=4) get GOCFD_ERR_ARG. Every
 *      handle on a device shares that array: calls that overlap on one
 *      device race, and the caller orders the call after step work of the
 *      same handle. ---- */
int gocfd_euler2d_rhs_noalloc(gocfd_euler2d *h, const double *Q, double t, double *RHS, void *stream);

/* ---- nsteps SSP54 steps: (*RungeKutta5SSP).Step (euler.go:408-418) with
 *      the Solve loop's Time += GlobalDT, StepCount++ (euler.go:175-182).
 *      Stops early, like CheckIfFinished (euler.go:796-801), once Time >=
 *      FinalTime. Returns GOCFD_ERR_NAN if a NaN appeared ("NAN found"). ---- */
int gocfd_euler2d_step(gocfd_euler2d *h, int32_t nsteps);
/* Asynchronous variant: enqueues the steps on the handle's stream and returns;
 * gocfd_euler2d_sync waits and reports a NaN. */
int gocfd_euler2d_step_async(gocfd_euler2d *h, int32_t nsteps);
int gocfd_euler2d_sync(gocfd_euler2d *h);

/* RK.Time, RK.GlobalDT (last stage's dt), RK.StepCount */
int gocfd_euler2d_time(gocfd_euler2d *h, double *time, double *global_dt, int32_t *steps);
int gocfd_euler2d_set_time(gocfd_euler2d *h, double time);
/* RK.Residual (euler.go:325; printed by PrintUpdate euler.go:814-849) */
int gocfd_euler2d_get_residual(gocfd_euler2d *h, double *R);

/* ---- Setup introspection for parity of the integer structure and operators ---- */
/* Unique edges sorted by key (min vertex + max vertex << 32), 10 int64 per edge:
 * [v_lo, v_hi, nconn, tri0, tri1, edgenum0, edgenum1, dir0, dir1, bctype]
 * (DG2D/triangulation.go:179-187 Edge, types/cfd.go:13-26 BCFLAG) */
int gocfd_euler2d_get_edges(gocfd_euler2d *h, int64_t *out);
/* EtoE[K][3] (triangulation.go:39-51) and the face -> edge index map [3][K]
 * (DFR2D.EdgeNumber, dfr_startup.go:295-302, as an index into the sorted edges) */
int gocfd_euler2d_get_connectivity(gocfd_euler2d *h, int32_t *EtoE, int32_t *FaceEdge);
/* DivInt[NpInt][NpFlux], Div[NpFlux][NpFlux], FluxEdgeInterp[3NpEdge][NpInt], R,S[NpFlux]
 * (raviart_thomas_element.go:192-297, dfr_startup.go:49-64) */
int gocfd_euler2d_get_operators(gocfd_euler2d *h, double *DivInt, double *Div, double *Interp,
                                double *R, double *S);
/* Jdet[K], Jinv[K][4], FaceNorm x/y [3][K], IInII[3][K], EdgeLenMax[K]
 * (dfr_startup.go:125-147, 256-314) */
int gocfd_euler2d_get_geometry(gocfd_euler2d *h, double *Jdet, double *Jinv, double *FaceNx,
                               double *FaceNy, double *IInII, double *EdgeLenMax);
/* Solution-point and flux-point coordinates [NpInt][K], [NpFlux][K]
 * (DFR2D.SolutionX/Y, FluxX/Y, dfr_startup.go:154-159) */
int gocfd_euler2d_get_coords(gocfd_euler2d *h, double *SolX, double *SolY, double *FluxX,
                             double *FluxY);

/* The solution basis interpolated to npts reference points (r, s):
 * out[npts][NpInt] = V(r, s) Vinv (JacobiBasis2D.GetInterpMatrix,
 * DG2D/basis_polynomials.go:49-53); what the reference's output writers
 * (SODShockTube.calculateInterpolation, sod_shock_tube/shock_tube.go:60-79;
 * GraphInterp, DG2D/dfr_startup.go:62-63) evaluate the state with. */
int gocfd_euler2d_interp_matrix(gocfd_euler2d *h, int32_t npts, const double *r, const double *s,
                                double *out);

/* ---- Shock capturing (Limiter: PerssonC0) and local time stepping ---- */
/* Per element, of the last stage evaluated: the sensor Se (UpdateSeMoment,
 * DG2D/dfr_shock_capturing.go:142-166), sigma after the vertex averaging
 * (sd.SigmaScalar, model_problems/Euler2D/euler.go:594-599) and the element
 * viscosity before it (sd.EpsilonScalar, dissipation.go:397-412). [K] each;
 * any pointer may be NULL. */
int gocfd_euler2d_get_shock(gocfd_euler2d *h, double *Se, double *Sigma, double *Eps);
/* Element time steps of the last stage with LocalTimeStepping (rk.DT after
 * CalculateLocalDT, euler.go:973-1002) [K] */
int gocfd_euler2d_get_element_dt(gocfd_euler2d *h, double *DT);

/* ---- Tangent-linear and discrete adjoint of the FP64 step (no counterpart
 *      in the reference, which has no derivative code; the differentiated
 *      map is its SSP54 step, rkAdvance euler.go:460-566, with the
 *      per-stage global dt of calculateGlobalDT euler.go:945-971). Both start
 *      from the solver's current state and time and leave the solver
 *      unchanged. Inputs of the map: the state Q [4][NpInt][K] and the
 *      far-field state FSFar.Qinf [4] (rho, rhoU, rhoV, E; FSFar's pressure
 *      and sound speed follow it); the In/Out states follow it too except in
 *      the shock tube. Needs an FP64 solver without IVortex boundary and
 *      steps that stay short of FinalTime; with the PerssonC0 limiter and
 *      global time stepping the whole shock-capturing path is
 *      differentiated (sensor, sigma ramp, element viscosity, vertex
 *      max-merges, sigma-dependent modal filters, eps grad Q, the viscous
 *      edge flux and the viscous dt; a vertex maximum follows the first
 *      element that attains it); with LocalTimeStepping a sensor that
 *      stays below S0 - Kappa in every element and an element viscous dt
 *      not yet set (GOCFD_ERR_ARG otherwise). The dt of a stage (of an
 *      element, with LocalTimeStepping) follows the face point of its
 *      maximum wave speed, or of its maximum viscous speed where that dt is
 *      smaller (ties: the first in the device order; with
 *      LocalTimeStepping the previous stage's element dt when that is
 *      larger, as in CalculateLocalDT). ---- */
/* Vector-Jacobian product over nsteps steps: dQ0 = (dQn/dQ)^T dQn_bar and
 * dQinf = (dQn/dQinf)^T dQn_bar (dQinf may be NULL). The reverse sweep
 * recomputes the steps from a checkpoint every checkpoint_every steps (0:
 * ceil(sqrt(nsteps))) and each step's stages from its start; with
 * checkpoint_every >= nsteps (one block) whose stage records fit in device
 * memory, the one recording sweep is also the forward sweep. */
int gocfd_euler2d_vjp(gocfd_euler2d *h, int32_t nsteps, int32_t checkpoint_every, const double *dQn_bar,
                      double *dQ0, double *dQinf);
/* Jacobian-vector product over nsteps steps (tangent-linear step in dual
 * arithmetic): dQn = (dQn/dQ) dQ0 + (dQn/dQinf) dQinf (dQinf may be NULL for
 * zero); Qn (may be NULL) receives the state after the steps. */
int gocfd_euler2d_jvp(gocfd_euler2d *h, int32_t nsteps, const double *dQ0, const double *dQinf, double *Qn,
                      double *dQn);

/* ---- Batched ensembles (no counterpart in the reference, which runs one
 *      case per process). ncases independent cases -- each its own mesh,
 *      order, initial state, Mach number, angle, CFL and FinalTime -- stepped
 *      together: cases of one order, limiter (PerssonC0 and its Kappa) and
 *      time-stepping mode share the launches of a stage, and runs of 6 steps
 *      replay a captured HIP graph. Every case needs the same precision,
 *      flux type and device. ---- */
typedef struct gocfd_ensemble gocfd_ensemble;
int gocfd_ensemble_create(int32_t ncases, const gocfd_mesh *const *meshes, const gocfd_params *params,
                          gocfd_ensemble **out);
void gocfd_ensemble_destroy(gocfd_ensemble *h);
int gocfd_ensemble_case_dims(gocfd_ensemble *h, int32_t c, int32_t *K, int32_t *NpInt);
/* nsteps SSP54 steps of every case (each case stops at its own FinalTime) */
int gocfd_ensemble_step(gocfd_ensemble *h, int32_t nsteps);
/* State of case c in the reference layout [4][NpInt][K] */
int gocfd_ensemble_get_state(gocfd_ensemble *h, int32_t c, double *Q);
/* RK.Time, RK.GlobalDT and RK.StepCount of case c */
int gocfd_ensemble_time(gocfd_ensemble *h, int32_t c, double *time, double *global_dt, int32_t *steps);
/* Wall time in ms of the last gocfd_ensemble_step call measured with HIP
 * events on the ensemble's stream, and its number of steps */
int gocfd_ensemble_kernel_time(gocfd_ensemble *h, double *ms, int32_t *steps);
/* Launch configuration (no counterpart in the reference): parallel = 1:
 * the order groups run as parallel branches of the captured graph; 2: the
 * groups without shock capturing run as one launch per stage of a
 * multi-order kernel (each workgroup runs its tile's order; FP64 ensembles
 * only; GOCFD_ERR_ARG where the groups do not allow it), the PerssonC0
 * groups after it; 0: one
 * group after another on one stream; graph_steps = steps per captured graph, a
 * multiple of 6, or 0 for eager launches without a graph (the default when
 * rocprofv3 is attached: ROCPROF_KERNEL_TRACE or ROCPROF_COUNTER_COLLECTION
 * set); xcd_tiles = each XCD takes a contiguous tile range (else
 * tile = workgroup id). None of them changes the results. */
int gocfd_ensemble_set_launch(gocfd_ensemble *h, int32_t parallel, int32_t graph_steps, int32_t xcd_tiles);
int gocfd_ensemble_get_launch(gocfd_ensemble *h, int32_t *parallel, int32_t *graph_steps, int32_t *xcd_tiles);
/* Autotuner: times each candidate configuration (parallel 0/1/2, graph_steps
 * 6/24 -- only 0 under rocprofv3 --, xcd_tiles 1/0) over trial_blocks graph replays from the initial
 * state, keeps the fastest and restores the initial state. Only before the
 * first step. gocfd_ensemble_kernel_time then reports the chosen
 * configuration's ms per step (steps = 1). */
int gocfd_ensemble_tune(gocfd_ensemble *h, int32_t trial_blocks);

/* ---- Device-level figures for measurement (no reference counterpart) ---- */
/* Average duration in ms of the last timed batch of stage kernels, measured
 * with HIP events on the handle's stream, and the number of launches. */
int gocfd_euler2d_kernel_time(gocfd_euler2d *h, double *avg_stage_ms, int32_t *launches);
/* Internal tiling chosen for this mesh: elements per tile, tiles, max halo faces */
int gocfd_euler2d_tiling(gocfd_euler2d *h, int32_t *tile, int32_t *ntiles, int32_t *max_halo);

#ifdef __cplusplus
}
#endif
#endif
