/*
 * cfd_hip.h - C ABI of the MI355X (gfx950) implicit unstructured
 * finite-volume Euler hot path (libcfd_hip.so).
 *
 * The reference solver (ashishgupta02/cfdsolutions, Euler3D/MPFLOW3D) keeps
 * its state in globals and exposes argument-less C++ functions; this ABI
 * takes the same arrays explicitly, keeps them resident in HBM behind an
 * opaque handle, and returns status codes instead of calling exit().  Each
 * entry point names the reference function it replaces (paths relative to
 * Trunk/ of the reference), so a C++ shim with the reference's exact
 * signatures maps each reference call onto one call here.
 *
 * Conventions
 *   - Arrays use the reference layouts: conserved state and residuals are
 *     structure-of-arrays [5][n] (Q1..Q5, Res1..Res5), ghost nodes follow the
 *     physical nodes (index n_node + k for boundary edge k), area vectors and
 *     coordinates are interleaved x,y,z, BSR blocks are 25 doubles row-major
 *     (A[blk][row][col]), the right-hand side and solution are [N][5].
 *   - All pointers passed in are host pointers unless the name says device.
 *   - Every call is ordered on the handle's own HIP stream; calls on one
 *     handle are not thread-safe, different handles are independent.
 *   - Return value: CFD_OK or a negative CFD_E_* code; cfd_last_error()
 *     describes the last failure seen by the calling thread (else the last
 *     failure on the context).
 */
#ifndef CFD_HIP_H
#define CFD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_E_ARG -1
#define CFD_E_HIP -2
#define CFD_E_OOM -3
#define CFD_E_NONFINITE -4          /* NaN/Inf in residual or linear solve */
#define CFD_E_NEGATIVE_PRESSURE -5  /* characteristic BC produced p < 0 */
#define CFD_E_SINGULAR -6           /* zero pivot in a 5x5 diagonal block */
#define CFD_E_HLLC -7               /* no HLLC wave configuration applies */
#define CFD_E_STATE -8              /* call order: required data not computed */

/* flux schemes (values of FLUX_SCHEME, Euler3D/include/SolverParameters.h:109-119) */
#define CFD_FLUX_ROE 0
#define CFD_FLUX_HLLC 1
#define CFD_FLUX_AUSMP 2
/* limiters (LIMITER_METHOD_NONE, _VENKATAKRISHNAN and _BERTH_JESPERSEN are -1, 1 and 0 in
 * Euler3D/include/SolverParameters.h:343-345; the shim maps them by name) */
#define CFD_LIM_NONE 0
#define CFD_LIM_VENKAT 1
#define CFD_LIM_BJ 2
/* gradients */
#define CFD_GRAD_LSQ 0
#define CFD_GRAD_GG 1
/* precision of residual-path arithmetic and storage */
#define CFD_FP64 0
#define CFD_FP32 1
/* arithmetic of the gradient, limiter and residual kernels:
 * EXACT: the reference's operation order, each edge flux evaluated by both
 *        of its nodes and summed in the reference's edge order; FP64
 *        results are bit-identical to the reference.
 * FAST:  fused multiply-adds, one reciprocal per shared denominator, each
 *        interior edge flux evaluated once per patch of nodes; sums in a
 *        fixed order of their own (deterministic), FP64 results within 1e-12
 *        of EXACT relative to the largest residual entry. */
#define CFD_ARITH_EXACT 0
#define CFD_ARITH_FAST 1
/* preconditioners */
#define CFD_PREC_NONE 0
#define CFD_PREC_BJACOBI 1
#define CFD_PREC_ILU0 2
#define CFD_PREC_ILU0_MC 3  /* ILU(0) of the matrix with rows in multicolour order */
/* linear solvers */
#define CFD_SOLVER_FGMRES 0
#define CFD_SOLVER_BLOCK_GS 1  /* MC_Iterative_Block_LU_Jacobi_CRS, forward */
#define CFD_SOLVER_LUSGS 2     /* MC_Iterative_Block_LUSGS_CRS */

/* fields for cfd_field_set / cfd_field_get */
#define CFD_FIELD_Q 0          /* [5][N+B] */
#define CFD_FIELD_GRAD 1       /* [15][N] Q1x,Q1y,Q1z,...,Q5z */
#define CFD_FIELD_PHI 2        /* [5][N] */
#define CFD_FIELD_RES_CONV 3   /* [5][N] */
#define CFD_FIELD_RES_DISS 4   /* [5][N] */
#define CFD_FIELD_DT_ACC 5     /* [N] sum of area*max|lambda| */
#define CFD_FIELD_DT 6         /* [N] local time step */
#define CFD_FIELD_JAC_B 7      /* [N][5] */
#define CFD_FIELD_X 8          /* [N][5] linear-solve solution */

typedef struct cfd_ctx cfd_ctx;
typedef struct cfd_mesh cfd_mesh;

/* Arrays of the preprocessed mesh, as produced by the reference's
 * Create_Connectivity_Maps + Calculate_Area_Volume
 * (Euler3D/src/Connectivity_Maps.cpp:991-1059, Euler3D/src/Area_Volume.cpp:903-917). */
typedef struct cfd_mesh_desc {
    int32_t n_node, n_edge, n_bedge;
    const int32_t *edge;        /* [2E] intEdge[].node, L < R */
    const double *edge_area;    /* [3E] intEdge[].areav */
    const int32_t *bedge;       /* [2B] bndEdge[].node (node, n_node + k) */
    const double *bedge_area;   /* [3B] bndEdge[].areav, outward */
    const int32_t *bedge_type;  /* [B]  bndEdge[].type (BC_TYPE values) */
    const int32_t *bedge_tag;   /* [B]  surface tag, may be NULL */
    const double *xyz;          /* [3N] coordXYZ */
    const double *vol;          /* [N]  cVolume */
    const int32_t *n2n_ia;      /* [N+1] crs_IA_Node2Node */
    const int32_t *n2n_ja;      /*       crs_JA_Node2Node */
} cfd_mesh_desc;

/* Solver parameters (globals of Euler3D/src/SolverParameters.cpp:14-95 and
 * Euler3D/src/Material.cpp:16-58 for an ideal gas, GENERIC non-dimensionalisation). */
typedef struct cfd_params {
    double gamma;
    double inf_rho, inf_u, inf_v, inf_w, inf_p;
    double outflow_p;
    int32_t flux;           /* CFD_FLUX_* */
    int32_t order;          /* 1 or 2 */
    int32_t limiter;        /* CFD_LIM_* */
    int32_t limiter_order;  /* 1: reconstruction fraction 0, 2: fraction 1/2 */
    int32_t entropy_fix;    /* 0 or 1 (Kermani) */
    int32_t gradient;       /* CFD_GRAD_* */
    int32_t zpg_iter;
    int32_t precision;      /* CFD_FP64 / CFD_FP32 for gradient, limiter, residual */
    double venkat_k;
    double cfl_min, cfl_max;
    int32_t cfl_ramp;
    int32_t arith;          /* CFD_ARITH_* of gradient, limiter and residual */
} cfd_params;

typedef struct cfd_linsolve_opts {
    int32_t solver;         /* CFD_SOLVER_* */
    int32_t precond;        /* CFD_PREC_* (FGMRES only) */
    int32_t krylov_m;       /* restart length, <= 64 */
    int32_t max_restarts;
    double rtol;            /* FGMRES: stop at ||r|| <= rtol*||b|| */
    int32_t sweeps;         /* block GS / LU-SGS sweeps */
    int32_t mixed;          /* 0: FP64; 1: FP32 matrix, preconditioner and Krylov basis;
                             * 2: as 1, with the block-SELL copies of the multicolour
                             *    ILU(0) factors (systems of 131,072 rows or more)
                             *    stored in FP16 (per-block FP32 scales, FP32
                             *    arithmetic);
                             * 3: as 2, with the block-SELL copy of the matrix for
                             *    the SpMV in FP16 as well */
} cfd_linsolve_opts;

typedef struct cfd_linsolve_stats {
    int32_t iterations;
    int32_t restarts;
    double final_res;       /* FGMRES residual estimate, GS sweep RMS */
    double b_norm;
} cfd_linsolve_stats;

typedef struct cfd_solve_opts {
    int32_t max_iter;
    double tol;             /* absolute RMS of Conv+Diss, RMS_Res of Solver_Steady_Implicit */
    int32_t first_order_iters;
    double relaxation;
    int32_t wall_tag;       /* surface tag for forces (<= 0: every Euler wall) */
    double alpha_deg;       /* for lift/drag axes */
    cfd_linsolve_opts lin;
    int32_t jacobian;       /* CFD_SOLVE_JAC_* */
    int32_t autotune;       /* 1: cfd_autotune the mesh at the first Newton iteration (same results) */
} cfd_solve_opts;

/* Jacobian of the steady loop (always of the first-order flux, as the
 * reference's Compute_Jacobian switches to first order,
 * Euler3D/src/Jacobian.cpp:18-24):
 * APPROX: the approximate Roe Jacobian (Compute_Jacobian_Approximate_Roe),
 *         the reference's, for every scheme;
 * FD:     central finite differences of the scheme's own flux
 *         (cfd_jacobian_fd), the Jacobian of HLLC and AUSM+ themselves. */
#define CFD_SOLVE_JAC_APPROX 0
#define CFD_SOLVE_JAC_FD 1

typedef struct cfd_solve_result {
    int32_t iterations;     /* pseudo-time (Newton) iterations */
    int32_t krylov_total;
    int32_t converged;
    int32_t status;
    double first_rms, final_rms;
    double force[3];
    double lift, drag, pressure_integral;
    double seconds;
} cfd_solve_result;

/* ---- context and mesh ------------------------------------------------- */
int cfd_ctx_create(int device, cfd_ctx **out);
void cfd_ctx_destroy(cfd_ctx *ctx);
const char *cfd_last_error(const cfd_ctx *ctx);

/* Median-dual preprocessing on the GPU: from the UGRID content (nodes,
 * boundary triangles in UGRID winding with their tags, tetrahedra) to the
 * arrays of cfd_mesh_desc, bit-identical to the host library's
 * cfdm_preprocess and to the reference's Create_Connectivity_Maps
 * (Euler3D/src/Connectivity_Maps.cpp:44-340, 733-910) and
 * Calculate_Area_Volume (Euler3D/src/Area_Volume.cpp:462-821, 903-917).
 * The result is held on the host until cfd_prep_destroy. */
typedef struct cfd_prep cfd_prep;
typedef struct cfd_prep_counts {
    int32_t n_node, n_tri, n_tet, n_edge, n_bedge, n_n2n;
} cfd_prep_counts;
int cfd_preprocess(cfd_ctx *ctx, int32_t n_node, const double *xyz, int32_t n_tri, const int32_t *tri,
                   const int32_t *tri_tag, int32_t n_tet, const int32_t *tet, cfd_prep **out);
int cfd_prep_info(const cfd_prep *prep, cfd_prep_counts *counts);
/* Copies out the arrays (any pointer may be NULL), laid out as
 * cfdm_get_arrays: tri [3*n_tri] after winding translation and orientation,
 * edge [2E], edge_area [3E], bedge [2B], bedge_area [3B], bedge_tag [B],
 * vol [N], n2n_ia [N+1], n2n_ja [n_n2n]. */
int cfd_prep_get(const cfd_prep *prep, int32_t *tri, int32_t *edge, double *edge_area, int32_t *bedge,
                 double *bedge_area, int32_t *bedge_tag, double *vol, int32_t *n2n_ia, int32_t *n2n_ja);
void cfd_prep_destroy(cfd_prep *prep);

/* Uploads a mesh and builds the device structures: node incidence lists in
 * the reference's accumulation order, LSQ coefficients
 * (Euler3D/src/Gradient.cpp:137-222), BSR structure
 * (Create_CRS_SolverBlockMatrix, Euler3D/src/Solver.cpp:612-704) and the
 * level schedules of the triangular sweeps.  The state is set to the
 * freestream of params (Solver_Set_Initial_Conditions,
 * Euler3D/src/Solver.cpp:340-362). */
int cfd_mesh_create(cfd_ctx *ctx, const cfd_mesh_desc *desc, const cfd_params *params, cfd_mesh **out);
void cfd_mesh_destroy(cfd_mesh *mesh);
/* HBM bytes held by the handle (device allocations) */
int64_t cfd_mesh_device_bytes(const cfd_mesh *mesh);
/* Bytes a mesh of this size would hold with a Jacobian/Krylov workspace
 * (admission control); mixed = 1 for FP32 solver storage. */
int64_t cfd_footprint_bytes(int32_t n_node, int32_t n_edge, int32_t n_bedge, int32_t krylov_m, int32_t mixed);
/* The handle's HIP stream (hipStream_t) for interop and timing. */
void *cfd_mesh_stream(cfd_mesh *mesh);
/* Blocks until all work on the handle's stream is done. */
int cfd_mesh_sync(cfd_mesh *mesh);

int cfd_field_set(cfd_mesh *mesh, int field, const double *host);
int cfd_field_get(cfd_mesh *mesh, int field, double *host);

/* ---- hot path ---------------------------------------------------------- */
/* Apply_Boundary_Condition (Euler3D/src/BC.cpp:1440-1487, 1847-1860):
 * characteristic ghost states for every boundary edge. */
int cfd_bc(cfd_mesh *mesh, const cfd_params *p, int iter);
/* Compute_Least_Square_Gradient(0) (Euler3D/src/Gradient.cpp:227-295) or
 * the median-dual Green-Gauss gradient (p->gradient). */
int cfd_gradient(cfd_mesh *mesh, const cfd_params *p);
/* Compute_Limiter (Euler3D/src/Limiters.cpp:20-110). */
int cfd_limiter(cfd_mesh *mesh, const cfd_params *p);
/* cfd_gradient followed by cfd_limiter, as the steady loop runs them before
 * the residual (Euler3D/src/Solver_Steady_Implicit.cpp:94-99); with the
 * Green-Gauss gradient both run in one pass over the node incidences.  Same
 * results as the two calls. */
int cfd_gradient_limiter(cfd_mesh *mesh, const cfd_params *p);
/* Compute_Residual(AddTime) (Euler3D/src/Residual.cpp:18-49): convective and
 * dissipative residual and the time-step accumulator, left on the device.
 * eig6 (host, may be NULL) receives the eigenvalue extrema of the Roe flux. */
int cfd_residual(cfd_mesh *mesh, const cfd_params *p, int add_time, double *eig6);
/* Compute_DeltaT (Euler3D/src/Time_Step.cpp:23-50); returns CFL in *cfl. */
int cfd_delta_t(cfd_mesh *mesh, const cfd_params *p, int iter, double *cfl);
/* RMS of Conv+Diss: RMS_Writer (MPFLOW3D/src/RMSIO.cpp:98-153); Euler3D squares Conv only
 * (Euler3D/src/Solver_Steady_Implicit.cpp:112-126), which equals this when Diss = 0. */
int cfd_residual_rms(cfd_mesh *mesh, double *rms);
/* Compute_Jacobian (Euler3D/src/Jacobian.cpp:18-47 ->
 * Compute_Jacobian_Approximate_Roe, Euler3D/src/Roe_Jacobian.cpp:29-228):
 * first-order approximate Roe Jacobian with V/dt on the diagonal.
 * rhs_mode 0: B = -Conv (Euler3D), 1: B = -(Conv+Diss). */
int cfd_jacobian(cfd_mesh *mesh, const cfd_params *p, int rhs_mode, int mixed);
/* Compute_Jacobian_FiniteDifference_Roe (Euler3D/src/Roe_Jacobian.cpp:236-776),
 * selected by Compute_Jacobian_Roe (Euler3D/src/Roe_Jacobian.cpp:780-804) for
 * the central, forward, backward and alternate methods: finite differences
 * (eps 1e-8) of the matrix-form Roe flux Compute_RoeFlux
 * (Euler3D/src/Roe_Fluxes.cpp:552-905) with respect to each end's state; for
 * HLLC and AUSM+ (p->flux) the same perturbations differentiate the scheme's
 * own flux, as MPFLOW3D's scheme-agnostic Compute_Jacobian_FiniteDifference
 * (MPFLOW3D/src/FiniteDifference_Jacobian.cpp:112-667) does;
 * second-order face states when p->order is 2.  Boundary edges are always
 * central, with the characteristic ghost state recomputed for each
 * perturbation.  CFD_JAC_FD_ALTERNATE is forward on even iterations and
 * backward on odd ones.  add_time puts V/dt on the diagonal; rhs_mode as for
 * cfd_jacobian.  Needs the FP64 state path (p->precision CFD_FP64) and a
 * current residual (and gradient and limiter for second order); mixed stores
 * the blocks in FP32. */
#define CFD_JAC_FD_CENTRAL 0
#define CFD_JAC_FD_FORWARD 1
#define CFD_JAC_FD_BACKWARD 2
#define CFD_JAC_FD_ALTERNATE 3
int cfd_jacobian_fd(cfd_mesh *mesh, const cfd_params *p, int method, int iter, int add_time, int rhs_mode,
                    int mixed);
/* Copies the BSR structure (ia[N+1], ja[nnz], iau[N]; any may be NULL) and
 * returns nnz. */
int32_t cfd_bsr_structure(cfd_mesh *mesh, int32_t *ia, int32_t *ja, int32_t *iau);
int cfd_jacobian_get(cfd_mesh *mesh, double *a, double *b);
int cfd_jacobian_set(cfd_mesh *mesh, const double *a, const double *b);
/* Linear solve A X = B: FGMRES + ILU(0)/block-Jacobi, or the reference's
 * MC_Iterative_Block_LU_Jacobi_CRS / MC_Iterative_Block_LUSGS_CRS
 * (MATH/src/MC_Iterative_Jacobi.c:346-717). X stays on the device
 * (CFD_FIELD_X). */
int cfd_linsolve(cfd_mesh *mesh, const cfd_linsolve_opts *o, cfd_linsolve_stats *st);
/* No counterpart in the reference: cfd_linsolve with CFD_SOLVER_BLOCK_GS
 * (MC_Iterative_Block_LU_Jacobi_CRS, forward) for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing: the LU factors of
 * the diagonal blocks and the per-row sums live in a file-scope device array
 * sized for meshes of at most 32,768 nodes (6.8 MB per GPU, held until the
 * process ends); larger meshes get CFD_E_ARG.  Same results as cfd_linsolve,
 * bit for bit.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race. */
int cfd_linsolve_gs_noalloc(cfd_mesh *mesh, int sweeps, cfd_linsolve_stats *st);
/* No counterpart in the reference: autotuning of the kernel variants that
 * compute the same bits.  On this mesh and these parameters it times the
 * fused gradient+limiter with direct and with LDS-staged neighbour gathers,
 * and, when a Jacobian is assembled and lin is FGMRES, whole linear solves
 * with the BSR and the block-SELL Krylov kernels (and, for the natural-order
 * ILU(0), sync-free and level-launched sweeps); the mesh keeps the fastest of
 * each for its later calls.  Results are unchanged bit for bit; the
 * gradient, limiter and linear-solve outputs are overwritten.  A variable
 * CFD_STAGE_GL, CFD_SELL or CFD_SYNC_FREE in the environment fixes that
 * choice instead.  rep (may be NULL) receives the times in ms (0: not
 * timed) and the choices (-1: not tuned). */
typedef struct cfd_tune_report {
    double gl_ms[2];          /* fused gradient+limiter: direct, staged */
    double krylov_ms[2][2];   /* linear solve: [BSR, block-SELL][sync-free, level-launched] */
    int32_t gl_stage, sell, sync_free, pad_;
} cfd_tune_report;
int cfd_autotune(cfd_mesh *mesh, const cfd_params *params, const cfd_linsolve_opts *lin, cfd_tune_report *rep);
/* Q += relaxation * X (Solver_Steady_Implicit.cpp:150-156). */
int cfd_update(cfd_mesh *mesh, double relaxation);
/* Pressure force on wall boundary edges: F = sum p_node S_out. */
int cfd_forces(cfd_mesh *mesh, const cfd_params *p, int wall_tag, double alpha_deg, double *force3, double *lift,
               double *drag, double *pint);
/* Solver_Steady_Implicit (Euler3D/src/Solver_Steady_Implicit.cpp:33-168)
 * with B = -(Conv+Diss): BC, gradient, limiter, residual, time step, RMS,
 * Jacobian, linear solve, update, until RMS < tol.  rms_hist (host,
 * max_iter entries) may be NULL. */
int cfd_solve_steady(cfd_mesh *mesh, const cfd_params *p, const cfd_solve_opts *o, cfd_solve_result *res,
                     double *rms_hist);

/* Resets the state of a resident mesh to the freestream of p (the initial
 * condition of Solver_Set_Initial_Conditions, Euler3D/src/Solver.cpp:340-362)
 * and invalidates its Jacobian and preconditioner, so the mesh can serve
 * another solve without rebuilding its structure. */
int cfd_mesh_reset(cfd_mesh *mesh, const cfd_params *p);

/* ---- batched small cases ---------------------------------------------------
 * No counterpart in the reference.  Solves ncase independent steady cases in
 * one union mesh: every kernel launch (and every captured Krylov step)
 * covers all cases, while each case keeps its own freestream, residual norm,
 * Krylov dots and rotations, convergence test and forces.  Each case's
 * iterates are bit-identical to cfd_solve_steady on that case alone.  All
 * cases must share the scheme fields of cfd_params (everything except the
 * freestream) and the solver fields of cfd_solve_opts (everything except
 * alpha_deg).  Once the cases still iterating hold half of the union's
 * nodes or less, the union is rebuilt from just those cases (their states
 * carried over; CFD_BATCH_COMPACT=0 disables this), so results[c].seconds is
 * the time from the start of the batch until case c's result was final. */
int cfd_batch_solve(cfd_ctx *ctx, int32_t ncase, const cfd_mesh_desc *meshes, const cfd_params *params,
                    const cfd_solve_opts *opts, cfd_solve_result *results);

/* ---- serving -------------------------------------------------------------
 * No counterpart in the reference, which solves one case per process.  A
 * server owns worker threads that run queued steady solves concurrently on
 * one GPU: requests are admitted in arrival order (with bounded look-ahead)
 * when a worker is free and their HBM footprint fits the budget; device
 * meshes stay resident after a solve and serve later requests with the same
 * mesh_key.  With batch_max >= 2 the server batches continuously: while
 * queued requests outnumber the idle workers, a worker that admits a request
 * also takes compatible requests queued in its look-ahead window (same scheme
 * and solver options, FGMRES, meshes of at most batch_case_nodes nodes), as
 * many as needed for the queue to fit the idle workers and at most batch_max,
 * and solves them together as one union mesh (as cfd_batch_solve).  Each
 * batched result is bit-identical to the same request solved alone and is
 * posted as soon as it is final; requests that arrive meanwhile are admitted
 * by the other workers.  Results do not depend on concurrency (deterministic
 * kernels). */
typedef struct cfd_server cfd_server;

typedef struct cfd_server_opts {
    int32_t max_inflight;   /* concurrent solves (worker threads, one stream each), 1..64 */
    int32_t lookahead;      /* queue entries considered for admission (<= 0: 16) */
    int32_t max_bypass;     /* admissions ahead of a request before it blocks the window (<= 0: 8) */
    int32_t keep_resident;  /* 1: keep device meshes for reuse by requests with the same mesh_key */
    int64_t hbm_budget;     /* device bytes the server may hold (meshes, workspaces) */
    int32_t batch_max;      /* >= 2: up to this many requests per union-mesh batch; else one per solve */
    int32_t batch_nodes;    /* nodes per batch at most (<= 0: 2,000,000) */
    int32_t batch_wait_us;  /* while other solves run, a batchable request with fewer companions queued
                               than its batch would take waits up to this long for more (0: never waits) */
    int32_t batch_case_nodes; /* only requests on meshes of at most this many nodes are batched
                                 (<= 0: 131,072); a larger mesh fills the GPU on its own */
} cfd_server_opts;

typedef struct cfd_request {
    int64_t id;
    int64_t mesh_key;           /* requests with equal keys must pass the same mesh */
    const cfd_mesh_desc *mesh;  /* host arrays, valid until the completion is polled */
    cfd_params params;
    cfd_solve_opts opts;
    double *q_out;              /* host [5][N+B] final state, or NULL */
    double *rms_hist;           /* host [opts.max_iter] residual history, or NULL (not filled when batched) */
} cfd_request;

typedef struct cfd_completion {
    int64_t id;
    int32_t status;             /* CFD_OK or CFD_E_* */
    int32_t reused;             /* 1 if a resident mesh was reused */
    int32_t batch_size;         /* requests solved together with this one (1: alone) */
    int32_t pad_;
    cfd_solve_result result;
    double t_submit, t_admit, t_done;  /* seconds on the server clock */
    int64_t footprint;          /* bytes reserved for the request's mesh (its batch's union mesh) */
    int64_t device_bytes;       /* bytes the mesh actually held */
    char message[160];          /* error text when status != CFD_OK */
} cfd_completion;

typedef struct cfd_server_stats {
    int64_t bytes_in_use, peak_bytes;
    int32_t queued, running, resident, pad_;
    int64_t admitted, reused, evicted;
    double clock;
} cfd_server_stats;

int cfd_server_create(cfd_ctx *ctx, const cfd_server_opts *opts, cfd_server **out);
/* Queues a request; returns CFD_E_OOM at once if it can never fit the budget. */
int cfd_server_submit(cfd_server *server, const cfd_request *req);
/* Moves up to max finished requests into out, waiting up to wait_seconds for
 * the first; returns how many were written. */
int cfd_server_poll(cfd_server *server, cfd_completion *out, int32_t max, double wait_seconds);
int cfd_server_get_stats(cfd_server *server, cfd_server_stats *st);
/* Drops queued requests, waits for running ones, frees every resident mesh. */
void cfd_server_destroy(cfd_server *server);

#ifdef __cplusplus
}
#endif

#endif /* CFD_HIP_H */
