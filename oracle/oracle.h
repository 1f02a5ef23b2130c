/*
 * oracle.h - CPU restatement of the reference hot path.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker; the product never links it.
 *
 * Every function states the reference lines it follows (paths relative to
 * Trunk/ of ashishgupta02/cfdsolutions).  Parity status per function:
 *   pinned by golden vectors of the compiled reference (bit-exact):
 *     orc_residual (Roe, HLLC; first and second order), orc_bc,
 *     orc_lsq_gradient, orc_limiter (Venkatakrishnan, Barth-Jespersen,
 *     pressure corrections), orc_delta_t, orc_jacobian, orc_jacobian_fd
 *     (Roe), orc_block_gs, orc_lusgs
 *   pinned by closed forms only (no reference arithmetic exists), each
 *   chosen so that a plausible mistake fails it:
 *     AUSM+ flux (Liou 1996): exact stationary normal shock (pins the
 *       interface sound speed), the expanded subsonic split polynomials
 *       (pin alpha = 3/16, beta = 1/8), consistency, supersonic upwinding,
 *       stationary contact;
 *     orc_gg_gradient: exact for linear fields at interior nodes, zero for a
 *       nodal constant at every node whatever the ghost states (pins the
 *       boundary term);
 *     orc_ilu0 / orc_fgmres (Saad 2003): (L U)_ij = A_ij on A's pattern for
 *       a Jacobian with fill, exact LU without fill, dense solves;
 *     orc_forces: closed surface under uniform pressure gives F = 0; lift
 *       and drag are F along the reference's freestream axes at 30, 90 and
 *       -2 degrees;
 *     orc_jacobian_fd for HLLC and AUSM+: its product with a small state
 *       change equals the change of the (pinned) first-order residual to
 *       first order, which fixes the flux, signs and block placement;
 *     orc_solve_steady (iteration counts): the GPU runs the same algorithm.
 *
 * Layouts: Q, residuals, limiter: structure of arrays [5][n];
 * gradients [15][N] ordered Q1x,Q1y,Q1z,Q2x,...,Q5z; coordinates and area
 * vectors interleaved [3][...] per entity as in the reference; BSR blocks are
 * 25 doubles row-major.
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ROE = 0, ORC_HLLC = 1, ORC_AUSMP = 2 };
enum { ORC_LIM_NONE = 0, ORC_LIM_VENKAT = 1, ORC_LIM_BJ = 2 };
enum { ORC_GRAD_LSQ = 0, ORC_GRAD_GG = 1 };
/* boundary types: values of BC_TYPE in Euler3D/include/SolverParameters.h:40-58 */
enum {
    ORC_BC_EULER_WALL = 0, ORC_BC_FAR_FIELD = 1, ORC_BC_INFLOW = 2, ORC_BC_OUTFLOW = 3,
    ORC_BC_SUBSONIC_INFLOW = 4, ORC_BC_SUBSONIC_OUTFLOW = 5, ORC_BC_SUPERSONIC_INFLOW = 6,
    ORC_BC_SUPERSONIC_OUTFLOW = 7, ORC_BC_FF_SUBSONIC_INFLOW = 8, ORC_BC_FF_SUBSONIC_OUTFLOW = 9,
    ORC_BC_FF_SUPERSONIC_INFLOW = 10, ORC_BC_FF_SUPERSONIC_OUTFLOW = 11
};

typedef struct orc_mesh {
    int32_t n_node, n_edge, n_bedge;
    const int32_t *edge;      /* [2E] (L,R) */
    const double *edge_area;  /* [3E] L->R */
    const int32_t *bedge;     /* [2B] (node, ghost) */
    const double *bedge_area; /* [3B] outward */
    const int32_t *bedge_type;/* [B] */
    const double *xyz;        /* [3N] */
    const double *vol;        /* [N] */
    const int32_t *n2n_ia;    /* [N+1] */
    const int32_t *n2n_ja;
} orc_mesh;

typedef struct orc_params {
    double gamma;
    double r_gas;           /* NonDim_R */
    double inf_rho, inf_u, inf_v, inf_w, inf_p;
    double outflow_p;
    int32_t flux;           /* ORC_ROE / ORC_HLLC / ORC_AUSMP */
    int32_t order;          /* 1 or 2 */
    int32_t limiter;        /* ORC_LIM_* */
    int32_t limiter_order;  /* 1: reconstruction fraction 0, 2: fraction 1/2 */
    int32_t entropy_fix;    /* 0 or 1 (Kermani) */
    int32_t zpg_iter;
    double venkat_k;
    double cfl_min, cfl_max;
    int32_t cfl_ramp;
} orc_params;

/* Residual edge loop.  res_conv/res_diss [5][N] and dt_acc [N] are
 * accumulated into (caller zeroes them); eig[6] holds min/max of the three
 * eigenvalue magnitudes as in Euler3D/src/FluxTest.cpp:3116-3124 (Roe only). */
void orc_residual(const orc_mesh *m, const orc_params *p, const double *q, const double *grad,
                  const double *phi, double *res_conv, double *res_diss, double *dt_acc, double *eig);

/* One-edge flux (for closed-form tests): normal area vector s[3], states
 * ql[5], qr[5]; conv[5], diss[5]; returns area * max eigenvalue. */
double orc_edge_flux(const orc_params *p, const double *s, const double *ql, const double *qr,
                     double *conv, double *diss);

/* Characteristic boundary condition, updates ghost states in q. Returns the
 * number of boundary edges with negative pressure (reference: error()). */
int orc_bc(const orc_mesh *m, const orc_params *p, int iter, double *q);

void orc_lsq_gradient(const orc_mesh *m, const double *q, double *grad);
void orc_gg_gradient(const orc_mesh *m, const double *q, double *grad);
void orc_limiter(const orc_mesh *m, const orc_params *p, const double *q, const double *grad, double *phi);
/* returns CFL used */
double orc_delta_t(const orc_mesh *m, const orc_params *p, int iter, const double *dt_acc, double *dt);

/* BSR structure of Create_CRS_SolverBlockMatrix (Euler3D/src/Solver.cpp:612-704) */
int32_t orc_bsr_nnz(const orc_mesh *m);
void orc_bsr_structure(const orc_mesh *m, int32_t *ia, int32_t *ja, int32_t *iau);
/* Approximate Roe Jacobian; rhs_mode 0: B = -Conv (Euler3D), 1: B = -(Conv+Diss) */
void orc_jacobian(const orc_mesh *m, const orc_params *p, const double *q, const double *dt,
                  const double *res_conv, const double *res_diss, int rhs_mode, const int32_t *ia,
                  const int32_t *ja, const int32_t *iau, double *a, double *b);

/* Finite-difference Jacobian (Compute_Jacobian_FiniteDifference_Roe,
 * Euler3D/src/Roe_Jacobian.cpp:236-776) of the matrix-form Roe flux
 * (Compute_RoeFlux, Euler3D/src/Roe_Fluxes.cpp:552-905), eps 1e-8; for HLLC
 * and AUSM+ the same perturbation sequence differentiates the scheme's own
 * flux (the scheme-agnostic Compute_Jacobian_FiniteDifference,
 * MPFLOW3D/src/FiniteDifference_Jacobian.cpp:112-667).  method:
 * ORC_FD_CENTRAL / FORWARD / BACKWARD / ALTERNATE (forward on even
 * iterations, backward on odd); boundary edges are central.  q is [5][N+B]
 * with the ghost states of the last boundary-condition pass; grad/phi are
 * used for second order.  Returns the count of negative ghost pressures. */
enum { ORC_FD_CENTRAL = 0, ORC_FD_FORWARD = 1, ORC_FD_BACKWARD = 2, ORC_FD_ALTERNATE = 3 };
int orc_jacobian_fd(const orc_mesh *m, const orc_params *p, const double *q, const double *grad,
                    const double *phi, const double *dt, const double *res_conv, const double *res_diss,
                    int rhs_mode, int method, int iter, int add_time, const int32_t *ia, const int32_t *ja,
                    const int32_t *iau, double *a, double *b);
/* One edge of the matrix-form Roe flux (Compute_RoeFlux), for closed-form tests. */
double orc_edge_flux_matrix(const orc_params *p, const double *s, const double *ql, const double *qr,
                            double *conv, double *diss);

/* Reference linear solvers on BSR (x [N][5]) */
double orc_block_gs(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
                    const double *b, int sweeps, double *x);
void orc_lusgs(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
               const double *b, int sweeps, double *x);

/* Block ILU(0) in place: lower blocks L_ik (unit lower, multiplied by the
 * inverse of the pivot), upper blocks; dinv [N][25] inverse diagonal blocks. */
int orc_ilu0(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, double *lu, double *dinv);
void orc_ilu0_apply(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *lu,
                    const double *dinv, const double *r, double *z);
void orc_bsr_mv(int32_t n, const int32_t *ia, const int32_t *ja, const double *a, const double *x, double *y);

/* Restarted right-preconditioned FGMRES(m) with CGS2 orthogonalisation.
 * precond 0: none, 1: block-Jacobi, 2: ILU(0).  Stops when the Arnoldi
 * residual estimate <= rtol*||b||.  Returns total Krylov iterations. */
int orc_fgmres(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
               const double *b, int m, int max_restarts, double rtol, int precond, double *x,
               double *final_res);

typedef struct orc_solve_opts {
    int32_t max_iter;
    double tol;            /* absolute RMS of Conv+Diss */
    int32_t krylov_m;
    int32_t max_restarts;
    double krylov_rtol;
    int32_t precond;       /* as orc_fgmres */
    int32_t gradient;      /* ORC_GRAD_* */
    int32_t first_order_iters;
    double relaxation;
} orc_solve_opts;

typedef struct orc_solve_result {
    int32_t iterations;    /* Newton (pseudo-time) iterations */
    int32_t krylov_total;
    int32_t converged;
    double final_rms;
    double first_rms;
    double force[3];
    double lift, drag, pressure_integral;
} orc_solve_result;

/* Full steady implicit solve (Euler3D/src/Solver_Steady_Implicit.cpp:45-168
 * with B = -(Conv+Diss) and FGMRES instead of block GS).  q [5][N+B] is the
 * initial state and is updated.  rms_hist[max_iter] may be NULL. */
int orc_solve_steady(const orc_mesh *m, const orc_params *p, const orc_solve_opts *o, double alpha_deg,
                     int wall_type_tag_filter, const int32_t *bedge_tag, double *q, double *rms_hist,
                     int32_t *krylov_hist, orc_solve_result *res);

/* Pressure force on boundary edges whose tag equals wall_tag (or all walls if
 * wall_tag <= 0): F = sum p_node * S_out. */
void orc_forces(const orc_mesh *m, const orc_params *p, const double *q, const int32_t *bedge_tag,
                int wall_tag, double alpha_deg, double *force, double *lift, double *drag,
                double *pint);

#ifdef __cplusplus
}
#endif

#endif
