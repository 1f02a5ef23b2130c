/*
 * oracle.c - CPU restatement of the reference hot path (test infrastructure
 * only; see oracle.h for the parity status of each function).
 *
 * Follows Euler3D's conservative-variable path with the GENERIC
 * non-dimensionalisation and an ideal gas.  Expressions keep the reference's
 * operation order so that, compiled without FMA contraction, results are
 * bit-identical to the compiled reference.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MINR(a, b) ((a) < (b) ? (a) : (b)) /* UTILS/include/Trim_Utils.h:95 */
#define MAXR(a, b) ((a) > (b) ? (a) : (b)) /* UTILS/include/Trim_Utils.h:102 */

typedef struct {
    double rho, u, v, w, et, q2, p, ht, c, ubar;
} eos_t;

/* Compute_EOS_Variables_Face, Euler3D/src/Material.cpp:216-240 (GENERIC, CONSERVATIVE) */
static void eos_face(double g, const double *Q, double nx, double ny, double nz, eos_t *e) {
    e->rho = Q[0];
    e->u = Q[1] / e->rho;
    e->v = Q[2] / e->rho;
    e->w = Q[3] / e->rho;
    e->et = Q[4] / e->rho;
    e->q2 = e->u * e->u + e->v * e->v + e->w * e->w;
    e->p = (g - 1.0) * (e->rho * e->et - 0.5 * e->rho * e->q2);
    e->ht = (e->rho * e->et + e->p) / e->rho;
    e->c = sqrt((g * e->p) / e->rho);
    e->ubar = e->u * nx + e->v * ny + e->w * nz;
}

/* Vector3D::magnitude / normalize, MATH/src/Vector3D.cpp:216-229 */
static double area_normal(const double *s, double *n) {
    double area = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    double mag = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    n[0] = s[0];
    n[1] = s[1];
    n[2] = s[2];
    if (mag > 1.0e-20) {
        n[0] = s[0] / mag;
        n[1] = s[1] / mag;
        n[2] = s[2] / mag;
    }
    return area;
}

/* Get_SpeedSound(u,v,w,ht), Euler3D/src/Material.cpp:939-975 (GENERIC) */
static double roe_c(double g, double u, double v, double w, double ht) {
    double phi = 0.5 * (g - 1.0) * (u * u + v * v + w * w);
    return sqrt((g - 1.0) * ht - phi);
}

/* Roe_EntropyFix (Kermani), Euler3D/src/Roe_EntropyFix.cpp:19-59 */
static void entropy_fix(int fix, double ubar_L, double c_L, double ubar_R, double c_R, double ubar, double c,
                        double *E) {
    if (fix == 1) {
        double lL, lR, l, eps;
        E[0] = fabs(ubar);
        lL = ubar_L + c_L;
        lR = ubar_R + c_R;
        l = ubar + c;
        eps = 4.0 * MAXR(0.0, MAXR((l - lL), (lR - l)));
        E[1] = (fabs(l) < eps) ? (l * l + eps * eps) / (2.0 * eps) : fabs(l);
        lL = ubar_L - c_L;
        lR = ubar_R - c_R;
        l = ubar - c;
        eps = 4.0 * MAXR(0.0, MAXR((l - lL), (lR - l)));
        E[2] = (fabs(l) < eps) ? (l * l + eps * eps) / (2.0 * eps) : fabs(l);
    } else {
        E[0] = fabs(ubar);
        E[1] = fabs(ubar + c);
        E[2] = fabs(ubar - c);
    }
}

/* Compute_Flux_Roe_Optimized, Euler3D/src/FluxTest.cpp:3026-3200.
 * Returns area*max|lambda|; E[3] receives the three eigenvalue magnitudes. */
static double roe_flux(const orc_params *P, const double *s, const double *QL, const double *QR, double *conv,
                       double *diss, double *E) {
    const double g = P->gamma;
    double n[3];
    double area = area_normal(s, n);
    const double nx = n[0], ny = n[1], nz = n[2];
    eos_t L, R;
    eos_face(g, QL, nx, ny, nz, &L);
    eos_face(g, QR, nx, ny, nz, &R);
    double fL[5], fR[5];
    fL[0] = L.rho * L.ubar;
    fL[1] = L.u * fL[0] + L.p * nx;
    fL[2] = L.v * fL[0] + L.p * ny;
    fL[3] = L.w * fL[0] + L.p * nz;
    fL[4] = L.ht * fL[0];
    fR[0] = R.rho * R.ubar;
    fR[1] = R.u * fR[0] + R.p * nx;
    fR[2] = R.v * fR[0] + R.p * ny;
    fR[3] = R.w * fR[0] + R.p * nz;
    fR[4] = R.ht * fR[0];
    double rho = sqrt(R.rho * L.rho);
    double sigma = rho / (L.rho + rho);
    double u = L.u + sigma * (R.u - L.u);
    double v = L.v + sigma * (R.v - L.v);
    double w = L.w + sigma * (R.w - L.w);
    double ht = L.ht + sigma * (R.ht - L.ht);
    double c = roe_c(g, u, v, w, ht);
    double ubar = u * nx + v * ny + w * nz;
    entropy_fix(P->entropy_fix, L.ubar, L.c, R.ubar, R.c, ubar, c, E);
    const double E0 = E[0], E1 = E[1], E2 = E[2];
    double maxlambda = MAXR(E0, MAXR(E1, E2));
    /* M: Compute_Transformation_Conservative_To_RUP, Euler3D/src/Material.cpp:1172-1235 */
    double phi = 0.5 * (g - 1.0) * (u * u + v * v + w * w);
    double M[5][5] = {{1.0, 0.0, 0.0, 0.0, 0.0},
                      {u, rho, 0.0, 0.0, 0.0},
                      {v, 0.0, rho, 0.0, 0.0},
                      {w, 0.0, 0.0, rho, 0.0},
                      {phi / (g - 1.0), rho * u, rho * v, rho * w, 1.0 / (g - 1.0)}};
    double dQ[5] = {R.rho - L.rho, R.u - L.u, R.v - L.v, R.w - L.w, R.p - L.p};
    double dU = dQ[1] * nx + dQ[2] * ny + dQ[3] * nz;
    double dw[5];
    dw[0] = E0 * dQ[0] + (E1 + E2 - 2.0 * E0) * dQ[4] / (2.0 * c * c) + rho * (E1 - E2) * dU / (2.0 * c);
    dw[1] = E0 * dQ[1] + 0.5 * nx * dU * ((E1 + E2) - 2.0 * E0) + nx * (E1 - E2) * dQ[4] / (2.0 * rho * c);
    dw[2] = E0 * dQ[2] + 0.5 * ny * dU * ((E1 + E2) - 2.0 * E0) + ny * (E1 - E2) * dQ[4] / (2.0 * rho * c);
    dw[3] = E0 * dQ[3] + 0.5 * nz * dU * ((E1 + E2) - 2.0 * E0) + nz * (E1 - E2) * dQ[4] / (2.0 * rho * c);
    dw[4] = 0.5 * dQ[4] * (E1 + E2) + 0.5 * rho * c * dU * (E1 - E2);
    for (int i = 0; i < 5; i++) {
        double fa = 0.0; /* MC_Matrix_Mul_Vector, MATH/src/MC_Iterative_Jacobi.c:174-187 */
        for (int j = 0; j < 5; j++) fa += M[i][j] * dw[j];
        conv[i] = 0.5 * (fL[i] + fR[i]) * area;
        diss[i] = -0.5 * fa * area;
    }
    return area * maxlambda;
}

/* Compute_HLLCFlux, Euler3D/src/HLLC_Fluxes.cpp:143-326.  Returns
 * area*max|lambda| (NaN flux if no branch applies: reference error()). */
static double hllc_flux(const orc_params *P, const double *s, const double *QL, const double *QR, double *F) {
    const double g = P->gamma;
    double n[3];
    double area = area_normal(s, n);
    const double nx = n[0], ny = n[1], nz = n[2];
    eos_t L, R;
    eos_face(g, QL, nx, ny, nz, &L);
    eos_face(g, QR, nx, ny, nz, &R);
    double rho = sqrt(R.rho * L.rho);
    double sigma = rho / (L.rho + rho);
    double u = L.u + sigma * (R.u - L.u);
    double v = L.v + sigma * (R.v - L.v);
    double w = L.w + sigma * (R.w - L.w);
    double ht = L.ht + sigma * (R.ht - L.ht);
    double phi = 0.5 * (g - 1.0) * (u * u + v * v + w * w);
    double c = (g - 1.0) * ht - phi;
    c = sqrt(c);
    double ubar = u * nx + v * ny + w * nz;
    double maxlambda = MAXR(fabs(ubar), MAXR(fabs(ubar + c), fabs(ubar - c)));
    double eig5_L = L.ubar - L.c, eig4_R = R.ubar + R.c, eig4 = ubar + c, eig5 = ubar - c;
    double SL = MINR(eig5_L, eig5);
    double SR = MAXR(eig4_R, eig4);
    double SM = (R.rho * R.ubar * (SR - R.ubar) - L.rho * L.ubar * (SL - L.ubar) + L.p - R.p) /
                (R.rho * (SR - R.ubar) - L.rho * (SL - L.ubar));
    if (SL > 0.0) {
        F[0] = L.rho * L.ubar;
        F[1] = L.u * F[0] + L.p * nx;
        F[2] = L.v * F[0] + L.p * ny;
        F[3] = L.w * F[0] + L.p * nz;
        F[4] = L.ht * F[0];
    } else if (SR < 0.0) {
        F[0] = R.rho * R.ubar;
        F[1] = R.u * F[0] + R.p * nx;
        F[2] = R.v * F[0] + R.p * ny;
        F[3] = R.w * F[0] + R.p * nz;
        F[4] = R.ht * F[0];
    } else if ((SL <= 0.0) && (SM > 0.0)) {
        double pstar = L.p + L.rho * (L.ubar - SL) * (L.ubar - SM);
        double omega = 1.0 / (SL - SM);
        double cnst1 = SL - L.ubar;
        double cnst2 = pstar - L.p;
        double rs = omega * (cnst1 * L.rho);
        double rus = omega * (cnst1 * L.rho * L.u + cnst2 * nx);
        double rvs = omega * (cnst1 * L.rho * L.v + cnst2 * ny);
        double rws = omega * (cnst1 * L.rho * L.w + cnst2 * nz);
        double res = omega * (cnst1 * L.rho * L.et - L.p * L.ubar + pstar * SM);
        F[0] = rs * SM;
        F[1] = rus * SM + pstar * nx;
        F[2] = rvs * SM + pstar * ny;
        F[3] = rws * SM + pstar * nz;
        F[4] = (res + pstar) * SM;
    } else if ((SM <= 0.0) && (SR > 0.0)) {
        double pstar = R.p + R.rho * (R.ubar - SR) * (R.ubar - SM);
        double omega = 1.0 / (SR - SM);
        double cnst1 = SR - R.ubar;
        double cnst2 = pstar - R.p;
        double rs = omega * (cnst1 * R.rho);
        double rus = omega * (cnst1 * R.rho * R.u + cnst2 * nx);
        double rvs = omega * (cnst1 * R.rho * R.v + cnst2 * ny);
        double rws = omega * (cnst1 * R.rho * R.w + cnst2 * nz);
        double res = omega * (cnst1 * R.rho * R.et - R.p * R.ubar + pstar * SM);
        F[0] = rs * SM;
        F[1] = rus * SM + pstar * nx;
        F[2] = rvs * SM + pstar * ny;
        F[3] = rws * SM + pstar * nz;
        F[4] = (res + pstar) * SM;
    } else {
        for (int i = 0; i < 5; i++) F[i] = NAN;
    }
    for (int i = 0; i < 5; i++) F[i] *= area;
    return area * maxlambda;
}

/* AUSM+ (M.-S. Liou, J. Comput. Phys. 129:364-382, 1996), eqs. (19a-b) split
 * Mach numbers with beta = 1/8, (21a-b) split pressures with alpha = 3/16,
 * interface sound speed (30)-(31) from the critical speed c*^2 = 2(g-1)/(g+1) H.
 * The reference's AUSM files are empty stubs (Euler3D/src/AUSM_Fluxes.cpp:24-54).
 * Time-step eigenvalue as the HLLC routine (Roe-averaged). */
static double ausmp_mach_p(double M) {
    return fabs(M) >= 1.0 ? 0.5 * (M + fabs(M)) : 0.25 * (M + 1.0) * (M + 1.0) + 0.125 * (M * M - 1.0) * (M * M - 1.0);
}
static double ausmp_mach_m(double M) {
    return fabs(M) >= 1.0 ? 0.5 * (M - fabs(M)) : -0.25 * (M - 1.0) * (M - 1.0) - 0.125 * (M * M - 1.0) * (M * M - 1.0);
}
static double ausmp_pres_p(double M) {
    return fabs(M) >= 1.0 ? (M > 0.0 ? 1.0 : 0.0)
                          : 0.25 * (M + 1.0) * (M + 1.0) * (2.0 - M) + 0.1875 * M * (M * M - 1.0) * (M * M - 1.0);
}
static double ausmp_pres_m(double M) {
    return fabs(M) >= 1.0 ? (M < 0.0 ? 1.0 : 0.0)
                          : 0.25 * (M - 1.0) * (M - 1.0) * (2.0 + M) - 0.1875 * M * (M * M - 1.0) * (M * M - 1.0);
}
static double ausmp_flux(const orc_params *P, const double *s, const double *QL, const double *QR, double *F) {
    const double g = P->gamma;
    double n[3];
    double area = area_normal(s, n);
    const double nx = n[0], ny = n[1], nz = n[2];
    eos_t L, R;
    eos_face(g, QL, nx, ny, nz, &L);
    eos_face(g, QR, nx, ny, nz, &R);
    double cs2L = 2.0 * (g - 1.0) / (g + 1.0) * L.ht;
    double cs2R = 2.0 * (g - 1.0) / (g + 1.0) * R.ht;
    double csL = sqrt(cs2L), csR = sqrt(cs2R);
    double chL = cs2L / MAXR(csL, L.ubar);
    double chR = cs2R / MAXR(csR, -R.ubar);
    double c12 = MINR(chL, chR);
    double ML = L.ubar / c12, MR = R.ubar / c12;
    double m12 = ausmp_mach_p(ML) + ausmp_mach_m(MR);
    double p12 = ausmp_pres_p(ML) * L.p + ausmp_pres_m(MR) * R.p;
    double mp = 0.5 * (m12 + fabs(m12)), mm = 0.5 * (m12 - fabs(m12));
    double mL = c12 * mp * L.rho, mR = c12 * mm * R.rho;
    F[0] = mL + mR;
    F[1] = mL * L.u + mR * R.u + p12 * nx;
    F[2] = mL * L.v + mR * R.v + p12 * ny;
    F[3] = mL * L.w + mR * R.w + p12 * nz;
    F[4] = mL * L.ht + mR * R.ht;
    for (int i = 0; i < 5; i++) F[i] *= area;
    double rho = sqrt(R.rho * L.rho);
    double sigma = rho / (L.rho + rho);
    double u = L.u + sigma * (R.u - L.u);
    double v = L.v + sigma * (R.v - L.v);
    double w = L.w + sigma * (R.w - L.w);
    double ht = L.ht + sigma * (R.ht - L.ht);
    double c = roe_c(g, u, v, w, ht);
    double ubar = u * nx + v * ny + w * nz;
    return area * MAXR(fabs(ubar), MAXR(fabs(ubar + c), fabs(ubar - c)));
}

double orc_edge_flux(const orc_params *p, const double *s, const double *ql, const double *qr, double *conv,
                     double *diss) {
    double E[3];
    for (int i = 0; i < 5; i++) conv[i] = diss[i] = 0.0;
    if (p->flux == ORC_ROE) return roe_flux(p, s, ql, qr, conv, diss, E);
    if (p->flux == ORC_HLLC) return hllc_flux(p, s, ql, qr, conv);
    return ausmp_flux(p, s, ql, qr, conv);
}

/* Compute_SecondOrderReconstructQ, Euler3D/src/HigherOrderReconstructQ.cpp:20-71 */
static void reconstruct(const orc_mesh *m, const orc_params *P, const double *q, const double *grad,
                        const double *phi, int nl, int nr, double *QL, double *QR) {
    const int NT = m->n_node + m->n_bedge, N = m->n_node;
    const double frac = (P->limiter_order == 2) ? 0.5 : 0.0;
    const double *X = m->xyz;
    double RxL = 0.5 * (X[3 * nr + 0] - X[3 * nl + 0]);
    double RyL = 0.5 * (X[3 * nr + 1] - X[3 * nl + 1]);
    double RzL = 0.5 * (X[3 * nr + 2] - X[3 * nl + 2]);
    double RxR = -RxL, RyR = -RyL, RzR = -RzL;
    for (int k = 0; k < 5; k++) {
        double pL = (P->limiter != ORC_LIM_NONE) ? phi[(size_t)k * N + nl] : 1.0;
        double pR = (P->limiter != ORC_LIM_NONE) ? phi[(size_t)k * N + nr] : 1.0;
        const double *gx = grad + (size_t)(3 * k) * N, *gy = gx + N, *gz = gy + N;
        double qL = q[(size_t)k * NT + nl], qR = q[(size_t)k * NT + nr];
        QL[k] = qL + pL * (frac * (qR - qL) + (1.0 - frac) * (gx[nl] * RxL + gy[nl] * RyL + gz[nl] * RzL));
        QR[k] = qR + pR * (frac * (qL - qR) + (1.0 - frac) * (gx[nr] * RxR + gy[nr] * RyR + gz[nr] * RzR));
    }
}

/* Compute_Residual_Roe_New (Euler3D/src/FluxTest.cpp:7968-8151) and
 * Compute_Residual_HLLC (Euler3D/src/HLLC_Fluxes.cpp:519-573); the residual
 * transform at FluxTest.cpp:1273-1340 is the identity for conservative
 * variables. */
void orc_residual(const orc_mesh *m, const orc_params *p, const double *q, const double *grad,
                  const double *phi, double *res_conv, double *res_diss, double *dt_acc, double *eig) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    double QL[5], QR[5], conv[5], diss[5], E[3];
    if (eig) {
        eig[0] = eig[2] = eig[4] = DBL_MAX;
        eig[1] = eig[3] = eig[5] = DBL_MIN;
    }
    for (int pass = 0; pass < 2; pass++) {
        const int ne = pass ? m->n_bedge : m->n_edge;
        const int32_t *en = pass ? m->bedge : m->edge;
        const double *es = pass ? m->bedge_area : m->edge_area;
        for (int e = 0; e < ne; e++) {
            const int nl = en[2 * e], nr = en[2 * e + 1];
            if (!pass && p->order == 2) {
                reconstruct(m, p, q, grad, phi, nl, nr, QL, QR);
            } else {
                for (int k = 0; k < 5; k++) {
                    QL[k] = q[(size_t)k * NT + nl];
                    QR[k] = q[(size_t)k * NT + nr];
                }
            }
            double lam;
            for (int k = 0; k < 5; k++) conv[k] = diss[k] = 0.0;
            if (p->flux == ORC_ROE) {
                lam = roe_flux(p, es + 3 * (size_t)e, QL, QR, conv, diss, E);
                if (eig) {
                    eig[0] = MINR(eig[0], E[0]);
                    eig[1] = MAXR(eig[1], E[0]);
                    eig[2] = MINR(eig[2], E[1]);
                    eig[3] = MAXR(eig[3], E[1]);
                    eig[4] = MINR(eig[4], E[2]);
                    eig[5] = MAXR(eig[5], E[2]);
                }
            } else if (p->flux == ORC_HLLC) {
                lam = hllc_flux(p, es + 3 * (size_t)e, QL, QR, conv);
            } else {
                lam = ausmp_flux(p, es + 3 * (size_t)e, QL, QR, conv);
            }
            if (dt_acc) {
                dt_acc[nl] += lam;
                if (nr < N) dt_acc[nr] += lam;
            }
            for (int k = 0; k < 5; k++) {
                res_conv[(size_t)k * N + nl] += conv[k];
                if (p->flux == ORC_ROE) res_diss[(size_t)k * N + nl] += diss[k];
                if (!pass) {
                    res_conv[(size_t)k * N + nr] -= conv[k];
                    if (p->flux == ORC_ROE) res_diss[(size_t)k * N + nr] -= diss[k];
                }
            }
        }
    }
}

/* Compute_Characteristic_BoundaryCondition_New, Euler3D/src/BC.cpp:381-613
 * (conservative variables, gauge pressure 0), applied per boundary edge as
 * Apply_Characteristic_Boundary_Condition, Euler3D/src/BC.cpp:1440-1487. */
static int bc_edge(const orc_params *P, const double *QL, const double *QR, const double *s, int type, int iter,
                   double *QB) {
    const double g = P->gamma;
    double n[3];
    area_normal(s, n);
    const double nx = n[0], ny = n[1], nz = n[2];
    eos_t I, Z;
    eos_face(g, QL, nx, ny, nz, &I);
    double Qa[5];
    for (int i = 0; i < 5; i++) Qa[i] = 0.5 * (QL[i] + QR[i]);
    eos_face(g, Qa, nx, ny, nz, &Z);
    const double l1 = Z.ubar, l4 = l1 + Z.c, l5 = l1 - Z.c;
    const double infP = P->inf_p;
    double pb = infP, rb = P->inf_rho, ub = P->inf_u, vb = P->inf_v, wb = P->inf_w;
    int inst = type;
    switch (type) {
    case ORC_BC_EULER_WALL:
        break;
    case ORC_BC_FAR_FIELD:
        if ((l1 <= 0.0) && (l4 <= 0.0) && (l5 <= 0.0)) inst = ORC_BC_FF_SUPERSONIC_INFLOW;
        else if ((l1 >= 0.0) && (l4 >= 0.0) && (l5 >= 0.0)) inst = ORC_BC_FF_SUPERSONIC_OUTFLOW;
        else if ((l1 >= 0.0) && (l5 <= 0.0) && (l4 >= 0.0)) inst = ORC_BC_FF_SUBSONIC_OUTFLOW;
        else if ((l1 <= 0.0) && (l4 >= 0.0) && (l5 <= 0.0)) inst = ORC_BC_FF_SUBSONIC_INFLOW;
        else return 1;
        break;
    case ORC_BC_INFLOW:
        if ((l1 <= 0.0) && (l4 >= 0.0) && (l5 <= 0.0)) inst = ORC_BC_SUBSONIC_INFLOW;
        else if ((l1 <= 0.0) && (l4 <= 0.0) && (l5 <= 0.0)) inst = ORC_BC_SUPERSONIC_INFLOW;
        else inst = (fabs(l1) < Z.c) ? ORC_BC_SUBSONIC_INFLOW : ORC_BC_SUPERSONIC_INFLOW;
        break;
    case ORC_BC_OUTFLOW:
        if ((l1 >= 0.0) && (l5 <= 0.0) && (l4 >= 0.0)) inst = ORC_BC_SUBSONIC_OUTFLOW;
        else if ((l1 >= 0.0) && (l4 >= 0.0) && (l5 >= 0.0)) inst = ORC_BC_SUPERSONIC_OUTFLOW;
        else inst = (fabs(l1) < Z.c) ? ORC_BC_SUBSONIC_OUTFLOW : ORC_BC_SUPERSONIC_OUTFLOW;
        break;
    case ORC_BC_SUBSONIC_INFLOW:
    case ORC_BC_SUBSONIC_OUTFLOW:
    case ORC_BC_SUPERSONIC_INFLOW:
    case ORC_BC_SUPERSONIC_OUTFLOW:
        break;
    default:
        return 1;
    }
    switch (inst) {
    case ORC_BC_EULER_WALL:
        pb = I.p + Z.rho * Z.c * I.ubar;
        if (iter < P->zpg_iter && pb < 0.0) pb = I.p;
        rb = I.rho + (pb - I.p) / (Z.c * Z.c);
        ub = I.u - nx * (pb - I.p) / (Z.rho * Z.c);
        vb = I.v - ny * (pb - I.p) / (Z.rho * Z.c);
        wb = I.w - nz * (pb - I.p) / (Z.rho * Z.c);
        break;
    case ORC_BC_SUBSONIC_INFLOW:
    case ORC_BC_SUPERSONIC_INFLOW:
    case ORC_BC_FF_SUPERSONIC_INFLOW:
        break;
    case ORC_BC_SUBSONIC_OUTFLOW:
    case ORC_BC_FF_SUBSONIC_OUTFLOW:
        pb = (inst == ORC_BC_SUBSONIC_OUTFLOW) ? P->outflow_p : infP;
        rb = I.rho + (pb - I.p) / (Z.c * Z.c);
        ub = I.u - nx * (pb - I.p) / (Z.rho * Z.c);
        vb = I.v - ny * (pb - I.p) / (Z.rho * Z.c);
        wb = I.w - nz * (pb - I.p) / (Z.rho * Z.c);
        break;
    case ORC_BC_SUPERSONIC_OUTFLOW:
    case ORC_BC_FF_SUPERSONIC_OUTFLOW:
        pb = I.p;
        rb = I.rho;
        ub = I.u;
        vb = I.v;
        wb = I.w;
        break;
    case ORC_BC_FF_SUBSONIC_INFLOW: {
        double temp = nx * (P->inf_u - I.u) + ny * (P->inf_v - I.v) + nz * (P->inf_w - I.w);
        pb = 0.5 * (infP + I.p - Z.rho * Z.c * temp);
        rb = P->inf_rho + (pb - infP) / (Z.c * Z.c);
        ub = P->inf_u + nx * (pb - infP) / (Z.rho * Z.c);
        vb = P->inf_v + ny * (pb - infP) / (Z.rho * Z.c);
        wb = P->inf_w + nz * (pb - infP) / (Z.rho * Z.c);
        break;
    }
    default:
        return 1;
    }
    int rv = (pb < 0.0) ? 1 : 0;
    /* Get_TotalEnergy, Euler3D/src/Material.cpp:1012-1040 */
    double et = ub * ub + vb * vb + wb * wb;
    et = pb / ((g - 1.0) * rb) + 0.5 * et;
    QB[0] = rb;
    QB[1] = rb * ub;
    QB[2] = rb * vb;
    QB[3] = rb * wb;
    QB[4] = rb * et;
    return rv;
}

int orc_bc(const orc_mesh *m, const orc_params *p, int iter, double *q) {
    const int NT = m->n_node + m->n_bedge;
    int bad = 0;
    for (int e = 0; e < m->n_bedge; e++) {
        const int nl = m->bedge[2 * e], nr = m->bedge[2 * e + 1];
        double QL[5], QR[5], QB[5];
        for (int k = 0; k < 5; k++) {
            QL[k] = q[(size_t)k * NT + nl];
            QR[k] = q[(size_t)k * NT + nr];
            QB[k] = QR[k];
        }
        bad += bc_edge(p, QL, QR, m->bedge_area + 3 * (size_t)e, m->bedge_type[e], iter, QB);
        for (int k = 0; k < 5; k++) q[(size_t)k * NT + nr] = QB[k];
    }
    return bad;
}

/* Compute_Least_Square_Gradient(_Coefficients), unweighted,
 * Euler3D/src/Gradient.cpp:137-295 */
void orc_lsq_gradient(const orc_mesh *m, const double *q, double *grad) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    const double *X = m->xyz;
    for (int i = 0; i < N; i++) {
        double r11 = 0, r12 = 0, r13 = 0, r22 = 0, r23 = 0, ro22 = 0, ro33 = 0, r33 = 0;
        const int b = m->n2n_ia[i], e = m->n2n_ia[i + 1];
        for (int k = b; k < e; k++) {
            int j = m->n2n_ja[k];
            double dx = 1.0 * (X[3 * j] - X[3 * i]), dy = 1.0 * (X[3 * j + 1] - X[3 * i + 1]),
                   dz = 1.0 * (X[3 * j + 2] - X[3 * i + 2]);
            r11 += dx * dx;
            r12 += dx * dy;
            r13 += dx * dz;
        }
        for (int k = b; k < e; k++) {
            int j = m->n2n_ja[k];
            double dx = 1.0 * (X[3 * j] - X[3 * i]), dy = 1.0 * (X[3 * j + 1] - X[3 * i + 1]),
                   dz = 1.0 * (X[3 * j + 2] - X[3 * i + 2]);
            double t = dy - (r12 / r11) * dx;
            r22 += t * t;
            r23 += t * dz;
            ro22 += t * dy;
        }
        for (int k = b; k < e; k++) {
            int j = m->n2n_ja[k];
            double dx = 1.0 * (X[3 * j] - X[3 * i]), dy = 1.0 * (X[3 * j + 1] - X[3 * i + 1]),
                   dz = 1.0 * (X[3 * j + 2] - X[3 * i + 2]);
            double t = dz - (r13 / r11) * dx - (r23 / r22) * (dy - (r12 / r11) * dx);
            r33 += t * t;
            ro33 += t * dz;
        }
        (void)r33;
        double gsum[15];
        for (int f = 0; f < 15; f++) gsum[f] = 0.0;
        for (int k = b; k < e; k++) {
            int j = m->n2n_ja[k];
            double alpha = 1.0;
            double dx = alpha * (X[3 * j] - X[3 * i]), dy = alpha * (X[3 * j + 1] - X[3 * i + 1]),
                   dz = alpha * (X[3 * j + 2] - X[3 * i + 2]);
            double Wz = (dz - (r13 / r11) * dx - (r23 / r22) * (dy - (r12 / r11) * dx)) / ro33;
            double Wy = (dy - (r12 / r11) * dx - r23 * Wz) / ro22;
            double Wx = (dx - r12 * Wy - r13 * Wz) / r11;
            for (int f = 0; f < 5; f++) {
                double d = q[(size_t)f * NT + j] - q[(size_t)f * NT + i];
                gsum[3 * f + 0] += Wx * alpha * d;
                gsum[3 * f + 1] += Wy * alpha * d;
                gsum[3 * f + 2] += Wz * alpha * d;
            }
        }
        for (int f = 0; f < 15; f++) grad[(size_t)f * N + i] = gsum[f];
    }
}

/* Median-dual edge-based Green-Gauss (the reference has no 3D Green-Gauss;
 * the enum GRADINET_METHOD_GREEN_GAUSS is unused):
 *   grad_i = (1/V_i) [ sum_edges +-1/2 (Q_L+Q_R) S_LR + sum_bnd Q_i S_b ],
 * accumulated in edge order (interior, then boundary). */
void orc_gg_gradient(const orc_mesh *m, const double *q, double *grad) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    memset(grad, 0, sizeof(double) * 15 * (size_t)N);
    for (int e = 0; e < m->n_edge; e++) {
        const int L = m->edge[2 * e], R = m->edge[2 * e + 1];
        const double *s = m->edge_area + 3 * (size_t)e;
        for (int f = 0; f < 5; f++) {
            double qf = 0.5 * (q[(size_t)f * NT + L] + q[(size_t)f * NT + R]);
            for (int d = 0; d < 3; d++) {
                grad[(size_t)(3 * f + d) * N + L] += qf * s[d];
                grad[(size_t)(3 * f + d) * N + R] -= qf * s[d];
            }
        }
    }
    for (int e = 0; e < m->n_bedge; e++) {
        const int L = m->bedge[2 * e];
        const double *s = m->bedge_area + 3 * (size_t)e;
        for (int f = 0; f < 5; f++) {
            double qf = q[(size_t)f * NT + L];
            for (int d = 0; d < 3; d++) grad[(size_t)(3 * f + d) * N + L] += qf * s[d];
        }
    }
    for (int f = 0; f < 15; f++)
        for (int i = 0; i < N; i++) grad[(size_t)f * N + i] = grad[(size_t)f * N + i] / m->vol[i];
}

/* Second-order face state of node c towards neighbour n without limiter, as
 * in Euler3D/src/Limiters.cpp:130-143 / 205-218. */
static void face_q(const orc_mesh *m, const double *q, const double *grad, double frac, int c, int nb,
                   const double *phic, double *Qc) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    const double *X = m->xyz;
    double Rx = 0.5 * (X[3 * nb + 0] - X[3 * c + 0]);
    double Ry = 0.5 * (X[3 * nb + 1] - X[3 * c + 1]);
    double Rz = 0.5 * (X[3 * nb + 2] - X[3 * c + 2]);
    for (int k = 0; k < 5; k++) {
        const double *gx = grad + (size_t)(3 * k) * N, *gy = gx + N, *gz = gy + N;
        double qc = q[(size_t)k * NT + c], qn = q[(size_t)k * NT + nb];
        if (phic)
            Qc[k] = qc + phic[k] * (frac * (qn - qc) + (1.0 - frac) * (gx[c] * Rx + gy[c] * Ry + gz[c] * Rz));
        else
            Qc[k] = qc + frac * (qn - qc) + (1.0 - frac) * (gx[c] * Rx + gy[c] * Ry + gz[c] * Rz);
    }
}

/* Compute_RoeVariables, Euler3D/src/Roe_Fluxes.cpp:3609-3634 */
static double roe_pressure(double g, const double *QL, const double *QR) {
    eos_t L, R;
    eos_face(g, QL, 0.0, 0.0, 0.0, &L);
    eos_face(g, QR, 0.0, 0.0, 0.0, &R);
    double rho = sqrt(R.rho * L.rho);
    double sigma = rho / (L.rho + rho);
    double u = L.u + sigma * (R.u - L.u);
    double v = L.v + sigma * (R.v - L.v);
    double w = L.w + sigma * (R.w - L.w);
    double ht = L.ht + sigma * (R.ht - L.ht);
    double Q[5];
    Q[0] = rho;
    Q[1] = rho * u;
    Q[2] = rho * v;
    Q[3] = rho * w;
    Q[4] = (rho / g) * (ht + 0.5 * (g - 1.0) * (u * u + v * v + w * w));
    return (g - 1.0) * (Q[4] - (0.5 / Q[0]) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]));
}

/* Compute_Limiter and its variants, Euler3D/src/Limiters.cpp:20-356 */
void orc_limiter(const orc_mesh *m, const orc_params *p, const double *q, const double *grad, double *phi) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    const double frac = (p->limiter_order == 2) ? 0.5 : 0.0;
    const double g = p->gamma;
    const double K3 = pow(p->venkat_k, 3.0);
    for (int c = 0; c < N; c++) {
        double ph[5] = {1.0, 1.0, 1.0, 1.0, 1.0};
        double Qmin[5], Qmax[5], Qt[5], Qc[5];
        for (int k = 0; k < 5; k++) Qmin[k] = Qmax[k] = Qt[k] = q[(size_t)k * NT + c];
        const int b = m->n2n_ia[c], e = m->n2n_ia[c + 1];
        if (p->limiter == ORC_LIM_VENKAT || p->limiter == ORC_LIM_BJ) {
            for (int n = b; n < e; n++) {
                int j = m->n2n_ja[n];
                for (int k = 0; k < 5; k++) {
                    Qmin[k] = MINR(Qmin[k], q[(size_t)k * NT + j]);
                    Qmax[k] = MAXR(Qmax[k], q[(size_t)k * NT + j]);
                }
            }
            double eps2 = K3;
            eps2 = 6.0 * eps2 * m->vol[c] / M_PI;
            for (int n = b; n < e; n++) {
                int j = m->n2n_ja[n];
                face_q(m, q, grad, frac, c, j, NULL, Qc);
                for (int k = 0; k < 5; k++) {
                    if (p->limiter == ORC_LIM_VENKAT) {
                        if ((Qc[k] - Qt[k]) > 0.0) {
                            double dP = Qmax[k] - Qt[k], dM = Qc[k] - Qt[k];
                            double t = (dP * dP + eps2) + 2.0 * (dM * dP);
                            t = t / (dP * dP + 2.0 * dM * dM + dM * dP + eps2);
                            ph[k] = MINR(ph[k], t);
                        } else if ((Qc[k] - Qt[k]) < 0.0) {
                            double dP = Qmin[k] - Qt[k], dM = Qc[k] - Qt[k];
                            double t = (dP * dP + eps2) + 2.0 * (dM * dP);
                            t = t / (dP * dP + 2.0 * dM * dM + dM * dP + eps2);
                            ph[k] = MINR(ph[k], t);
                        } else
                            ph[k] = MINR(ph[k], 1.0);
                    } else {
                        if ((Qc[k] - Qt[k]) > 0.0)
                            ph[k] = MINR(ph[k], MINR(1.0, (Qmax[k] - Qt[k]) / (Qc[k] - Qt[k])));
                        else if ((Qc[k] - Qt[k]) < 0.0)
                            ph[k] = MINR(ph[k], MINR(1.0, (Qmin[k] - Qt[k]) / (Qc[k] - Qt[k])));
                        else
                            ph[k] = MINR(ph[k], 1.0);
                    }
                    ph[k] = MAXR(0.0, ph[k]);
                }
            }
        }
        /* Compute_Limiter_PressureCorrection, Limiters.cpp:273-304 */
        for (int n = b; n < e; n++) {
            int j = m->n2n_ja[n];
            face_q(m, q, grad, frac, c, j, ph, Qc);
            double P = (g - 1.0) * (Qc[4] - (0.5 / Qc[0]) * (Qc[1] * Qc[1] + Qc[2] * Qc[2] + Qc[3] * Qc[3]));
            if (P < 0.0)
                for (int k = 0; k < 5; k++) ph[k] = 0.0;
        }
        for (int k = 0; k < 5; k++) phi[(size_t)k * N + c] = ph[k];
    }
    /* Roe pressure correction, sequential over nodes and neighbours,
     * Limiters.cpp:71-110 and 311-356 */
    if (p->flux == ORC_ROE) {
        for (int c = 0; c < N; c++) {
            for (int n = m->n2n_ia[c]; n < m->n2n_ia[c + 1]; n++) {
                int j = m->n2n_ja[n];
                double pc[5], pn[5], QL[5], QR[5];
                for (int k = 0; k < 5; k++) {
                    pc[k] = phi[(size_t)k * N + c];
                    pn[k] = phi[(size_t)k * N + j];
                }
                face_q(m, q, grad, frac, c, j, pc, QL);
                face_q(m, q, grad, frac, j, c, pn, QR);
                if (roe_pressure(g, QL, QR) < 0.0)
                    for (int k = 0; k < 5; k++) phi[(size_t)k * N + c] = phi[(size_t)k * N + j] = 0.0;
            }
        }
    }
    (void)NT;
}

/* Compute_DeltaT, Euler3D/src/Time_Step.cpp:23-50 (local time stepping) */
double orc_delta_t(const orc_mesh *m, const orc_params *p, int iter, const double *dt_acc, double *dt) {
    double CFL;
    if ((p->cfl_ramp > 1) && (p->cfl_max > p->cfl_min)) {
        if (iter < p->cfl_ramp)
            CFL = p->cfl_min + (p->cfl_max - p->cfl_min) * (((double)iter) / ((double)(p->cfl_ramp - 1)));
        else
            CFL = p->cfl_max;
    } else
        CFL = p->cfl_max;
    for (int i = 0; i < m->n_node; i++) dt[i] = m->vol[i] * CFL / dt_acc[i];
    return CFL;
}

int32_t orc_bsr_nnz(const orc_mesh *m) { return m->n_node + m->n2n_ia[m->n_node]; }

void orc_bsr_structure(const orc_mesh *m, int32_t *ia, int32_t *ja, int32_t *iau) {
    const int N = m->n_node;
    ia[0] = 0;
    for (int i = 0; i < N; i++) ia[i + 1] = ia[i] + (m->n2n_ia[i + 1] - m->n2n_ia[i]) + 1;
    for (int i = 0; i < N; i++) {
        int idx = ia[i];
        ja[idx] = i;
        for (int j = m->n2n_ia[i]; j < m->n2n_ia[i + 1]; j++) ja[++idx] = m->n2n_ja[j];
    }
    for (int i = 0; i < N; i++) {
        for (int j = ia[i]; j < ia[i + 1]; j++) {
            int mn = ja[j], save = ja[j], ks = j;
            for (int k = j + 1; k < ia[i + 1]; k++)
                if (ja[k] < mn) {
                    mn = ja[k];
                    ks = k;
                }
            ja[j] = mn;
            ja[ks] = save;
            if (ja[j] == i) iau[i] = j;
        }
    }
}

/* ConservativeEulerFluxJacobian, Euler3D/src/CompressibleUtils.cpp:558-609 */
static void flux_jac(double g, const double *Qc, const double *n3, double J[5][5]) {
    const double nx = n3[0], ny = n3[1], nz = n3[2];
    double rho = Qc[0], u = Qc[1] / rho, v = Qc[2] / rho, w = Qc[3] / rho, rhoet = Qc[4];
    double ek = 0.5 * (u * u + v * v + w * w);
    double p = (g - 1.0) * (rhoet - rho * ek);
    double ht = (rhoet + p) / rho;
    double ubar = u * nx + v * ny + w * nz;
    J[0][0] = 0.0;
    J[1][0] = (g - 1.0) * ek * nx - u * ubar;
    J[2][0] = (g - 1.0) * ek * ny - v * ubar;
    J[3][0] = (g - 1.0) * ek * nz - w * ubar;
    J[4][0] = ((g - 1.0) * ek - ht) * ubar;
    J[0][1] = nx;
    J[1][1] = ubar - (g - 2.0) * u * nx;
    J[2][1] = v * nx - (g - 1.0) * u * ny;
    J[3][1] = w * nx - (g - 1.0) * u * nz;
    J[4][1] = ht * nx - (g - 1.0) * u * ubar;
    J[0][2] = ny;
    J[1][2] = u * ny - (g - 1.0) * v * nx;
    J[2][2] = ubar - (g - 2.0) * v * ny;
    J[3][2] = w * ny - (g - 1.0) * v * nz;
    J[4][2] = ht * ny - (g - 1.0) * v * ubar;
    J[0][3] = nz;
    J[1][3] = u * nz - (g - 1.0) * w * nx;
    J[2][3] = v * nz - (g - 1.0) * w * ny;
    J[3][3] = ubar - (g - 2.0) * w * nz;
    J[4][3] = ht * nz - (g - 1.0) * w * ubar;
    J[0][4] = 0.0;
    J[1][4] = (g - 1.0) * nx;
    J[2][4] = (g - 1.0) * ny;
    J[3][4] = (g - 1.0) * nz;
    J[4][4] = g * ubar;
}

/* MC_Matrix_Mul_Matrix, MATH/src/MC_Iterative_Jacobi.c:192-207 */
static void mm5(double A[5][5], double B[5][5], double C[5][5]) {
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            C[i][j] = 0.0;
            for (int k = 0; k < 5; k++) C[i][j] += A[i][k] * B[k][j];
        }
}

/* Compute_RoeAJacobian, Euler3D/src/Roe_Fluxes.cpp:149-273 */
static void roe_ajac(const orc_params *P, const double *QL, const double *QR, const double *n3, double A[5][5]) {
    const double g = P->gamma, nx = n3[0], ny = n3[1], nz = n3[2];
    eos_t L, R;
    eos_face(g, QL, nx, ny, nz, &L);
    eos_face(g, QR, nx, ny, nz, &R);
    double rho = sqrt(R.rho * L.rho);
    double sigma = rho / (L.rho + rho);
    double u = L.u + sigma * (R.u - L.u);
    double v = L.v + sigma * (R.v - L.v);
    double w = L.w + sigma * (R.w - L.w);
    double ht = L.ht + sigma * (R.ht - L.ht);
    double c = roe_c(g, u, v, w, ht);
    double ubar = u * nx + v * ny + w * nz;
    double phi = 0.5 * (g - 1.0) * (u * u + v * v + w * w);
    double M[5][5] = {{1.0, 0.0, 0.0, 0.0, 0.0},
                      {u, rho, 0.0, 0.0, 0.0},
                      {v, 0.0, rho, 0.0, 0.0},
                      {w, 0.0, 0.0, rho, 0.0},
                      {phi / (g - 1.0), rho * u, rho * v, rho * w, 1.0 / (g - 1.0)}};
    double Mi[5][5] = {{1.0, 0.0, 0.0, 0.0, 0.0},
                       {-u / rho, 1.0 / rho, 0.0, 0.0, 0.0},
                       {-v / rho, 0.0, 1.0 / rho, 0.0, 0.0},
                       {-w / rho, 0.0, 0.0, 1.0 / rho, 0.0},
                       {phi, -u * (g - 1.0), -v * (g - 1.0), -w * (g - 1.0), (g - 1.0)}};
    double Pm[5][5] = {{nx, ny, nz, rho / c, rho / c},
                       {0.0, -nz, ny, nx, -nx},
                       {nz, 0.0, -nx, ny, -ny},
                       {-ny, nx, 0.0, nz, -nz},
                       {0.0, 0.0, 0.0, rho * c, rho * c}};
    double Pi[5][5] = {{nx, 0.0, nz, -ny, -nx / (c * c)},
                       {ny, -nz, 0.0, nx, -ny / (c * c)},
                       {nz, ny, -nx, 0.0, -nz / (c * c)},
                       {0.0, 0.5 * nx, 0.5 * ny, 0.5 * nz, 1.0 / (2.0 * rho * c)},
                       {0.0, -0.5 * nx, -0.5 * ny, -0.5 * nz, 1.0 / (2.0 * rho * c)}};
    double T[5][5], Ti[5][5], Ei[5][5], Eg[5][5];
    mm5(M, Pm, T);
    double E[3];
    entropy_fix(P->entropy_fix, L.ubar, L.c, R.ubar, R.c, ubar, c, E);
    memset(Eg, 0, sizeof(Eg));
    Eg[0][0] = Eg[1][1] = Eg[2][2] = E[0];
    Eg[3][3] = E[1];
    Eg[4][4] = E[2];
    mm5(Pi, Mi, Ti);
    mm5(Eg, Ti, Ei);
    mm5(T, Ei, A);
}

/* Compute_Jacobian_Approximate_Roe, Euler3D/src/Roe_Jacobian.cpp:29-228 */
void orc_jacobian(const orc_mesh *m, const orc_params *p, const double *q, const double *dt,
                  const double *res_conv, const double *res_diss, int rhs_mode, const int32_t *ia,
                  const int32_t *ja, const int32_t *iau, double *a, double *b) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    const double g = p->gamma;
    memset(a, 0, sizeof(double) * 25 * (size_t)ia[N]);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 5; k++) {
            double r = res_conv[(size_t)k * N + i];
            if (rhs_mode == 1 && res_diss) r = r + res_diss[(size_t)k * N + i];
            b[5 * (size_t)i + k] = -r;
        }
        for (int j = 0; j < 5; j++) a[25 * (size_t)iau[i] + 5 * j + j] = m->vol[i] / dt[i];
    }
    for (int pass = 0; pass < 2; pass++) {
        const int ne = pass ? m->n_bedge : m->n_edge;
        const int32_t *en = pass ? m->bedge : m->edge;
        const double *es = pass ? m->bedge_area : m->edge_area;
        for (int e = 0; e < ne; e++) {
            const int nl = en[2 * e], nr = en[2 * e + 1];
            double n3[3];
            double area = area_normal(es + 3 * (size_t)e, n3);
            double QL[5], QR[5];
            for (int k = 0; k < 5; k++) {
                QL[k] = q[(size_t)k * NT + nl];
                QR[k] = q[(size_t)k * NT + nr];
            }
            double dL[5][5], dR[5][5], AR[5][5];
            flux_jac(g, QL, n3, dL);
            if (!pass) flux_jac(g, QR, n3, dR);
            roe_ajac(p, QL, QR, n3, AR);
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) {
                    dL[i][j] = 0.5 * (dL[i][j] + AR[i][j]) * area;
                    if (!pass) dR[i][j] = 0.5 * (dR[i][j] - AR[i][j]) * area;
                }
            double *Ad = a + 25 * (size_t)iau[nl];
            if (pass) {
                for (int i = 0; i < 5; i++)
                    for (int j = 0; j < 5; j++) Ad[5 * i + j] += dL[i][j];
                continue;
            }
            int ofL = -1, ofR = -1;
            for (int k = ia[nl]; k < ia[nl + 1]; k++)
                if (ja[k] == nr) {
                    ofL = k;
                    break;
                }
            for (int k = ia[nr]; k < ia[nr + 1]; k++)
                if (ja[k] == nl) {
                    ofR = k;
                    break;
                }
            double *AdR = a + 25 * (size_t)iau[nr], *AoL = a + 25 * (size_t)ofL, *AoR = a + 25 * (size_t)ofR;
            for (int i = 0; i < 5; i++)
                for (int j = 0; j < 5; j++) {
                    Ad[5 * i + j] += dL[i][j];
                    AdR[5 * i + j] -= dR[i][j];
                    AoL[5 * i + j] += dR[i][j];
                    AoR[5 * i + j] -= dL[i][j];
                }
        }
    }
}

/* Compute_RoeFlux, Euler3D/src/Roe_Fluxes.cpp:552-905: Roe branch, conservative
 * variables, AddTime FALSE.  The dissipation is the matrix form
 * -1/2 (M P |Lambda| P^-1 M^-1)(Q_R - Q_L)|S|, with the matrix built exactly as
 * Compute_RoeAJacobian builds it (same products in the same order). */
static void roe_flux_matrix(const orc_params *P, const double *s, const double *QL, const double *QR,
                            double *conv, double *diss) {
    const double g = P->gamma;
    double n[3];
    const double area = area_normal(s, n);
    eos_t L, R;
    eos_face(g, QL, n[0], n[1], n[2], &L);
    eos_face(g, QR, n[0], n[1], n[2], &R);
    double fL[5], fR[5];
    fL[0] = L.rho * L.ubar;
    fL[1] = L.u * fL[0] + L.p * n[0];
    fL[2] = L.v * fL[0] + L.p * n[1];
    fL[3] = L.w * fL[0] + L.p * n[2];
    fL[4] = L.ht * fL[0];
    fR[0] = R.rho * R.ubar;
    fR[1] = R.u * fR[0] + R.p * n[0];
    fR[2] = R.v * fR[0] + R.p * n[1];
    fR[3] = R.w * fR[0] + R.p * n[2];
    fR[4] = R.ht * fR[0];
    double A[5][5];
    roe_ajac(P, QL, QR, n, A);
    double dQ[5];
    for (int i = 0; i < 5; i++) dQ[i] = QR[i] - QL[i];
    for (int i = 0; i < 5; i++) {
        double fa = 0.0; /* MC_Matrix_Mul_Vector */
        for (int j = 0; j < 5; j++) fa += A[i][j] * dQ[j];
        conv[i] = 0.5 * (fL[i] + fR[i]) * area;
        diss[i] = -0.5 * fa * area;
    }
}

double orc_edge_flux_matrix(const orc_params *p, const double *s, const double *ql, const double *qr, double *conv,
                            double *diss) {
    roe_flux_matrix(p, s, ql, qr, conv, diss);
    return 0.0;
}

/* The flux a finite-difference column differentiates for edge (nl, nr), from
 * the states Compute_RoeFlux takes (second-order reconstruction between
 * physical nodes, node values otherwise): the matrix-form Roe flux
 * (Compute_Jacobian_FiniteDifference_Roe), or the scheme's own flux for HLLC
 * and AUSM+ (the scheme-agnostic Compute_Jacobian_FiniteDifference,
 * MPFLOW3D/src/FiniteDifference_Jacobian.cpp:112-667, with this build's
 * residual fluxes), whose dissipative part is zero. */
static void fd_flux(const orc_mesh *m, const orc_params *P, const double *q, const double *grad, const double *phi,
                    int nl, int nr, const double *s, double *conv, double *diss) {
    const int N = m->n_node, NT = N + m->n_bedge;
    double QL[5], QR[5];
    if (nr < N && P->order == 2) {
        reconstruct(m, P, q, grad, phi, nl, nr, QL, QR);
    } else {
        for (int k = 0; k < 5; k++) {
            QL[k] = q[(size_t)k * NT + nl];
            QR[k] = q[(size_t)k * NT + nr];
        }
    }
    if (P->flux == ORC_ROE) {
        roe_flux_matrix(P, s, QL, QR, conv, diss);
        return;
    }
    if (P->flux == ORC_HLLC)
        hllc_flux(P, s, QL, QR, conv);
    else
        ausmp_flux(P, s, QL, QR, conv);
    for (int k = 0; k < 5; k++) diss[k] = 0.0;
}

/* Compute_Jacobian_FiniteDifference_Roe, Euler3D/src/Roe_Jacobian.cpp:236-776.
 * Each node value is perturbed in place (a working copy of q) and restored,
 * as the reference perturbs Q1..Q5.  One-sided methods evaluate the
 * unperturbed flux with the first component and reuse it for the others;
 * boundary edges are always central, with the characteristic ghost state
 * recomputed from the perturbed node (Compute_Characteristic_BoundaryCondition,
 * Euler3D/src/BC.cpp:33-380, the same ghost as bc_edge for gauge pressure 0)
 * and restored after each flux.  Returns the number of boundary evaluations
 * with a negative ghost pressure (reference: error()). */
int orc_jacobian_fd(const orc_mesh *m, const orc_params *p, const double *q0, const double *grad,
                    const double *phi, const double *dt, const double *res_conv, const double *res_diss,
                    int rhs_mode, int method, int iter, int add_time, const int32_t *ia, const int32_t *ja,
                    const int32_t *iau, double *a, double *b) {
    const int N = m->n_node, NT = m->n_node + m->n_bedge;
    const double eps = 1.0E-8;
    double *q = (double *)malloc(sizeof(double) * 5 * (size_t)NT);
    memcpy(q, q0, sizeof(double) * 5 * (size_t)NT);
    int bad = 0;
    memset(a, 0, sizeof(double) * 25 * (size_t)ia[N]);
    for (int i = 0; i < N; i++) {
        for (int k = 0; k < 5; k++) {
            double r = res_conv[(size_t)k * N + i];
            if (rhs_mode == 1 && res_diss) r = r + res_diss[(size_t)k * N + i];
            b[5 * (size_t)i + k] = -r;
        }
        if (add_time)
            for (int j = 0; j < 5; j++) a[25 * (size_t)iau[i] + 5 * j + j] = m->vol[i] / dt[i];
    }
    int fdm = method;
    if (fdm == ORC_FD_ALTERNATE) fdm = (iter % 2 == 0) ? ORC_FD_FORWARD : ORC_FD_BACKWARD;
    const double coeff = (fdm == ORC_FD_CENTRAL) ? 0.5 : 1.0;
    const int plus = (fdm == ORC_FD_CENTRAL) || (fdm == ORC_FD_FORWARD);
    const int minus = (fdm == ORC_FD_CENTRAL) || (fdm == ORC_FD_BACKWARD);
    for (int e = 0; e < m->n_edge; e++) {
        const int nl = m->edge[2 * e], nr = m->edge[2 * e + 1];
        const double *s = m->edge_area + 3 * (size_t)e;
        double d[2][5][5];
        for (int side = 0; side < 2; side++) {
            const int node = side ? nr : nl;
            double FPc[5], FPd[5], FMc[5], FMd[5];
            for (int c = 0; c < 5; c++) {
                double *qc = q + (size_t)c * NT + node;
                const double save = *qc;
                if (c == 0) {
                    *qc = plus ? save + eps : save;
                    fd_flux(m, p, q, grad, phi, nl, nr, s, FPc, FPd);
                    *qc = minus ? save - eps : save;
                    fd_flux(m, p, q, grad, phi, nl, nr, s, FMc, FMd);
                } else {
                    if (plus) {
                        *qc = save + eps;
                        fd_flux(m, p, q, grad, phi, nl, nr, s, FPc, FPd);
                    }
                    if (minus) {
                        *qc = save - eps;
                        fd_flux(m, p, q, grad, phi, nl, nr, s, FMc, FMd);
                    }
                }
                for (int i = 0; i < 5; i++) d[side][i][c] = coeff * ((FPc[i] + FPd[i]) - (FMc[i] + FMd[i])) / eps;
                *qc = save;
            }
        }
        int ofL = -1, ofR = -1;
        for (int k = ia[nl]; k < ia[nl + 1]; k++)
            if (ja[k] == nr) {
                ofL = k;
                break;
            }
        for (int k = ia[nr]; k < ia[nr + 1]; k++)
            if (ja[k] == nl) {
                ofR = k;
                break;
            }
        double *AdL = a + 25 * (size_t)iau[nl], *AdR = a + 25 * (size_t)iau[nr];
        double *AoL = a + 25 * (size_t)ofL, *AoR = a + 25 * (size_t)ofR;
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) {
                AdL[5 * i + j] += d[0][i][j];
                AdR[5 * i + j] -= d[1][i][j];
                AoL[5 * i + j] += d[1][i][j];
                AoR[5 * i + j] -= d[0][i][j];
            }
    }
    for (int e = 0; e < m->n_bedge; e++) {
        const int nl = m->bedge[2 * e], ng = m->bedge[2 * e + 1];
        const double *s = m->bedge_area + 3 * (size_t)e;
        double QG[5];
        for (int k = 0; k < 5; k++) QG[k] = q[(size_t)k * NT + ng];
        double d[5][5];
        for (int c = 0; c < 5; c++) {
            double F[2][2][5];
            double *qc = q + (size_t)c * NT + nl;
            const double save = *qc;
            for (int pm = 0; pm < 2; pm++) {
                *qc = pm ? save - eps : save + eps;
                double QL[5], QB[5];
                for (int k = 0; k < 5; k++) {
                    QL[k] = q[(size_t)k * NT + nl];
                    QB[k] = QG[k];
                }
                bad += bc_edge(p, QL, QG, s, m->bedge_type[e], iter, QB);
                for (int k = 0; k < 5; k++) q[(size_t)k * NT + ng] = QB[k];
                fd_flux(m, p, q, grad, phi, nl, ng, s, F[pm][0], F[pm][1]);
                for (int k = 0; k < 5; k++) q[(size_t)k * NT + ng] = QG[k];
            }
            for (int i = 0; i < 5; i++)
                d[i][c] = 0.5 * ((F[0][0][i] + F[0][1][i]) - (F[1][0][i] + F[1][1][i])) / eps;
            *qc = save;
        }
        double *AdL = a + 25 * (size_t)iau[nl];
        for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) AdL[5 * i + j] += d[i][j];
    }
    free(q);
    return bad;
}

/* MC_Iterative_Block_LU_Jacobi_CRS, forward direction,
 * MATH/src/MC_Iterative_Jacobi.c:346-535 */
double orc_block_gs(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
                    const double *b, int sweeps, double *x) {
    double *LU = (double *)malloc(sizeof(double) * 25 * (size_t)n);
    double rms = 0.0, rhs[5], tv[5];
    for (size_t i = 0; i < 5 * (size_t)n; i++) x[i] = 0.0;
    for (int r = 0; r < n; r++) {
        double *L = LU + 25 * (size_t)r;
        memcpy(L, a + 25 * (size_t)iau[r], 25 * sizeof(double));
        for (int k = 0; k < 4; k++) {
            for (int j = k + 1; j < 5; j++) L[5 * j + k] = L[5 * j + k] / L[5 * k + k];
            for (int j = k + 1; j < 5; j++)
                for (int i = k + 1; i < 5; i++) L[5 * i + j] -= L[5 * i + k] * L[5 * k + j];
        }
    }
    if (sweeps <= 0) sweeps = 10000;
    for (int it = 0; it < sweeps; it++) {
        rms = 0.0;
        for (int r = 0; r < n; r++) {
            for (int m = 0; m < 5; m++) rhs[m] = b[5 * (size_t)r + m];
            for (int j = ia[r]; j < ia[r + 1]; j++) {
                const double *A = a + 25 * (size_t)j, *xj = x + 5 * (size_t)ja[j];
                for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
                for (int ii = 0; ii < 5; ii++)
                    for (int jj = 0; jj < 5; jj++) tv[ii] += A[5 * ii + jj] * xj[jj];
                for (int m = 0; m < 5; m++) rhs[m] -= tv[m];
            }
            for (int m = 0; m < 5; m++) rms += rhs[m] * rhs[m];
            const double *D = a + 25 * (size_t)iau[r];
            double *xr = x + 5 * (size_t)r;
            for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
            for (int ii = 0; ii < 5; ii++)
                for (int jj = 0; jj < 5; jj++) tv[ii] += D[5 * ii + jj] * xr[jj];
            for (int m = 0; m < 5; m++) rhs[m] += tv[m];
            const double *L = LU + 25 * (size_t)r;
            for (int ii = 0; ii < 5; ii++) {
                tv[ii] = rhs[ii];
                for (int jj = 0; jj < ii; jj++) tv[ii] -= L[5 * ii + jj] * tv[jj];
            }
            for (int ii = 4; ii >= 0; ii--) {
                xr[ii] = tv[ii];
                for (int jj = 4; jj > ii; jj--) xr[ii] -= L[5 * ii + jj] * xr[jj];
                xr[ii] = xr[ii] / L[5 * ii + ii];
            }
        }
        rms = sqrt(rms / ((double)n));
        if (rms < 100.0 * DBL_EPSILON) break;
    }
    free(LU);
    return rms;
}

/* GaussainLUDecompose and pivoted substitutions, MATH/src/LinearAlgebra.c:20-41,117-155 */
static void glu(double A[5][5], int *P) {
    for (int i = 0; i < 5; i++) P[i] = i;
    for (int c = 0; c < 5; c++) {
        double pm = A[P[c]][c];
        int pi = c;
        for (int i = c + 1; i < 5; i++)
            if (fabs(A[P[i]][c]) > fabs(pm)) {
                pm = A[P[i]][c];
                pi = i;
            }
        if (pi != c) {
            int t = P[c];
            P[c] = P[pi];
            P[pi] = t;
        }
        for (int i = c + 1; i < 5; i++) A[P[i]][c] /= A[P[c]][c];
        for (int i = c + 1; i < 5; i++)
            for (int j = c + 1; j < 5; j++) A[P[i]][j] -= A[P[i]][c] * A[P[c]][j];
    }
}
static void glu_solve(double A[5][5], const int *P, const double *b, double *x) {
    double y[5];
    y[0] = b[P[0]];
    for (int i = 1; i < 5; i++) {
        y[i] = b[P[i]];
        for (int j = 0; j < i; j++) y[i] -= y[j] * ((i == j) ? 1.0 : A[P[i]][j]);
    }
    x[4] = y[4] / A[P[4]][4];
    for (int i = 3; i >= 0; i--) {
        x[i] = y[i];
        for (int j = 4; j > i; j--) x[i] -= x[j] * A[P[i]][j];
        x[i] /= A[P[i]][i];
    }
}

/* MC_Iterative_Block_LUSGS_CRS, MATH/src/MC_Iterative_Jacobi.c:540-717 (X only) */
void orc_lusgs(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
               const double *b, int sweeps, double *x) {
    for (int r = 0; r < n; r++) {
        double s = 0.0;
        for (int j = 0; j < 5; j++) s += b[5 * (size_t)r + j];
        s /= 5;
        for (int j = 0; j < 5; j++) x[5 * (size_t)r + j] = s;
    }
    if (sweeps <= 0) sweeps = 100;
    for (int it = 0; it < sweeps; it++) {
        for (int dir = 0; dir < 2; dir++) {
            for (int rr = 0; rr < n; rr++) {
                const int r = dir ? n - 1 - rr : rr;
                double rhs[5], tv[5];
                for (int j = 0; j < 5; j++) rhs[j] = b[5 * (size_t)r + j];
                for (int j = ia[r]; j < ia[r + 1]; j++) {
                    if (j == iau[r]) continue;
                    const double *A = a + 25 * (size_t)j, *xj = x + 5 * (size_t)ja[j];
                    for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
                    for (int ii = 0; ii < 5; ii++)
                        for (int jj = 0; jj < 5; jj++) tv[ii] += A[5 * ii + jj] * xj[jj];
                    for (int m = 0; m < 5; m++) rhs[m] -= tv[m];
                }
                double D[5][5];
                int P[5];
                for (int i = 0; i < 5; i++)
                    for (int j = 0; j < 5; j++) D[i][j] = a[25 * (size_t)iau[r] + 5 * i + j];
                glu(D, P);
                double sol[5];
                glu_solve(D, P, rhs, sol);
                for (int j = 0; j < 5; j++) x[5 * (size_t)r + j] = sol[j];
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Block ILU(0), block-Jacobi and FGMRES (no reference arithmetic exists;
 * Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 10.4
 * (ILU(0)), Alg. 9.6 (FGMRES) with classical Gram-Schmidt applied twice).
 * The GPU implementation performs the same per-row operations in the same
 * order, so factors, triangular solves and SpMV are bit-identical.
 * ------------------------------------------------------------------------- */

/* 5x5 inverse by Gauss-Jordan elimination with partial pivoting.  Returns 0
 * when a zero pivot is met. */
static int inv5(const double *A, double *Ainv) {
    double a[5][10];
    for (int r = 0; r < 5; r++)
        for (int c = 0; c < 5; c++) {
            a[r][c] = A[5 * r + c];
            a[r][c + 5] = (r == c) ? 1.0 : 0.0;
        }
    for (int c = 0; c < 5; c++) {
        int piv = c;
        for (int r = c + 1; r < 5; r++)
            if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 10; j++) {
                double t = a[c][j];
                a[c][j] = a[piv][j];
                a[piv][j] = t;
            }
        double d = a[c][c];
        if (d == 0.0) return 0;
        for (int j = 0; j < 10; j++) a[c][j] = a[c][j] / d;
        for (int r = 0; r < 5; r++) {
            if (r == c) continue;
            double f = a[r][c];
            for (int j = 0; j < 10; j++) a[r][j] = a[r][j] - f * a[c][j];
        }
    }
    for (int r = 0; r < 5; r++)
        for (int c = 0; c < 5; c++) Ainv[5 * r + c] = a[r][c + 5];
    return 1;
}

/* C = A*B for 5x5 row-major blocks, sum over t in ascending order */
static void bmul(const double *A, const double *B, double *C) {
    for (int r = 0; r < 5; r++)
        for (int c = 0; c < 5; c++) {
            double s = 0.0;
            for (int t = 0; t < 5; t++) s += A[5 * r + t] * B[5 * t + c];
            C[5 * r + c] = s;
        }
}

int orc_ilu0(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, double *lu, double *dinv) {
    double tmp[25], prod[25];
    for (int i = 0; i < n; i++) {
        for (int kk = ia[i]; kk < iau[i]; kk++) {
            const int k = ja[kk];
            bmul(lu + 25 * (size_t)kk, dinv + 25 * (size_t)k, tmp);
            memcpy(lu + 25 * (size_t)kk, tmp, sizeof(tmp));
            int pp = iau[k] + 1, qq = kk + 1;
            while (pp < ia[k + 1] && qq < ia[i + 1]) {
                if (ja[pp] == ja[qq]) {
                    bmul(lu + 25 * (size_t)kk, lu + 25 * (size_t)pp, prod);
                    double *T = lu + 25 * (size_t)qq;
                    for (int t = 0; t < 25; t++) T[t] = T[t] - prod[t];
                    pp++;
                    qq++;
                } else if (ja[pp] < ja[qq])
                    pp++;
                else
                    qq++;
            }
        }
        if (!inv5(lu + 25 * (size_t)iau[i], dinv + 25 * (size_t)i)) return -1 - i;
    }
    return 0;
}

static void bmv_sub(const double *A, const double *x, double *y) {
    for (int r = 0; r < 5; r++) {
        double s = 0.0;
        for (int t = 0; t < 5; t++) s += A[5 * r + t] * x[t];
        y[r] = y[r] - s;
    }
}

void orc_ilu0_apply(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *lu,
                    const double *dinv, const double *r, double *z) {
    for (int i = 0; i < n; i++) {
        double y[5];
        for (int t = 0; t < 5; t++) y[t] = r[5 * (size_t)i + t];
        for (int kk = ia[i]; kk < iau[i]; kk++) bmv_sub(lu + 25 * (size_t)kk, z + 5 * (size_t)ja[kk], y);
        for (int t = 0; t < 5; t++) z[5 * (size_t)i + t] = y[t];
    }
    for (int i = n - 1; i >= 0; i--) {
        double y[5];
        for (int t = 0; t < 5; t++) y[t] = z[5 * (size_t)i + t];
        for (int kk = iau[i] + 1; kk < ia[i + 1]; kk++) bmv_sub(lu + 25 * (size_t)kk, z + 5 * (size_t)ja[kk], y);
        const double *D = dinv + 25 * (size_t)i;
        for (int rr = 0; rr < 5; rr++) {
            double s = 0.0;
            for (int t = 0; t < 5; t++) s += D[5 * rr + t] * y[t];
            z[5 * (size_t)i + rr] = s;
        }
    }
}

void orc_bsr_mv(int32_t n, const int32_t *ia, const int32_t *ja, const double *a, const double *x, double *y) {
    for (int i = 0; i < n; i++)
        for (int r = 0; r < 5; r++) {
            double s = 0.0;
            for (int kk = ia[i]; kk < ia[i + 1]; kk++) {
                const double *A = a + 25 * (size_t)kk + 5 * r, *xj = x + 5 * (size_t)ja[kk];
                for (int c = 0; c < 5; c++) s += A[c] * xj[c];
            }
            y[5 * (size_t)i + r] = s;
        }
}

static double vdot(size_t n, const double *a, const double *b) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int orc_fgmres(int32_t n, const int32_t *ia, const int32_t *ja, const int32_t *iau, const double *a,
               const double *b, int m, int max_restarts, double rtol, int precond, double *x,
               double *final_res) {
    const size_t L = 5 * (size_t)n;
    double *V = (double *)calloc(L * (m + 1), sizeof(double));
    double *Z = (double *)calloc(L * m, sizeof(double));
    double *w = (double *)calloc(L, sizeof(double));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double *)calloc(m, sizeof(double)), *sn = (double *)calloc(m, sizeof(double));
    double *gv = (double *)calloc(m + 1, sizeof(double)), *hcol = (double *)calloc(m + 1, sizeof(double));
    double *y = (double *)calloc(m, sizeof(double));
    double *lu = NULL, *dinv = NULL;
    const size_t nnz = (size_t)ia[n];
    if (precond == 2) {
        lu = (double *)malloc(sizeof(double) * 25 * nnz);
        memcpy(lu, a, sizeof(double) * 25 * nnz);
    }
    if (precond) dinv = (double *)malloc(sizeof(double) * 25 * (size_t)n);
    if (precond == 2) {
        if (orc_ilu0(n, ia, ja, iau, lu, dinv) != 0) precond = 0;
    } else if (precond == 1) {
        for (int i = 0; i < n; i++) inv5(a + 25 * (size_t)iau[i], dinv + 25 * (size_t)i);
    }
    for (size_t i = 0; i < L; i++) x[i] = 0.0;
    double bnorm = sqrt(vdot(L, b, b));
    double tol = rtol * bnorm;
    int total = 0;
    double res = bnorm;
    /* r0 = b (x0 = 0) */
    memcpy(V, b, L * sizeof(double));
    double beta = bnorm;
    for (int cycle = 0; cycle <= max_restarts && beta > tol && beta > 0.0; cycle++) {
        for (size_t i = 0; i < L; i++) V[i] = V[i] / beta;
        for (int i = 0; i <= m; i++) gv[i] = 0.0;
        gv[0] = beta;
        int j;
        int kdim = 0;
        for (j = 0; j < m; j++) {
            double *vj = V + L * j, *zj = Z + L * j;
            if (precond == 2)
                orc_ilu0_apply(n, ia, ja, iau, lu, dinv, vj, zj);
            else if (precond == 1) {
                for (int i = 0; i < n; i++)
                    for (int r = 0; r < 5; r++) {
                        double s = 0.0;
                        for (int t = 0; t < 5; t++) s += dinv[25 * (size_t)i + 5 * r + t] * vj[5 * (size_t)i + t];
                        zj[5 * (size_t)i + r] = s;
                    }
            } else
                memcpy(zj, vj, L * sizeof(double));
            orc_bsr_mv(n, ia, ja, a, zj, w);
            /* CGS2 */
            for (int k = 0; k <= j; k++) hcol[k] = 0.0;
            for (int pass = 0; pass < 2; pass++) {
                double h[64 + 1];
                for (int k = 0; k <= j; k++) h[k] = vdot(L, V + L * k, w);
                for (int k = 0; k <= j; k++) {
                    const double *vk = V + L * k;
                    for (size_t i = 0; i < L; i++) w[i] = w[i] - h[k] * vk[i];
                    hcol[k] += h[k];
                }
            }
            double hn = sqrt(vdot(L, w, w));
            hcol[j + 1] = hn;
            for (int k = 0; k < j; k++) {
                double t = cs[k] * hcol[k] + sn[k] * hcol[k + 1];
                hcol[k + 1] = -sn[k] * hcol[k] + cs[k] * hcol[k + 1];
                hcol[k] = t;
            }
            double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
            cs[j] = hcol[j] / den;
            sn[j] = hcol[j + 1] / den;
            hcol[j] = den;
            hcol[j + 1] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            for (int k = 0; k <= j; k++) H[(size_t)k * m + j] = hcol[k];
            total++;
            kdim = j + 1;
            res = fabs(gv[j + 1]);
            if (hn != 0.0) {
                double *vn = V + L * (j + 1);
                for (size_t i = 0; i < L; i++) vn[i] = w[i] / hn;
            }
            if (res <= tol || hn == 0.0) break;
        }
        for (int k = kdim - 1; k >= 0; k--) {
            double s = gv[k];
            for (int t = k + 1; t < kdim; t++) s -= H[(size_t)k * m + t] * y[t];
            y[k] = s / H[(size_t)k * m + k];
        }
        for (int k = 0; k < kdim; k++) {
            const double *zk = Z + L * k;
            for (size_t i = 0; i < L; i++) x[i] = x[i] + y[k] * zk[i];
        }
        if (res <= tol) break;
        /* true residual for the restart */
        orc_bsr_mv(n, ia, ja, a, x, w);
        for (size_t i = 0; i < L; i++) V[i] = b[i] - w[i];
        beta = sqrt(vdot(L, V, V));
        res = beta;
    }
    if (final_res) *final_res = res;
    free(V);
    free(Z);
    free(w);
    free(H);
    free(cs);
    free(sn);
    free(gv);
    free(hcol);
    free(y);
    free(lu);
    free(dinv);
    return total;
}

void orc_forces(const orc_mesh *m, const orc_params *p, const double *q, const int32_t *bedge_tag, int wall_tag,
                double alpha_deg, double *force, double *lift, double *drag, double *pint) {
    const int NT = m->n_node + m->n_bedge;
    double F[3] = {0.0, 0.0, 0.0}, pi_ = 0.0;
    for (int e = 0; e < m->n_bedge; e++) {
        if (wall_tag > 0 ? bedge_tag[e] != wall_tag : m->bedge_type[e] != ORC_BC_EULER_WALL) continue;
        const int nl = m->bedge[2 * e];
        double Q[5];
        for (int k = 0; k < 5; k++) Q[k] = q[(size_t)k * NT + nl];
        eos_t s;
        eos_face(p->gamma, Q, 0.0, 0.0, 0.0, &s);
        const double *S = m->bedge_area + 3 * (size_t)e;
        for (int d = 0; d < 3; d++) F[d] += s.p * S[d];
        pi_ += s.p * sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    }
    const double a = alpha_deg * M_PI / 180.0;
    for (int d = 0; d < 3; d++) force[d] = F[d];
    *drag = F[0] * cos(a) + F[1] * sin(a);
    *lift = -F[0] * sin(a) + F[1] * cos(a);
    *pint = pi_;
}

int orc_solve_steady(const orc_mesh *m, const orc_params *p0, const orc_solve_opts *o, double alpha_deg,
                     int wall_tag, const int32_t *bedge_tag, double *q, double *rms_hist, int32_t *krylov_hist,
                     orc_solve_result *res) {
    const int N = m->n_node;
    const int32_t nnz = orc_bsr_nnz(m);
    int32_t *ia = (int32_t *)malloc(sizeof(int32_t) * (N + 1)), *ja = (int32_t *)malloc(sizeof(int32_t) * nnz),
            *iau = (int32_t *)malloc(sizeof(int32_t) * N);
    orc_bsr_structure(m, ia, ja, iau);
    double *grad = (double *)calloc(15 * (size_t)N, sizeof(double));
    double *phi = (double *)calloc(5 * (size_t)N, sizeof(double));
    double *rc = (double *)malloc(sizeof(double) * 5 * N), *rd = (double *)malloc(sizeof(double) * 5 * N);
    double *dta = (double *)malloc(sizeof(double) * N), *dt = (double *)malloc(sizeof(double) * N);
    double *A = (double *)malloc(sizeof(double) * 25 * (size_t)nnz);
    double *B = (double *)malloc(sizeof(double) * 5 * N), *X = (double *)malloc(sizeof(double) * 5 * N);
    const int NT = N + m->n_bedge;
    int status = 0;
    memset(res, 0, sizeof(*res));
    for (int it = 0; it < o->max_iter; it++) {
        orc_params p = *p0;
        if (orc_bc(m, &p, it, q)) {
            status = -2;
            break;
        }
        if (p.order == 2 && it < o->first_order_iters) p.order = 1;
        if (p.order == 2) {
            if (o->gradient == ORC_GRAD_GG)
                orc_gg_gradient(m, q, grad);
            else
                orc_lsq_gradient(m, q, grad);
            if (p.limiter != ORC_LIM_NONE) orc_limiter(m, &p, q, grad, phi);
        }
        memset(rc, 0, sizeof(double) * 5 * N);
        memset(rd, 0, sizeof(double) * 5 * N);
        memset(dta, 0, sizeof(double) * N);
        orc_residual(m, &p, q, grad, phi, rc, rd, dta, NULL);
        double rk[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < 5; k++)
            for (int i = 0; i < N; i++) {
                double r = rc[(size_t)k * N + i] + rd[(size_t)k * N + i];
                rk[k] += r * r;
            }
        double rms = rk[0] + rk[1] + rk[2] + rk[3] + rk[4];
        rms = sqrt(rms / (5.0 * (double)N));
        if (rms_hist) rms_hist[it] = rms;
        if (it == 0) res->first_rms = rms;
        res->final_rms = rms;
        res->iterations = it;
        if (!(rms == rms) || isinf(rms)) {
            status = -1;
            break;
        }
        if (rms < o->tol) {
            res->converged = 1;
            break;
        }
        orc_delta_t(m, &p, it, dta, dt);
        orc_jacobian(m, &p, q, dt, rc, rd, 1, ia, ja, iau, A, B);
        double fr;
        int kit = orc_fgmres(N, ia, ja, iau, A, B, o->krylov_m, o->max_restarts, o->krylov_rtol, o->precond, X, &fr);
        if (krylov_hist) krylov_hist[it] = kit;
        res->krylov_total += kit;
        for (int k = 0; k < 5; k++)
            for (int i = 0; i < N; i++) q[(size_t)k * NT + i] += o->relaxation * X[5 * (size_t)i + k];
        res->iterations = it + 1;
    }
    orc_forces(m, p0, q, bedge_tag, wall_tag, alpha_deg, res->force, &res->lift, &res->drag, &res->pressure_integral);
    free(ia);
    free(ja);
    free(iau);
    free(grad);
    free(phi);
    free(rc);
    free(rd);
    free(dta);
    free(dt);
    free(A);
    free(B);
    free(X);
    return status;
}
