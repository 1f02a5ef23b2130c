// Implicit-side kernels: approximate Roe Jacobian assembly into BSR, block
// SpMV, block ILU(0) and its triangular solves, block-Jacobi, the reference's
// block Gauss-Seidel and LU-SGS sweeps, and restarted FGMRES.
//
// Row-sequential algorithms (ILU(0), Gauss-Seidel, LU-SGS, triangular solves)
// run level by level over the dependency DAG of the BSR pattern: rows of one
// level do not depend on each other and every row performs the same
// operations in the same order as the sequential CPU code, so factors and
// sweeps are bit-identical to it.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "cfd_internal.h"

#include <type_traits>
#include "flux.h"

using namespace cfd;

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Jacobian blocks
// ---------------------------------------------------------------------------
// ConservativeEulerFluxJacobian, Euler3D/src/CompressibleUtils.cpp:558-609
__device__ __forceinline__ void flux_jac(double g, const double *Qc, double nx, double ny, double nz, double *J) {
    const double rho = Qc[0], u = Qc[1] / rho, v = Qc[2] / rho, w = Qc[3] / rho, rhoet = Qc[4];
    const double ek = 0.5 * (u * u + v * v + w * w);
    const double p = (g - 1.0) * (rhoet - rho * ek);
    const double ht = (rhoet + p) / rho;
    const double ubar = u * nx + v * ny + w * nz;
    J[0] = 0.0;
    J[5] = (g - 1.0) * ek * nx - u * ubar;
    J[10] = (g - 1.0) * ek * ny - v * ubar;
    J[15] = (g - 1.0) * ek * nz - w * ubar;
    J[20] = ((g - 1.0) * ek - ht) * ubar;
    J[1] = nx;
    J[6] = ubar - (g - 2.0) * u * nx;
    J[11] = v * nx - (g - 1.0) * u * ny;
    J[16] = w * nx - (g - 1.0) * u * nz;
    J[21] = ht * nx - (g - 1.0) * u * ubar;
    J[2] = ny;
    J[7] = u * ny - (g - 1.0) * v * nx;
    J[12] = ubar - (g - 2.0) * v * ny;
    J[17] = w * ny - (g - 1.0) * v * nz;
    J[22] = ht * ny - (g - 1.0) * v * ubar;
    J[3] = nz;
    J[8] = u * nz - (g - 1.0) * w * nx;
    J[13] = v * nz - (g - 1.0) * w * ny;
    J[18] = ubar - (g - 2.0) * w * nz;
    J[23] = ht * nz - (g - 1.0) * w * ubar;
    J[4] = 0.0;
    J[9] = (g - 1.0) * nx;
    J[14] = (g - 1.0) * ny;
    J[19] = (g - 1.0) * nz;
    J[24] = g * ubar;
}

// MC_Matrix_Mul_Matrix, MATH/src/MC_Iterative_Jacobi.c:192-207
__device__ __forceinline__ void mm5(const double *A, const double *B, double *C) {
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 5; k++) s += A[5 * i + k] * B[5 * k + j];
            C[5 * i + j] = s;
        }
}

// Compute_RoeAJacobian, Euler3D/src/Roe_Fluxes.cpp:149-273: A = M P |Lambda| P^-1 M^-1
__device__ __forceinline__ void roe_ajac(double g, int efix, const double *QL, const double *QR, double nx, double ny,
                                         double nz, double *A) {
    Eos<double> L, R;
    eos_face(g, QL, nx, ny, nz, L);
    eos_face(g, QR, nx, ny, nz, R);
    const double rho = sqrt(R.rho * L.rho);
    const double sigma = rho / (L.rho + rho);
    const double u = L.u + sigma * (R.u - L.u);
    const double v = L.v + sigma * (R.v - L.v);
    const double w = L.w + sigma * (R.w - L.w);
    const double ht = L.ht + sigma * (R.ht - L.ht);
    const double c = roe_speed_sound(g, u, v, w, ht);
    const double ubar = u * nx + v * ny + w * nz;
    const double phi = 0.5 * (g - 1.0) * (u * u + v * v + w * w);
    const double M[25] = {1.0, 0.0, 0.0, 0.0, 0.0, u, rho, 0.0, 0.0, 0.0, v, 0.0, rho, 0.0, 0.0, w, 0.0, 0.0, rho, 0.0,
                          phi / (g - 1.0), rho * u, rho * v, rho * w, 1.0 / (g - 1.0)};
    const double Mi[25] = {1.0, 0.0, 0.0, 0.0, 0.0,
                           -u / rho, 1.0 / rho, 0.0, 0.0, 0.0,
                           -v / rho, 0.0, 1.0 / rho, 0.0, 0.0,
                           -w / rho, 0.0, 0.0, 1.0 / rho, 0.0,
                           phi, -u * (g - 1.0), -v * (g - 1.0), -w * (g - 1.0), (g - 1.0)};
    const double P[25] = {nx, ny, nz, rho / c, rho / c, 0.0, -nz, ny, nx, -nx, nz, 0.0, -nx, ny, -ny,
                          -ny, nx, 0.0, nz, -nz, 0.0, 0.0, 0.0, rho * c, rho * c};
    const double Pi[25] = {nx, 0.0, nz, -ny, -nx / (c * c),
                           ny, -nz, 0.0, nx, -ny / (c * c),
                           nz, ny, -nx, 0.0, -nz / (c * c),
                           0.0, 0.5 * nx, 0.5 * ny, 0.5 * nz, 1.0 / (2.0 * rho * c),
                           0.0, -0.5 * nx, -0.5 * ny, -0.5 * nz, 1.0 / (2.0 * rho * c)};
    double E0, E1, E2;
    roe_eigen(efix, L.ubar, L.c, R.ubar, R.c, ubar, c, E0, E1, E2);
    double T[25], Ti[25];
    mm5(M, P, T);
    mm5(Pi, Mi, Ti);
    // |Lambda| * Ti: the diagonal product of the reference's full multiply
    const double Ed[5] = {E0, E0, E0, E1, E2};
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) Ti[5 * i + j] = Ed[i] * Ti[5 * i + j];
    mm5(T, Ti, A);
}

template <class TA>
struct JacArgs {
    int N, NT;
    const int *order;
    const int *ip, *sptr, *nbr, *soff, *iau;
    const double *isx, *isy, *isz, *qa, *vol, *dt, *res;
    TA *a;
    double *b;
    double g;
    int efix, rhs_mode;
};

// Compute_Jacobian_Approximate_Roe, Euler3D/src/Roe_Jacobian.cpp:29-228,
// assembled by rows: node i writes its diagonal block (accumulated in edge
// order) and the off-diagonal block of every neighbour.
template <class TA>
__global__ __launch_bounds__(128) void k_jacobian(JacArgs<TA> a) {
    const int i = xcd_block() * 128 + threadIdx.x;  // internal node
    if (i >= a.N) return;
    const int ai = a.order[i];                      // its row of the BSR system
    double qi[5];
#pragma unroll
    for (int k = 0; k < 5; k++) qi[k] = a.qa[kQA * (size_t)i + k];
    double D[25];
#pragma unroll
    for (int t = 0; t < 25; t++) D[t] = 0.0;
    const double vd = a.vol[i] / a.dt[i];
#pragma unroll
    for (int t = 0; t < 5; t++) D[6 * t] = vd;
    const int deg = a.ip[i + 1] - a.ip[i];
    size_t idx = (size_t)a.sptr[i / kSlice] + (i % kSlice);
    for (int k = 0; k < deg; k++, idx += kSlice) {
        const int jr = ld_stream(a.nbr + idx);
        const bool iL = jr >= 0;  // ghosts (jr >= N) always pair with a physical L
        const int j = iL ? jr : ~jr;
        double nx, ny, nz;
        const double area = area_normal(ld_stream(a.isx + idx), ld_stream(a.isy + idx), ld_stream(a.isz + idx), nx,
                                        ny, nz);
        const double *rj = a.qa + kQA * (size_t)j;
        double QL[5], QR[5];
#pragma unroll
        for (int t = 0; t < 5; t++) {
            QL[t] = iL ? qi[t] : rj[t];
            QR[t] = iL ? rj[t] : qi[t];
        }
        double AR[25], J[25];
        roe_ajac(a.g, a.efix, QL, QR, nx, ny, nz, AR);
        if (j >= a.N) {
            flux_jac(a.g, QL, nx, ny, nz, J);
#pragma unroll
            for (int t = 0; t < 25; t++) D[t] += 0.5 * (J[t] + AR[t]) * area;
            continue;
        }
        TA *O = a.a + 25 * (size_t)ld_stream(a.soff + idx);
        if (iL) {
            flux_jac(a.g, QL, nx, ny, nz, J);
#pragma unroll
            for (int t = 0; t < 25; t++) D[t] += 0.5 * (J[t] + AR[t]) * area;
            flux_jac(a.g, QR, nx, ny, nz, J);
#pragma unroll
            for (int t = 0; t < 25; t++) O[t] = (TA)(0.0 + 0.5 * (J[t] - AR[t]) * area);
        } else {
            flux_jac(a.g, QR, nx, ny, nz, J);
#pragma unroll
            for (int t = 0; t < 25; t++) D[t] -= 0.5 * (J[t] - AR[t]) * area;
            flux_jac(a.g, QL, nx, ny, nz, J);
#pragma unroll
            for (int t = 0; t < 25; t++) O[t] = (TA)(0.0 - 0.5 * (J[t] + AR[t]) * area);
        }
    }
    TA *Dd = a.a + 25 * (size_t)a.iau[ai];
#pragma unroll
    for (int t = 0; t < 25; t++) Dd[t] = (TA)D[t];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        double r = a.res[kRES * (size_t)i + k];
        if (a.rhs_mode == 1) r = r + a.res[kRES * (size_t)i + 5 + k];
        a.b[5 * (size_t)ai + k] = -r;
    }
}

// Compute_RoeFlux, Euler3D/src/Roe_Fluxes.cpp:552-905 (Roe branch,
// conservative variables, no time term), returned as Conv + Diss per
// component: Conv = (F_L + F_R)|S|/2, Diss = -(M P |Lambda| P^-1 M^-1)(Q_R - Q_L)|S|/2.
__device__ __forceinline__ void roe_flux_matrix(double g, int efix, const double *QL, const double *QR, double sx,
                                                double sy, double sz, double *F) {
    double nx, ny, nz;
    const double area = area_normal(sx, sy, sz, nx, ny, nz);
    Eos<double> L, R;
    eos_face(g, QL, nx, ny, nz, L);
    eos_face(g, QR, nx, ny, nz, R);
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
    double A[25];
    roe_ajac(g, efix, QL, QR, nx, ny, nz, A);
    double dQ[5];
#pragma unroll
    for (int t = 0; t < 5; t++) dQ[t] = QR[t] - QL[t];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        double fa = 0.0;  // MC_Matrix_Mul_Vector
#pragma unroll
        for (int t = 0; t < 5; t++) fa += A[5 * r + t] * dQ[t];
        const double conv = 0.5 * (fL[r] + fR[r]) * area;
        const double diss = -0.5 * fa * area;
        F[r] = conv + diss;
    }
}

struct FdArgs {
    int N, NT;
    const int *order;
    const int *ip, *sptr, *nbr, *soff, *iau;
    const double *isx, *isy, *isz, *qa, *qb, *vol, *dt, *res;
    const int *btype, *bcase;
    const double *case_inf;
    double *a;
    float *af;
    double *b;
    double g, frac, coeff;
    int efix, rhs_mode, add_time, order2, lim, plus, minus, iter, zpg;
    BcFree<double> free;
    int *flag;
};

// Face states of an interior edge: Compute_SecondOrderReconstructQ
// (Euler3D/src/HigherOrderReconstructQ.cpp:20-71) for second order, the node
// values for first order.  o: this node, n: its neighbour, iL: this node is
// the L end.  Component c of the end `pend` (0 L, 1 R) is perturbed by dq if
// perturb is set.  All choices are selects, so nothing is indexed at run time.
struct FdNode {
    double q[5], x[3], g[15], p[5];
};

__device__ __forceinline__ void fd_faces(const FdArgs &a, const FdNode &o, const FdNode &n, bool iL, bool perturb,
                                         int pend, int c, double dq, double *QL, double *QR) {
    double qL[5], qR[5];
#pragma unroll
    for (int m = 0; m < 5; m++) {
        const double l = iL ? o.q[m] : n.q[m], r = iL ? n.q[m] : o.q[m];
        qL[m] = (perturb && pend == 0 && m == c) ? l + dq : l;
        qR[m] = (perturb && pend == 1 && m == c) ? r + dq : r;
    }
    if (!a.order2) {
#pragma unroll
        for (int m = 0; m < 5; m++) {
            QL[m] = qL[m];
            QR[m] = qR[m];
        }
        return;
    }
    const double xl0 = iL ? o.x[0] : n.x[0], xl1 = iL ? o.x[1] : n.x[1], xl2 = iL ? o.x[2] : n.x[2];
    const double xr0 = iL ? n.x[0] : o.x[0], xr1 = iL ? n.x[1] : o.x[1], xr2 = iL ? n.x[2] : o.x[2];
    const double RxL = 0.5 * (xr0 - xl0), RyL = 0.5 * (xr1 - xl1), RzL = 0.5 * (xr2 - xl2);
    const double RxR = -RxL, RyR = -RyL, RzR = -RzL;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const double pL = a.lim ? (iL ? o.p[k] : n.p[k]) : 1.0, pR = a.lim ? (iL ? n.p[k] : o.p[k]) : 1.0;
        const double glx = iL ? o.g[3 * k] : n.g[3 * k], gly = iL ? o.g[3 * k + 1] : n.g[3 * k + 1],
                     glz = iL ? o.g[3 * k + 2] : n.g[3 * k + 2];
        const double grx = iL ? n.g[3 * k] : o.g[3 * k], gry = iL ? n.g[3 * k + 1] : o.g[3 * k + 1],
                     grz = iL ? n.g[3 * k + 2] : o.g[3 * k + 2];
        QL[k] = qL[k] + pL * (a.frac * (qR[k] - qL[k]) + (1.0 - a.frac) * (glx * RxL + gly * RyL + glz * RzL));
        QR[k] = qR[k] + pR * (a.frac * (qL[k] - qR[k]) + (1.0 - a.frac) * (grx * RxR + gry * RyR + grz * RzR));
    }
}

// Loads the record of node j: state, coordinates, and for second order the
// gradient and limiter.
__device__ __forceinline__ void fd_load(const FdArgs &a, int j, FdNode &r) {
    const double *ra = a.qa + kQA * (size_t)j;
#pragma unroll
    for (int m = 0; m < 5; m++) r.q[m] = ra[m];
#pragma unroll
    for (int m = 0; m < 3; m++) r.x[m] = ra[5 + m];
    if (a.order2 && j < a.N) {
        const double *rb = a.qb + kQB * (size_t)j;
#pragma unroll
        for (int f = 0; f < 15; f++) r.g[f] = rb[f];
#pragma unroll
        for (int m = 0; m < 5; m++) r.p[m] = rb[kPHI + m];
    } else {
#pragma unroll
        for (int f = 0; f < 15; f++) r.g[f] = 0.0;
#pragma unroll
        for (int m = 0; m < 5; m++) r.p[m] = 1.0;
    }
}

// The flux one finite-difference column differentiates: the matrix-form Roe
// flux (Compute_RoeFlux), or the scheme's own flux for HLLC and AUSM+ (the
// scheme-agnostic Compute_Jacobian_FiniteDifference,
// MPFLOW3D/src/FiniteDifference_Jacobian.cpp:112-667), whose dissipative
// part is zero.  F = conv + diss, as the oracle sums them.
template <int FLUX>
__device__ __forceinline__ void fd_edge_flux(const FdArgs &a, const double *QL, const double *QR, double sx, double sy,
                                             double sz, double *F) {
    if (FLUX == CFD_FLUX_ROE) {
        roe_flux_matrix(a.g, a.efix, QL, QR, sx, sy, sz, F);
        return;
    }
    double cf[5];
    int bad = 0;
    if (FLUX == CFD_FLUX_HLLC) hllc_flux(a.g, sx, sy, sz, QL, QR, cf, bad);
    else ausmp_flux(a.g, sx, sy, sz, QL, QR, cf);
    if (bad) atomicAdd(a.flag + 1, 1);
#pragma unroll
    for (int r = 0; r < 5; r++) F[r] = cf[r] + 0.0;
}

// Compute_Jacobian_FiniteDifference_Roe, Euler3D/src/Roe_Jacobian.cpp:236-776,
// assembled by rows with ten lanes per node: lane (role, c) walks the node's
// incidences in the reference's edge order and forms column c of
//   role 0: the derivative of each incident flux with respect to this node,
//           summed into the diagonal block (+ as L, - as R, then boundary
//           edges, central, ghost recomputed per perturbation);
//   role 1: the derivative with respect to the neighbour, written to the
//           off-diagonal block (0 + dF/dR as L, 0 - dF/dL as R).
// Each column costs two flux evaluations per incidence: perturbed +/- eps,
// or perturbed and unperturbed for the one-sided methods (the reference
// evaluates the unperturbed flux once per edge end and reuses it; the value
// is the same).
template <int FLUX>
__global__ __launch_bounds__(128) void k_jacobian_fd(FdArgs a) {
    const int t = blockIdx.x * 128 + threadIdx.x;
    const int i = t / 10;  // internal node
    if (i >= a.N) return;
    const int role = (t % 10) / 5, c = t % 5;
    const int ai = a.order[i];
    const double eps = 1.0E-8;
    FdNode own;
    fd_load(a, i, own);
    double D[5];  // column c of the diagonal block
#pragma unroll
    for (int r = 0; r < 5; r++) D[r] = (a.add_time && r == c) ? a.vol[i] / a.dt[i] : 0.0;
    const int deg = a.ip[i + 1] - a.ip[i];
    size_t idx = (size_t)a.sptr[i / kSlice] + (i % kSlice);
    for (int k = 0; k < deg; k++, idx += kSlice) {
        const int jr = ld_stream(a.nbr + idx);
        const bool iL = jr >= 0;  // ghosts (jr >= N) always pair with a physical L
        const int j = iL ? jr : ~jr;
        const double sx = ld_stream(a.isx + idx), sy = ld_stream(a.isy + idx), sz = ld_stream(a.isz + idx);
        if (j >= a.N) {  // boundary edge: central, this node perturbed
            if (role) continue;
            const int be = j - a.N;
            const double *rg = a.qa + kQA * (size_t)j;
            double qg[5];
#pragma unroll
            for (int m = 0; m < 5; m++) qg[m] = rg[m];
            const BcFree<double> fr = a.case_inf ? bc_free_of_case<double>(a.case_inf, a.bcase[be]) : a.free;
            double Fp[2][5];
#pragma unroll
            for (int pm = 0; pm < 2; pm++) {
                double QP[5], QG[5];
#pragma unroll
                for (int m = 0; m < 5; m++) QP[m] = (m == c) ? (pm ? own.q[m] - eps : own.q[m] + eps) : own.q[m];
                const int st = bc_ghost(a.g, QP, qg, sx, sy, sz, a.btype[be], a.iter, a.zpg, fr, QG);
                if (st == 2) {
                    atomicAdd(a.flag + 3, 1);
#pragma unroll
                    for (int m = 0; m < 5; m++) QG[m] = qg[m];
                } else if (st == 1) {
                    atomicAdd(a.flag + 0, 1);
                }
                fd_edge_flux<FLUX>(a, QP, QG, sx, sy, sz, Fp[pm]);
            }
#pragma unroll
            for (int r = 0; r < 5; r++) D[r] += 0.5 * (Fp[0][r] - Fp[1][r]) / eps;
            continue;
        }
        FdNode nb;
        fd_load(a, j, nb);
        // the perturbed end: this node (role 0) or the neighbour (role 1)
        const int pend = ((role == 0) == iL) ? 0 : 1;
        double FP[5], FM[5];
        {
            double QL[5], QR[5];
            fd_faces(a, own, nb, iL, a.plus, pend, c, eps, QL, QR);
            fd_edge_flux<FLUX>(a, QL, QR, sx, sy, sz, FP);
            fd_faces(a, own, nb, iL, a.minus, pend, c, -eps, QL, QR);
            fd_edge_flux<FLUX>(a, QL, QR, sx, sy, sz, FM);
        }
        double col[5];
#pragma unroll
        for (int r = 0; r < 5; r++) col[r] = a.coeff * (FP[r] - FM[r]) / eps;
        if (role == 0) {
#pragma unroll
            for (int r = 0; r < 5; r++) D[r] = iL ? D[r] + col[r] : D[r] - col[r];
        } else {
            const size_t o = 25 * (size_t)ld_stream(a.soff + idx) + c;
#pragma unroll
            for (int r = 0; r < 5; r++) {
                const double v = iL ? 0.0 + col[r] : 0.0 - col[r];
                if (a.af) a.af[o + 5 * r] = (float)v;
                else a.a[o + 5 * r] = v;
            }
        }
    }
    if (role == 0) {
        const size_t o = 25 * (size_t)a.iau[ai] + c;
#pragma unroll
        for (int r = 0; r < 5; r++) {
            if (a.af) a.af[o + 5 * r] = (float)D[r];
            else a.a[o + 5 * r] = D[r];
        }
        if (c == 0) {
#pragma unroll
            for (int k = 0; k < 5; k++) {
                double r = a.res[kRES * (size_t)i + k];
                if (a.rhs_mode == 1) r = r + a.res[kRES * (size_t)i + 5 + k];
                a.b[5 * (size_t)ai + k] = -r;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Dense 5x5 helpers (same operation order as the CPU oracle)
// ---------------------------------------------------------------------------
// Gauss-Jordan inverse with partial pivoting; returns false on a zero pivot.
// Every loop is unrolled and the pivot-row exchange is a select per entry, so
// the working matrix stays in registers (a dynamically indexed row swap put it
// in scratch memory); the arithmetic is the textbook sequence, unchanged.
__device__ bool inv5(const double *A, double *Ai) {
    double a[5][10];
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 5; c++) {
            a[r][c] = A[5 * r + c];
            a[r][c + 5] = (r == c) ? 1.0 : 0.0;
        }
#pragma unroll
    for (int c = 0; c < 5; c++) {
        // first row of largest magnitude in column c (ties keep the earlier row)
        int piv = c;
        double best = fabs(a[c][c]);
#pragma unroll
        for (int r = c + 1; r < 5; r++) {
            const double v = fabs(a[r][c]);
            if (v > best) {
                piv = r;
                best = v;
            }
        }
#pragma unroll
        for (int r = c + 1; r < 5; r++) {
            const bool sw = piv == r;
#pragma unroll
            for (int j = 0; j < 10; j++) {
                const double t = a[c][j];
                a[c][j] = sw ? a[r][j] : t;
                a[r][j] = sw ? t : a[r][j];
            }
        }
        const double d = a[c][c];
        if (d == 0.0) return false;
#pragma unroll
        for (int j = 0; j < 10; j++) a[c][j] = a[c][j] / d;
#pragma unroll
        for (int r = 0; r < 5; r++) {
            if (r == c) continue;
            const double f = a[r][c];
#pragma unroll
            for (int j = 0; j < 10; j++) a[r][j] = a[r][j] - f * a[c][j];
        }
    }
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 5; c++) Ai[5 * r + c] = a[r][c + 5];
    return true;
}

template <class TM>
__device__ __forceinline__ void bmul(const TM *A, const TM *B, TM *C) {
#pragma unroll
    for (int r = 0; r < 5; r++)
#pragma unroll
        for (int c = 0; c < 5; c++) {
            TM s = TM(0.0);
#pragma unroll
            for (int t = 0; t < 5; t++) s += A[5 * r + t] * B[5 * t + c];
            C[5 * r + c] = s;
        }
}

// ---------------------------------------------------------------------------
// Block ILU(0), one level: Saad Alg. 10.4 in block form, identical to
// orc_ilu0.  Factors in TM; the pivot inverse is computed in FP64.
// ---------------------------------------------------------------------------
template <class TM>
__global__ __launch_bounds__(kBlock) void k_ilu0_level(int nrows, const int *__restrict__ rows,
                                                      const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const int *__restrict__ iau, const int *__restrict__ moff,
                                                      const int *__restrict__ mqq, TM *lu, TM *dinv,
                                                      const TM *__restrict__ luk, const TM *__restrict__ dinvk,
                                                      int *flag) {
    // luk / dinvk view the blocks and inverse pivots of rows of earlier
    // levels, which this launch only reads (it writes rows of one level).
    // 12 rows of 5 lanes per wavefront; lane r owns row r of every block of
    // its block row, so the products of two 5x5 blocks are split by rows with
    // the per-entry summation order of bmul.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / 5, r = lane - 5 * slot;
    const int t = (blockIdx.x * (kBlock / 64) + wave) * 12 + slot;
    const bool active = slot < 12 && t < nrows;
    const int i = active ? rows[t] : 0;
    TM drow[5] = {TM(0.0), TM(0.0), TM(0.0), TM(0.0), TM(0.0)};
    if (active) {
        // the lanes of a block row hold one row each of every block they
        // load; a product needs the other rows, which come from the group's
        // lanes by shuffles (the five lanes of a row run every loop together)
        const int gbase = lane - r;
        for (int kk = ia[i]; kk < iau[i]; kk++) {
            const int k = ja[kk];
            TM *Lk = lu + 25 * (size_t)kk + 5 * r;
            const TM *Dk = dinvk + 25 * (size_t)k + 5 * r;
            TM Lr[5], U[5], Dr[5];
#pragma unroll
            for (int q = 0; q < 5; q++) U[q] = Lk[q];
#pragma unroll
            for (int c = 0; c < 5; c++) Dr[c] = Dk[c];
#pragma unroll
            for (int c = 0; c < 5; c++) {
                TM s = TM(0.0);
#pragma unroll
                for (int q = 0; q < 5; q++) s += U[q] * __shfl(Dr[c], gbase + q, 64);
                Lr[c] = s;
            }
#pragma unroll
            for (int c = 0; c < 5; c++) Lk[c] = Lr[c];
            // upper blocks of row k that meet a block of row i, in ascending
            // column order (the precomputed merge of the two rows), CH at a
            // time: their slots, rows and targets are loaded before any
            // target is updated.  The targets are distinct blocks of row i, so
            // each update is the one-at-a-time arithmetic.
            constexpr int CH = sizeof(TM) == 4 ? 8 : 4;
            const int p0 = iau[k] + 1, np = ia[k + 1] - p0, m0 = moff[kk];
            for (int t2 = 0; t2 < np; t2 += CH) {
                int qs[CH];
                TM ur[CH][5], tv[CH][5];
#pragma unroll
                for (int u = 0; u < CH; u++) qs[u] = t2 + u < np ? mqq[m0 + t2 + u] : -1;
#pragma unroll
                for (int u = 0; u < CH; u++)
                    if (qs[u] >= 0) {
                        const TM *Up = luk + 25 * (size_t)(p0 + t2 + u) + 5 * r;
#pragma unroll
                        for (int c = 0; c < 5; c++) ur[u][c] = Up[c];
#pragma unroll
                        for (int c = 0; c < 5; c++) tv[u][c] = lu[25 * (size_t)qs[u] + 5 * r + c];
                    }
#pragma unroll
                for (int u = 0; u < CH; u++) {
                    if (qs[u] < 0) continue;
                    TM *T = lu + 25 * (size_t)qs[u] + 5 * r;
#pragma unroll
                    for (int c = 0; c < 5; c++) {
                        TM s = TM(0.0);
#pragma unroll
                        for (int q = 0; q < 5; q++) s += Lr[q] * __shfl(ur[u][c], gbase + q, 64);
                        T[c] = tv[u][c] - s;
                    }
                }
            }
        }
        const TM *Dp = lu + 25 * (size_t)iau[i] + 5 * r;
#pragma unroll
        for (int c = 0; c < 5; c++) drow[c] = Dp[c];
    }
    // every lane of the row gathers the diagonal block and inverts it (FP64)
    const int base = lane - r;
    double Dg[25], Di[25];
#pragma unroll
    for (int q = 0; q < 5; q++)
#pragma unroll
        for (int c = 0; c < 5; c++) Dg[5 * q + c] = (double)__shfl(drow[c], base + q, 64);
    if (!active) return;
    if (!inv5(Dg, Di)) {
        if (r == 0) atomicAdd(flag + 2, 1);
        for (int q = 0; q < 25; q++) Di[q] = 0.0;
    }
    // row r of the inverse, picked with selects (an index by r would put Di in
    // scratch memory)
    double dr[5] = {Di[0], Di[1], Di[2], Di[3], Di[4]};
#pragma unroll
    for (int rr = 1; rr < 5; rr++)
#pragma unroll
        for (int c = 0; c < 5; c++) dr[c] = r == rr ? Di[5 * rr + c] : dr[c];
#pragma unroll
    for (int c = 0; c < 5; c++) dinv[25 * (size_t)i + 5 * r + c] = (TM)dr[c];
}

// Inverse diagonal blocks (block-Jacobi).
template <class TA, class TM>
__global__ __launch_bounds__(kBlock) void k_block_inv(int N, const int *iau, const TA *a, TM *dinv, int *flag) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double Dg[25], Di[25];
    for (int q = 0; q < 25; q++) Dg[q] = (double)a[25 * (size_t)iau[i] + q];
    if (!inv5(Dg, Di)) {
        atomicAdd(flag + 2, 1);
        for (int q = 0; q < 25; q++) Di[q] = 0.0;
    }
    for (int q = 0; q < 25; q++) dinv[25 * (size_t)i + q] = (TM)Di[q];
}

// Triangular solves of ILU(0), one level; 5 lanes per row (lane r owns
// component r).  Forward: z_i = r_i - sum_{k<i} L_ik z_k.  Backward:
// z_i = D_i^-1 (z_i - sum_{k>i} U_ik z_k).
template <class TM, class TV>
__global__ __launch_bounds__(kBlock) void k_trsv_level(int nrows, const int *__restrict__ rows,
                                                       const int *__restrict__ ia, const int *__restrict__ ja,
                                                       const int *__restrict__ iau, const TM *__restrict__ lu,
                                                       const TM *__restrict__ dinv, const TV *__restrict__ rhs, TV *z,
                                                       int backward) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / 5, r = lane - 5 * slot;
    const int t = (blockIdx.x * (kBlock / 64) + wave) * 12 + slot;
    const bool active = slot < 12 && t < nrows;
    const int i = active ? rows[t] : 0;
    TV y = TV(0.0);
    if (active) {
        if (!backward) {
            y = rhs[5 * (size_t)i + r];
            for (int kk = ia[i]; kk < iau[i]; kk++) {
                const TM *A = lu + 25 * (size_t)kk + 5 * r;
                const TV *zj = z + 5 * (size_t)ja[kk];
                TV s = TV(0.0);
#pragma unroll
                for (int c = 0; c < 5; c++) s += (TV)A[c] * zj[c];
                y = y - s;
            }
        } else {
            y = z[5 * (size_t)i + r];
            for (int kk = iau[i] + 1; kk < ia[i + 1]; kk++) {
                const TM *A = lu + 25 * (size_t)kk + 5 * r;
                const TV *zj = z + 5 * (size_t)ja[kk];
                TV s = TV(0.0);
#pragma unroll
                for (int c = 0; c < 5; c++) s += (TV)A[c] * zj[c];
                y = y - s;
            }
        }
    }
    if (!backward) {
        if (active) z[5 * (size_t)i + r] = y;
        return;
    }
    // gather the row's 5 components from its lanes, then apply D_i^-1
    const int base = lane - r;
    TV yy[5];
#pragma unroll
    for (int c = 0; c < 5; c++) yy[c] = __shfl(y, base + c, 64);
    if (active) {
        const TM *D = dinv + 25 * (size_t)i + 5 * r;
        TV s = TV(0.0);
#pragma unroll
        for (int c = 0; c < 5; c++) s += (TV)D[c] * yy[c];
        z[5 * (size_t)i + r] = s;
    }
}

// ---------------------------------------------------------------------------
// Sync-free triangular sweeps (ILU(0) apply): one launch per sweep instead of
// one per level.  The factorization, once per Jacobian, stays one launch per
// level.
// Rows are taken in level order and every wavefront holds rows of a single
// level (levels are padded to whole wavefronts, pad = -1), so a row waits
// only on rows of earlier wavefronts.  Workgroups are dispatched in index
// order, so the earliest unfinished row always has a slot and the sweep
// always progresses.
//
// Each XCD has its own L2 and the L2s are not coherent with one another; an
// agent-scope acquire invalidates the whole local L2, so it must not sit in a
// polling loop.  The values one row hands to another are written with
// coherent (relaxed agent-scope atomic) stores that go past the L2, then the
// row sets ready[row] = epoch with an agent-scope release store (one L2
// write-back per finished row).  A waiting row polls the flag with coherent
// loads, orders
// its later loads after the poll (workgroup acquire fence: no cache
// invalidation) and reads the handed-over values with coherent loads, which
// bypass stale cache lines.  The epoch is read from device
// memory (bumped by k_epoch_bump before each sweep), so ready[] never needs
// clearing.  A wait that exceeds kSpinMax polls raises flag[3]; every waiter
// also watches flag[3] and gives up once it is set, so a sweep always ends,
// after at most one bounded wait.
// ---------------------------------------------------------------------------
constexpr int kSpinMax = 1 << 22;

__global__ void k_epoch_bump(int *epoch) {
    if (threadIdx.x == 0 && blockIdx.x == 0) epoch[0] = epoch[0] + 1;
}

template <class T>
__device__ __forceinline__ T ld_coh(const T *p) {
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_coh(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool wait_ready(const int *ready, int j, int ep, int *flag) {
    bool ok = false;
    for (int it = 0; it < kSpinMax; it++) {
        if (ld_coh(ready + j) == ep) {
            ok = true;
            break;
        }
        if ((it & 255) == 255 && ld_coh(flag + 3) != 0) break;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (!ok && ld_coh(flag + 3) == 0) atomicOr(flag + 3, 1);
    return ok;
}

// The row's stores are visible device-wide before its flag is set (an
// s_waitcnt on the coherent stores alone does not order them before the flag
// for readers on other XCDs).
__device__ __forceinline__ void publish_ready(int *ready, int i, int ep) {
    __hip_atomic_store(ready + i, ep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Same arithmetic as k_trsv_level; 12 rows of 5 lanes per wavefront.
template <class TM, class TV>
__global__ __launch_bounds__(kBlock) void k_trsv_sf(int nslots, const int *__restrict__ slots,
                                                    const int *__restrict__ ia, const int *__restrict__ ja,
                                                    const int *__restrict__ iau, const TM *__restrict__ lu,
                                                    const TM *__restrict__ dinv, const TV *__restrict__ rhs, TV *z,
                                                    int backward, int *ready, const int *epoch, int *flag) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / 5, r = lane - 5 * slot;
    const int t = (blockIdx.x * (kBlock / 64) + wave) * 12 + slot;
    const int i = (slot < 12 && t < nslots) ? slots[t] : -1;
    const int ep = epoch[0];
    TV y = TV(0.0);
    bool ok = true;
    if (i >= 0) {
        const int kb = backward ? iau[i] + 1 : ia[i];
        const int ke = backward ? ia[i + 1] : iau[i];
        y = backward ? ld_coh(z + 5 * (size_t)i + r) : rhs[5 * (size_t)i + r];
        for (int kk = kb; kk < ke; kk++) {
            const int j = ja[kk];
            ok = wait_ready(ready, j, ep, flag) && ok;
            const TM *A = lu + 25 * (size_t)kk + 5 * r;
            const TV *zj = z + 5 * (size_t)j;
            TV s = TV(0.0);
#pragma unroll
            for (int c = 0; c < 5; c++) s += (TV)A[c] * ld_coh(zj + c);
            y = y - s;
        }
    }
    if (backward) {
        // gather the row's 5 components from its lanes, then apply D_i^-1
        const int base = lane - r;
        TV yy[5];
#pragma unroll
        for (int c = 0; c < 5; c++) yy[c] = __shfl(y, base + c, 64);
        if (i >= 0) {
            const TM *D = dinv + 25 * (size_t)i + 5 * r;
            TV s = TV(0.0);
#pragma unroll
            for (int c = 0; c < 5; c++) s += (TV)D[c] * yy[c];
            y = s;
        }
    }
    if (i >= 0) st_coh(z + 5 * (size_t)i + r, y);
    if (i >= 0 && r == 0) publish_ready(ready, i, ep);
    if (!ok) atomicOr(flag + 3, 1);
}

// Blocks of the matrix in multicolour order: dst[s] = src[map[s]].
template <class TS, class TD>
__global__ __launch_bounds__(kBlock) void k_gather_blocks(size_t nnz, const int *__restrict__ map,
                                                          const TS *__restrict__ src, TD *__restrict__ dst) {
    for (size_t t = blockIdx.x * (size_t)kBlock + threadIdx.x; t < 25 * nnz; t += (size_t)gridDim.x * kBlock) {
        const size_t s = t / 25, q = t - 25 * s;
        dst[t] = (TD)src[25 * (size_t)map[s] + q];
    }
}

// One level of the multicolour ILU(0) sweeps; rows and columns are in
// multicolour order, the right-hand side and the result in the reference
// order (perm maps a multicolour row to its reference row).  Forward:
// w_i = r_perm(i) - sum_{k<i} L_ik w_k.  Backward: w_i = D_i^-1 (w_i -
// sum_{k>i} U_ik w_k), z_perm(i) = w_i.  5 lanes per row, as k_trsv_level.
template <class TM, class TV>
__global__ __launch_bounds__(kBlock) void k_trsv_mc(int nrows, const int *__restrict__ rows,
                                                    const int *__restrict__ ia, const int *__restrict__ ja,
                                                    const int *__restrict__ iau, const int *__restrict__ perm,
                                                    const TM *__restrict__ lu, const TM *__restrict__ dinv,
                                                    const TV *__restrict__ rhs, TV *w, TV *z, int backward) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / 5, r = lane - 5 * slot;
    const int t = (blockIdx.x * (kBlock / 64) + wave) * 12 + slot;
    const bool active = slot < 12 && t < nrows;
    const int i = active ? rows[t] : 0;
    TV y = TV(0.0);
    if (active) {
        const int kb = backward ? iau[i] + 1 : ia[i];
        const int ke = backward ? ia[i + 1] : iau[i];
        y = backward ? w[5 * (size_t)i + r] : rhs[5 * (size_t)perm[i] + r];
        // four blocks at a time: their loads are all issued before the
        // products are subtracted, in block order as one at a time
        for (int k0 = kb; k0 < ke; k0 += 4) {
            int jj[4];
#pragma unroll
            for (int u = 0; u < 4; u++) jj[u] = k0 + u < ke ? ja[k0 + u] : -1;
            TM av[4][5];
            TV wv[4][5];
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (jj[u] >= 0) {
#pragma unroll
                    for (int c = 0; c < 5; c++) {
                        av[u][c] = lu[25 * (size_t)(k0 + u) + 5 * r + c];
                        wv[u][c] = w[5 * (size_t)jj[u] + c];
                    }
                }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (jj[u] >= 0) {
                    TV s = TV(0.0);
#pragma unroll
                    for (int c = 0; c < 5; c++) s += (TV)av[u][c] * wv[u][c];
                    y = y - s;
                }
        }
    }
    if (!backward) {
        if (active) w[5 * (size_t)i + r] = y;
        return;
    }
    const int base = lane - r;
    TV yy[5];
#pragma unroll
    for (int c = 0; c < 5; c++) yy[c] = __shfl(y, base + c, 64);
    if (active) {
        const TM *D = dinv + 25 * (size_t)i + 5 * r;
        TV s = TV(0.0);
#pragma unroll
        for (int c = 0; c < 5; c++) s += (TV)D[c] * yy[c];
        w[5 * (size_t)i + r] = s;
        z[5 * (size_t)perm[i] + r] = s;
    }
}

// One colour of the multicolour ILU(0) sweeps with the factors in block-SELL
// layout: rows r0..r1-1 (permuted order, one colour, independent of each
// other), one row per lane.  Forward: w_i = r_P(i) - sum over the strict lower
// blocks L_ij w_j; backward: w_i = D_i^-1 (w_i - sum over the strict upper
// blocks U_ij w_j), also scattered to z in original order.  Same operation
// order as k_trsv_mc.  Workgroups are numbered XCD by XCD (the grid is a
// multiple of the XCD count), so each XCD sweeps a contiguous range of the
// colour's rows and gathers from fewer lines of w.
template <class TM, class TV>
__global__ __launch_bounds__(kBlock) void k_trsv_mc_sell(int r0, int r1, const int *__restrict__ ia,
                                                         const int *__restrict__ iau, const int *__restrict__ perm,
                                                         const int *__restrict__ off, const int *__restrict__ col,
                                                         const TM *__restrict__ val, const TM *__restrict__ dv,
                                                         const TV *__restrict__ rhs, TV *w, TV *z, int backward) {
    const int i = (r0 & ~63) + xcd_block() * kBlock + threadIdx.x;
    if (i < r0 || i >= r1) return;
    const int l = i & 63;
    const int len = backward ? ia[i + 1] - iau[i] - 1 : iau[i] - ia[i];
    TV y[5];
#pragma unroll
    for (int r = 0; r < 5; r++) y[r] = backward ? w[5 * (size_t)i + r] : rhs[5 * (size_t)perm[i] + r];
    // U blocks at a time: their column indices, then their x values and
    // entries (streamed: each is read once per sweep) are loaded before the
    // products, which are then subtracted in block order as before.
    constexpr int U = sizeof(TM) == 8 ? 2 : 4;
    size_t p = (size_t)off[i >> 6] + l;
    for (int k0 = 0; k0 < len; k0 += U, p += 64 * U) {
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; u++) jj[u] = k0 + u < len ? col[p + 64 * u] : -1;
        TV wv[U][5];
        TM av[U][25];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jj[u] >= 0) {
                const TV *wj = w + 5 * (size_t)jj[u];
#pragma unroll
                for (int c = 0; c < 5; c++) wv[u][c] = wj[c];
                const TM *A = val + 25 * (p + 64 * u - l) + l;
#pragma unroll
                for (int e = 0; e < 25; e++) av[u][e] = ld_stream(A + 64 * e);
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jj[u] >= 0) {
#pragma unroll
                for (int r = 0; r < 5; r++) {
                    TV s = TV(0.0);
#pragma unroll
                    for (int c = 0; c < 5; c++) s += (TV)av[u][5 * r + c] * wv[u][c];
                    y[r] = y[r] - s;
                }
            }
    }
    if (!backward) {
#pragma unroll
        for (int r = 0; r < 5; r++) w[5 * (size_t)i + r] = y[r];
        return;
    }
    const TM *D = dv + 25 * (size_t)(i - l) + l;
#pragma unroll
    for (int r = 0; r < 5; r++) {
        TV s = TV(0.0);
#pragma unroll
        for (int c = 0; c < 5; c++) s += (TV)D[64 * (5 * r + c)] * y[c];
        w[5 * (size_t)i + r] = s;
        z[5 * (size_t)perm[i] + r] = s;
    }
}

// FP16 block-SELL copy of FP32 blocks: each block's 25 entries over its
// largest magnitude (a zero block keeps scale 1), the scale kept per slot.
__global__ __launch_bounds__(kBlock) void k_to_sell_half(size_t nblk, const int *__restrict__ src,
                                                         const float *__restrict__ v, _Float16 *__restrict__ out,
                                                         float *__restrict__ scale) {
    for (size_t p = blockIdx.x * (size_t)kBlock + threadIdx.x; p < nblk; p += (size_t)gridDim.x * kBlock) {
        const int q = src ? src[p] : (int)p;
        if (q < 0) continue;
        const size_t l = p & 63;
        _Float16 *o = out + 25 * (p - l) + l;
        const float *in = v + 25 * (size_t)q;
        float mx = 0.0f;
#pragma unroll
        for (int e = 0; e < 25; e++) mx = fmaxf(mx, fabsf(in[e]));
        const float sc = mx > 0.0f ? mx : 1.0f, inv = 1.0f / sc;
#pragma unroll
        for (int e = 0; e < 25; e++) o[64 * e] = (_Float16)(in[e] * inv);
        scale[p] = sc;
    }
}

// k_trsv_mc_sell with the factors stored in FP16 (per-block FP32 scales):
// each block's product is formed in FP32 and scaled once.
__global__ __launch_bounds__(kBlock) void k_trsv_mc_sell_half(int r0, int r1, const int *__restrict__ ia,
                                                              const int *__restrict__ iau, const int *__restrict__ perm,
                                                              const int *__restrict__ off, const int *__restrict__ col,
                                                              const _Float16 *__restrict__ val,
                                                              const float *__restrict__ vsc,
                                                              const _Float16 *__restrict__ dv,
                                                              const float *__restrict__ dsc,
                                                              const float *__restrict__ rhs, float *w, float *z,
                                                              int backward) {
    const int i = (r0 & ~63) + xcd_block() * kBlock + threadIdx.x;
    if (i < r0 || i >= r1) return;
    const int l = i & 63;
    const int len = backward ? ia[i + 1] - iau[i] - 1 : iau[i] - ia[i];
    float y[5];
#pragma unroll
    for (int r = 0; r < 5; r++) y[r] = backward ? w[5 * (size_t)i + r] : rhs[5 * (size_t)perm[i] + r];
    constexpr int U = 4;  // blocks whose loads are in flight together, as in k_trsv_mc_sell
    size_t p = (size_t)off[i >> 6] + l;
    for (int k0 = 0; k0 < len; k0 += U, p += 64 * U) {
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; u++) jj[u] = k0 + u < len ? col[p + 64 * u] : -1;
        float wv[U][5], sc[U];
        _Float16 av[U][25];
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jj[u] >= 0) {
                const float *wj = w + 5 * (size_t)jj[u];
#pragma unroll
                for (int c = 0; c < 5; c++) wv[u][c] = wj[c];
                const _Float16 *A = val + 25 * (p + 64 * u - l) + l;
#pragma unroll
                for (int e = 0; e < 25; e++) av[u][e] = ld_stream(A + 64 * e);
                sc[u] = ld_stream(vsc + p + 64 * u);
            }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jj[u] >= 0) {
#pragma unroll
                for (int r = 0; r < 5; r++) {
                    float s = 0.0f;
#pragma unroll
                    for (int c = 0; c < 5; c++) s += (float)av[u][5 * r + c] * wv[u][c];
                    y[r] = y[r] - sc[u] * s;
                }
            }
    }
    if (!backward) {
#pragma unroll
        for (int r = 0; r < 5; r++) w[5 * (size_t)i + r] = y[r];
        return;
    }
    const _Float16 *D = dv + 25 * (size_t)(i - l) + l;
    const float sd = dsc[i];
#pragma unroll
    for (int r = 0; r < 5; r++) {
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 5; c++) s += (float)D[64 * (5 * r + c)] * y[c];
        s *= sd;
        w[5 * (size_t)i + r] = s;
        z[5 * (size_t)perm[i] + r] = s;
    }
}

// z = D^-1 r (block-Jacobi apply)
template <class TM, class TV>
__global__ __launch_bounds__(kBlock) void k_bjac_apply(int N, const TM *dinv, const TV *rhs, TV *z) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 5 * N) return;
    const int i = t / 5, r = t - 5 * i;
    const TM *D = dinv + 25 * (size_t)i + 5 * r;
    const TV *x = rhs + 5 * (size_t)i;
    TV s = TV(0.0);
#pragma unroll
    for (int c = 0; c < 5; c++) s += (TV)D[c] * x[c];
    z[t] = s;
}

// y = A x, one thread per (row, component); sum over blocks then columns.
template <class TA, class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_spmv(int N, const int *__restrict__ ia, const int *__restrict__ ja,
                                                 const TA *__restrict__ a, const TX *__restrict__ x, TY *__restrict__ y,
                                                 const double *__restrict__ bsub) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= 5 * N) return;
    const int i = t / 5, r = t - 5 * i;
    TY s = TY(0.0);
    // four blocks at a time (loads first, then the products in order)
    const int ke = ia[i + 1];
    for (int k0 = ia[i]; k0 < ke; k0 += 4) {
        int jj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) jj[u] = k0 + u < ke ? ja[k0 + u] : -1;
        TA av[4][5];
        TX xv[4][5];
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (jj[u] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; c++) {
                    av[u][c] = a[25 * (size_t)(k0 + u) + 5 * r + c];
                    xv[u][c] = x[5 * (size_t)jj[u] + c];
                }
            }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (jj[u] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; c++) s += (TY)av[u][c] * (TY)xv[u][c];
            }
    }
    if (bsub) s = (TY)bsub[t] - s;  // residual form r = b - A x
    y[t] = s;
}

// Copies BSR-ordered blocks (25 values each) into a block-SELL layout: slot
// p takes block src[p]; padding slots are left as they are (never read).
template <class T>
__global__ __launch_bounds__(kBlock) void k_to_sell(size_t nblk, const int *__restrict__ src, const T *__restrict__ v,
                                                    T *__restrict__ out) {
    for (size_t p = blockIdx.x * (size_t)kBlock + threadIdx.x; p < nblk; p += (size_t)gridDim.x * kBlock) {
        const int q = src ? src[p] : (int)p;
        if (q < 0) continue;
        const size_t l = p & 63;
        T *o = out + 25 * (p - l) + l;
        const T *in = v + 25 * (size_t)q;
#pragma unroll
        for (int e = 0; e < 25; e++) o[64 * e] = in[e];
    }
}

// y = A x (or y = b - A x) with A in block-SELL layout, one row per lane.  Each
// row's sum runs over its blocks in BSR order and each block's columns in
// order, as k_spmv and the oracle add them, so the bits are the same.  Rows
// go to the XCDs in contiguous ranges (grid a multiple of the XCD count) and
// the entries, read once per product, are streamed past the L2.
template <class TA, class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(int N, const int *__restrict__ ia, const int *__restrict__ off,
                                                      const int *__restrict__ col, const TA *__restrict__ val,
                                                      const TX *__restrict__ x, TY *__restrict__ y,
                                                      const double *__restrict__ bsub) {
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= N) return;
    const int l = i & 63;
    const int len = ia[i + 1] - ia[i];
    size_t p = (size_t)off[i >> 6] + l;
    TY s[5];
#pragma unroll
    for (int r = 0; r < 5; r++) s[r] = TY(0.0);
    for (int k = 0; k < len; k++, p += 64) {
        const TX *xj = x + 5 * (size_t)col[p];
        TX xv[5];
#pragma unroll
        for (int c = 0; c < 5; c++) xv[c] = xj[c];
        const TA *A = val + 25 * (p - l) + l;
#pragma unroll
        for (int r = 0; r < 5; r++)
#pragma unroll
            for (int c = 0; c < 5; c++) s[r] += (TY)ld_stream(A + 64 * (5 * r + c)) * (TY)xv[c];
    }
#pragma unroll
    for (int r = 0; r < 5; r++) {
        TY v = s[r];
        if (bsub) v = (TY)bsub[5 * (size_t)i + r] - v;  // residual form r = b - A x
        y[5 * (size_t)i + r] = v;
    }
}

// k_spmv_sell with A stored in FP16 (per-block FP32 scales): each block's
// product is formed in FP32 and scaled once.
template <class TX, class TY>
__global__ __launch_bounds__(kBlock) void k_spmv_sell_half(int N, const int *__restrict__ ia,
                                                           const int *__restrict__ off, const int *__restrict__ col,
                                                           const _Float16 *__restrict__ val,
                                                           const float *__restrict__ vsc, const TX *__restrict__ x,
                                                           TY *__restrict__ y, const double *__restrict__ bsub) {
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= N) return;
    const int l = i & 63;
    const int len = ia[i + 1] - ia[i];
    size_t p = (size_t)off[i >> 6] + l;
    TY s[5];
#pragma unroll
    for (int r = 0; r < 5; r++) s[r] = TY(0.0);
    for (int k = 0; k < len; k++, p += 64) {
        const TX *xj = x + 5 * (size_t)col[p];
        float xv[5];
#pragma unroll
        for (int c = 0; c < 5; c++) xv[c] = (float)xj[c];
        const _Float16 *A = val + 25 * (p - l) + l;
        const float sc = ld_stream(vsc + p);
#pragma unroll
        for (int r = 0; r < 5; r++) {
            float t = 0.0f;
#pragma unroll
            for (int c = 0; c < 5; c++) t += (float)ld_stream(A + 64 * (5 * r + c)) * xv[c];
            s[r] += (TY)(sc * t);
        }
    }
#pragma unroll
    for (int r = 0; r < 5; r++) {
        TY v = s[r];
        if (bsub) v = (TY)bsub[5 * (size_t)i + r] - v;
        y[5 * (size_t)i + r] = v;
    }
}

// ---------------------------------------------------------------------------
// Reference sweeps (MATH/src/MC_Iterative_Jacobi.c), one level per launch
// ---------------------------------------------------------------------------
// No-pivot LU of the diagonal blocks, MC_Iterative_Jacobi.c:381-396
__global__ __launch_bounds__(kBlock) void k_gs_lu(int N, const int *iau, const double *a, double *lu) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= N) return;
    double L[25];
    for (int q = 0; q < 25; q++) L[q] = a[25 * (size_t)iau[r] + q];
    for (int k = 0; k < 4; k++) {
        for (int j = k + 1; j < 5; j++) L[5 * j + k] = L[5 * j + k] / L[5 * k + k];
        for (int j = k + 1; j < 5; j++)
            for (int i = k + 1; i < 5; i++) L[5 * i + j] -= L[5 * i + k] * L[5 * k + j];
    }
    for (int q = 0; q < 25; q++) lu[25 * (size_t)r + q] = L[q];
}

// One forward Gauss-Seidel row update, MC_Iterative_Jacobi.c:410-450
__global__ __launch_bounds__(kBlock) void k_gs_level(int nrows, const int *rows, const int *ia, const int *ja,
                                                     const int *iau, const double *a, const double *b,
                                                     const double *lu, double *x, double *rowrms) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nrows) return;
    const int r = rows[t];
    double rhs[5], tv[5];
    for (int m = 0; m < 5; m++) rhs[m] = b[5 * (size_t)r + m];
    for (int j = ia[r]; j < ia[r + 1]; j++) {
        const double *A = a + 25 * (size_t)j, *xj = x + 5 * (size_t)ja[j];
        for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
        for (int ii = 0; ii < 5; ii++)
            for (int jj = 0; jj < 5; jj++) tv[ii] += A[5 * ii + jj] * xj[jj];
        for (int m = 0; m < 5; m++) rhs[m] -= tv[m];
    }
    double s = 0.0;
    for (int m = 0; m < 5; m++) s += rhs[m] * rhs[m];
    rowrms[r] = s;
    const double *D = a + 25 * (size_t)iau[r];
    double *xr = x + 5 * (size_t)r;
    for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
    for (int ii = 0; ii < 5; ii++)
        for (int jj = 0; jj < 5; jj++) tv[ii] += D[5 * ii + jj] * xr[jj];
    for (int m = 0; m < 5; m++) rhs[m] += tv[m];
    const double *L = lu + 25 * (size_t)r;
    for (int ii = 0; ii < 5; ii++) {
        tv[ii] = rhs[ii];
        for (int jj = 0; jj < ii; jj++) tv[ii] -= L[5 * ii + jj] * tv[jj];
    }
    double xn[5];
    for (int ii = 4; ii >= 0; ii--) {
        xn[ii] = tv[ii];
        for (int jj = 4; jj > ii; jj--) xn[ii] -= L[5 * ii + jj] * xn[jj];
        xn[ii] = xn[ii] / L[5 * ii + ii];
    }
    for (int ii = 0; ii < 5; ii++) xr[ii] = xn[ii];
}

// LU-SGS row update with pivoted LU of the diagonal recomputed per sweep,
// MC_Iterative_Jacobi.c:586-678 and MATH/src/LinearAlgebra.c:20-41,117-155
__global__ __launch_bounds__(kBlock) void k_lusgs_level(int nrows, const int *rows, const int *ia, const int *ja,
                                                        const int *iau, const double *a, const double *b, double *x) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nrows) return;
    const int r = rows[t];
    double rhs[5], tv[5];
    for (int j = 0; j < 5; j++) rhs[j] = b[5 * (size_t)r + j];
    const int diag = iau[r];
    for (int j = ia[r]; j < ia[r + 1]; j++) {
        if (j == diag) continue;
        const double *A = a + 25 * (size_t)j, *xj = x + 5 * (size_t)ja[j];
        for (int ii = 0; ii < 5; ii++) tv[ii] = 0.0;
        for (int ii = 0; ii < 5; ii++)
            for (int jj = 0; jj < 5; jj++) tv[ii] += A[5 * ii + jj] * xj[jj];
        for (int m = 0; m < 5; m++) rhs[m] -= tv[m];
    }
    double A[5][5];
    int P[5];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) A[i][j] = a[25 * (size_t)diag + 5 * i + j];
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
            int tt = P[c];
            P[c] = P[pi];
            P[pi] = tt;
        }
        for (int i = c + 1; i < 5; i++) A[P[i]][c] /= A[P[c]][c];
        for (int i = c + 1; i < 5; i++)
            for (int j = c + 1; j < 5; j++) A[P[i]][j] -= A[P[i]][c] * A[P[c]][j];
    }
    double y[5], xs[5];
    y[0] = rhs[P[0]];
    for (int i = 1; i < 5; i++) {
        y[i] = rhs[P[i]];
        for (int j = 0; j < i; j++) y[i] -= y[j] * ((i == j) ? 1.0 : A[P[i]][j]);
    }
    xs[4] = y[4] / A[P[4]][4];
    for (int i = 3; i >= 0; i--) {
        xs[i] = y[i];
        for (int j = 4; j > i; j--) xs[i] -= xs[j] * A[P[i]][j];
        xs[i] /= A[P[i]][i];
    }
    for (int j = 0; j < 5; j++) x[5 * (size_t)r + j] = xs[j];
}

__global__ __launch_bounds__(kBlock) void k_lusgs_init(int N, const double *b, double *x) {
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= N) return;
    double s = 0.0;
    for (int j = 0; j < 5; j++) s += b[5 * (size_t)r + j];
    s /= 5;
    for (int j = 0; j < 5; j++) x[5 * (size_t)r + j] = s;
}

// ---------------------------------------------------------------------------
// Krylov vector kernels (deterministic reductions)
// ---------------------------------------------------------------------------
// part[blk][k] = sum over the block's elements of V_k . w, k < nv
// Partial dot products of w with V_0..V_{nv-1} over this block's strided
// elements [o + i0, o + n) (step `stride`), stored in part[0..nv).  The
// products are accumulated in register chunks of 8 vectors (w is re-read per
// chunk): each (thread, vector) sum visits the same elements in the same order
// and each vector goes through the same block tree, whatever the chunking.
template <class TV, bool DEEP = false>
__device__ __forceinline__ void block_dots(size_t o, size_t i0, size_t n, size_t stride, int nv,
                                           const TV *__restrict__ V, size_t ld, const TV *__restrict__ w,
                                           double *__restrict__ part) {
    __shared__ double wsum[8][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k0 = 0; k0 < nv; k0 += 8) {
        double acc[8];
#pragma unroll
        for (int kk = 0; kk < 8; kk++) acc[kk] = 0.0;
        // DEEP: four strides at a time, all their loads issued before the
        // products are added in the same order as one stride at a time (for
        // threads with many strides to walk; it costs registers, so short
        // walks keep the single-stride loop and its occupancy)
        size_t i = i0;
        for (; DEEP && i + 3 * stride < n; i += 4 * stride) {
            TV wv[4], vv[4][8];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                wv[u] = w[o + i + u * stride];
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (k0 + kk < nv) vv[u][kk] = V[ld * (k0 + kk) + o + i + u * stride];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double wi = (double)wv[u];
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (k0 + kk < nv) acc[kk] += (double)vv[u][kk] * wi;
            }
        }
        for (; i < n; i += stride) {
            const double wi = (double)w[o + i];
#pragma unroll
            for (int kk = 0; kk < 8; kk++)
                if (k0 + kk < nv) acc[kk] += (double)V[ld * (k0 + kk) + o + i] * wi;
        }
        // block sum: a butterfly within each wavefront (every lane ends with
        // the same value), then the four wavefront sums in order
#pragma unroll
        for (int kk = 0; kk < 8; kk++) {
            if (k0 + kk >= nv) break;
            double v = acc[kk];
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) v += __shfl_xor(v, sh, 64);
            if (lane == 0) wsum[kk][wave] = v;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int kk = 0; kk < 8 && k0 + kk < nv; kk++)
                part[k0 + kk] = ((wsum[kk][0] + wsum[kk][1]) + wsum[kk][2]) + wsum[kk][3];
        __syncthreads();
    }
}

template <class TV, bool DEEP>
__global__ __launch_bounds__(kBlock) void k_mdot(size_t n, int nv, const TV *__restrict__ V, size_t ld,
                                                 const TV *__restrict__ w, double *__restrict__ part) {
    block_dots<TV, DEEP>(0, blockIdx.x * (size_t)kBlock + threadIdx.x, n, (size_t)gridDim.x * kBlock, nv, V, ld, w,
                   part + (size_t)blockIdx.x * (kMaxKrylov + 1));
}

// Sum of the partials of dot k (block k): lane l adds partials l, l+64, ...
// in order, then the 64 lane sums are combined by a fixed butterfly.
__device__ __forceinline__ double wave_sum_parts(const double *part, int p0, int np, int k) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int p = lane; p < np; p += 64) s += part[(size_t)(p0 + p) * (kMaxKrylov + 1) + k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

__global__ __launch_bounds__(64) void k_mdot_final(int nparts, int nv, const double *part, double *out) {
    const int k = blockIdx.x;
    const double s = wave_sum_parts(part, 0, nparts, k);
    if (threadIdx.x == 0) out[k] = s;
}

// U entries i, i + stride, ... of the pass: their updates (the loads of 8
// vectors for all U entries issued together), stores, and products.
// hold (LDS, [NVMAX][kBlock], one entry per call): the entry's V values are
// kept there between the update and the products instead of being read again
// from memory (the same values).
template <class TV, int NVMAX, bool NORM, int U, bool LHOLD = false>
__device__ __forceinline__ void cgs_entries(size_t o, size_t i, size_t stride, int nv, const TV *__restrict__ V,
                                            size_t ld, const double *h, TV *__restrict__ w, double *acc,
                                            TV *hold = nullptr) {
    static_assert(!LHOLD || (U == 1 && !NORM), "hold: one entry, dots pass");
    constexpr int NC = (NVMAX + 7) / 8;
    double wi[U];
#pragma unroll
    for (int u = 0; u < U; u++) wi[u] = (double)w[o + i + u * stride];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (8 * c < nv) {
            TV vv[U][8];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (8 * c + kk < nv) vv[u][kk] = V[ld * (8 * c + kk) + o + i + u * stride];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (8 * c + kk < nv) wi[u] = wi[u] - h[8 * c + kk] * (double)vv[u][kk];
            if (LHOLD) {
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (8 * c + kk < nv) hold[(8 * c + kk) * kBlock + threadIdx.x] = vv[0][kk];
            }
        }
    }
    double wd[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const TV wt = (TV)wi[u];
        w[o + i + u * stride] = wt;
        wd[u] = (double)wt;
    }
    if (NORM) {
#pragma unroll
        for (int u = 0; u < U; u++) acc[0] += wd[u] * wd[u];
        return;
    }
    // read V again rather than hold its values across the update
    asm volatile("" ::: "memory");
#pragma unroll
    for (int c = 0; c < NC; c++) {
        if (8 * c < nv) {
            TV vv[U][8];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (8 * c + kk < nv)
                        vv[u][kk] = LHOLD ? hold[(8 * c + kk) * kBlock + threadIdx.x]
                                          : V[ld * (8 * c + kk) + o + i + u * stride];
#pragma unroll
            for (int u = 0; u < U; u++)
#pragma unroll
                for (int kk = 0; kk < 8; kk++)
                    if (8 * c + kk < nv) acc[8 * c + kk] += (double)vv[u][kk] * wd[u];
        }
    }
}

// One fused CGS pass over this thread's strided entries of [o + i0, o + n)
// (the entries block_dots visits, in its order): each entry first loses its
// projections on V_0..V_{nv-1} (w_i -= h_k V_ki for k ascending, as CGS in
// the oracle, rounded once to TV) and is stored; then its products with
// V_0..V_{nv-1} (or, NORM, its square) are added to the partial sums, which
// go through block_dots' block tree into part[0..nv) (part[0]).  The results
// are those of the update over all entries followed by block_dots, bit for
// bit, with one pass over w instead of three and the second read of each
// V_ki served by the cache.  DEEP (threads with many entries to walk):
// entries go four at a time (two for long bases) while four remain, so their
// loads overlap; it costs registers, so short walks take one entry at a time.
// h: nv coefficients in LDS; nv <= NVMAX.
template <class TV, int NVMAX, bool NORM, bool DEEP, bool LHOLD = false>
__device__ __forceinline__ void block_cgs_pass(size_t o, size_t i0, size_t n, size_t stride, int nv,
                                               const TV *__restrict__ V, size_t ld, const double *h,
                                               TV *__restrict__ w, double *__restrict__ part) {
    static_assert(!LHOLD || (!DEEP && !NORM), "hold: one entry at a time, dots pass");
    constexpr int NA = NORM ? 1 : NVMAX;
    constexpr int U = !DEEP ? 1 : NVMAX > 24 ? 2 : 4;
    __shared__ double wsum[NA][kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[NA];
#pragma unroll
    for (int k = 0; k < NA; k++) acc[k] = 0.0;
    size_t i = i0;
    if (U > 1)
        for (; i + (U - 1) * stride < n; i += U * stride)
            cgs_entries<TV, NVMAX, NORM, U>(o, i, stride, nv, V, ld, h, w, acc);
    if (LHOLD) {
        __shared__ TV hold[LHOLD ? NVMAX : 1][kBlock];
        for (; i < n; i += stride) cgs_entries<TV, NVMAX, NORM, 1, LHOLD>(o, i, stride, nv, V, ld, h, w, acc, &hold[0][0]);
    } else {
        for (; i < n; i += stride) cgs_entries<TV, NVMAX, NORM, 1>(o, i, stride, nv, V, ld, h, w, acc);
    }
    const int nd = NORM ? 1 : nv;
#pragma unroll
    for (int k = 0; k < NA; k++) {
        if (k < nd) {
            double v = acc[k];
#pragma unroll
            for (int sh = 32; sh > 0; sh >>= 1) v += __shfl_xor(v, sh, 64);
            if (lane == 0) wsum[k][wave] = v;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nd; k += kBlock) part[k] = ((wsum[k][0] + wsum[k][1]) + wsum[k][2]) + wsum[k][3];
}

// w -= V h, then the dots of w with V_0..V_{nv-1} (DOTS) or with itself
// (NORM) into part[block][..] (grid: dot_parts(n) blocks, as k_mdot).
template <class TV, int NVMAX, bool NORM, bool DEEP>
__global__ __launch_bounds__(kBlock) void k_cgs_pass(size_t n, int nv, const TV *__restrict__ V, size_t ld,
                                                     const double *__restrict__ hdev, TV *__restrict__ w,
                                                     double *__restrict__ part) {
    __shared__ double h[kMaxKrylov + 1];
    for (int k = threadIdx.x; k < nv; k += kBlock) h[k] = hdev[k];
    __syncthreads();
    block_cgs_pass<TV, NVMAX, NORM, DEEP>(0, blockIdx.x * (size_t)kBlock + threadIdx.x, n, (size_t)gridDim.x * kBlock, nv,
                                    V, ld, h, w, part + (size_t)blockIdx.x * (kMaxKrylov + 1));
}

// The same over a union of cases: block (c, b) of the segment map updates
// and reduces case c's range with case c's coefficients out[c][slot..].
template <class TV, int NVMAX, bool NORM>
__global__ __launch_bounds__(kBlock) void k_cgs_pass_seg(const int2 *__restrict__ blk, const int *__restrict__ first,
                                                         const int *__restrict__ off, const int *__restrict__ len,
                                                         int nv, const TV *__restrict__ V, size_t ld,
                                                         const double *__restrict__ out, int slot, TV *__restrict__ w,
                                                         double *__restrict__ part) {
    const int2 cb = blk[blockIdx.x];
    __shared__ double h[kMaxKrylov + 1];
    for (int k = threadIdx.x; k < nv; k += kBlock) h[k] = out[(size_t)cb.x * kBOut + slot + k];
    __syncthreads();
    const size_t P = (size_t)(first[cb.x + 1] - first[cb.x]);
    // (the FP32 dots pass with up to 41 vectors keeps each entry's V values
    // in LDS: a case's blocks give each thread one or a few entries)
    constexpr bool hold = !NORM && sizeof(TV) == 4 && NVMAX <= 41;
    block_cgs_pass<TV, NVMAX, NORM, false, hold>((size_t)off[cb.x], cb.y * (size_t)kBlock + threadIdx.x,
                                                 (size_t)len[cb.x], P * kBlock, nv, V, ld, h, w,
                                                 part + (size_t)blockIdx.x * (kMaxKrylov + 1));
}

// dst = src * (1/scale) with scale read from the device (scale = out[idx])
template <class TS, class TD>
__global__ __launch_bounds__(kBlock) void k_scale(size_t n, const TS *src, TD *dst, double denom) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        dst[i] = (TD)((double)src[i] / denom);
}

// dst = src / sqrt(*norm2): the Arnoldi normalisation with the norm read from
// device memory (the host evaluates the same correctly rounded sqrt for its
// Givens rotations), so the step can be enqueued without a host round trip.
template <class TS, class TD>
__global__ __launch_bounds__(kBlock) void k_scale_dev(size_t n, const TS *src, TD *dst, const double *norm2) {
    const double denom = sqrt(norm2[0]);
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        dst[i] = (TD)((double)src[i] / denom);
}

// x += sum_k y_k Z_k
template <class TV>
__global__ __launch_bounds__(kBlock) void k_xupdate(size_t n, int nv, const TV *__restrict__ Z, size_t ld,
                                                    const double *__restrict__ y, double *__restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        double xi = x[i];
        for (int k = 0; k < nv; k++) xi = xi + y[k] * (double)Z[ld * k + i];
        x[i] = xi;
    }
}

__global__ __launch_bounds__(kBlock) void k_vsum_partial(size_t n, const double *v, double *part) {
    __shared__ double sh[kBlock];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * (kMaxKrylov + 1)] = sh[0];
}

__global__ void k_fill(size_t n, double v, double *x) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) x[i] = v;
}

// ---- union of independent cases (batched solves) ------------------------
// Per-case dot products: block (c, b) of the segment map reduces over case
// c's range with the stride of an nb[c]-block launch on that case alone, so
// each case's dots are bit-identical to k_mdot on the case by itself.
template <class TV>
__global__ __launch_bounds__(kBlock) void k_mdot_seg(const int2 *__restrict__ blk, const int *__restrict__ first,
                                                     const int *__restrict__ off, const int *__restrict__ len, int nv,
                                                     const TV *__restrict__ V, size_t ld, const TV *__restrict__ w,
                                                     double *__restrict__ part) {
    const int2 cb = blk[blockIdx.x];
    const size_t P = (size_t)(first[cb.x + 1] - first[cb.x]);
    block_dots<TV>((size_t)off[cb.x], cb.y * (size_t)kBlock + threadIdx.x, (size_t)len[cb.x], P * kBlock, nv, V, ld,
                   w, part + (size_t)blockIdx.x * (kMaxKrylov + 1));
}

// out[c][slot + k] = sum of case c's partials, combined as k_mdot_final
// (grid: dots x cases)
__global__ __launch_bounds__(64) void k_mdot_seg_final(const int *first, int nv, const double *part, double *out,
                                                       int slot) {
    const int k = blockIdx.x, c = blockIdx.y;
    const double s = wave_sum_parts(part, first[c], first[c + 1] - first[c], k);
    if (threadIdx.x == 0) out[(size_t)c * kBOut + slot + k] = s;
}

// dst = src / sqrt(out[c][slot]) (Arnoldi normalisation, per case)
template <class TS, class TD>
__global__ __launch_bounds__(kBlock) void k_scale_dev_seg(size_t n, const int *__restrict__ ecase, const TS *src,
                                                          TD *dst, const double *__restrict__ out, int slot) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double denom = sqrt(out[(size_t)ecase[i / 5] * kBOut + slot]);
        dst[i] = (TD)((double)src[i] / denom);
    }
}

// dst = src / den[c] (first Krylov vector, per case)
template <class TS, class TD>
__global__ __launch_bounds__(kBlock) void k_scale_seg(size_t n, const int *__restrict__ ecase, const TS *src, TD *dst,
                                                      const double *__restrict__ den) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        dst[i] = (TD)((double)src[i] / den[(size_t)ecase[i / 5] * (kMaxKrylov + 1)]);
}

// x += sum_{k < kdim_c} y_{c,k} Z_k
template <class TV>
__global__ __launch_bounds__(kBlock) void k_xupdate_seg(size_t n, const int *__restrict__ ecase,
                                                        const TV *__restrict__ Z, size_t ld,
                                                        const double *__restrict__ y, const int *__restrict__ kdim,
                                                        double *__restrict__ x) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int c = ecase[i / 5];
        const double *yc = y + (size_t)c * (kMaxKrylov + 1);
        double xi = x[i];
        for (int k = 0; k < kdim[c]; k++) xi = xi + yc[k] * (double)Z[ld * k + i];
        x[i] = xi;
    }
}

int grid_cap(size_t n, int cap) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return (int)(g < (size_t)cap ? g : cap);
}

}  // namespace

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
int launch_jacobian(cfd_mesh *m, const cfd_params *p, int rhs_mode, int mixed) {
    const int nb = xcd_grid((m->N + 127) / 128);
    if (mixed) {
        JacArgs<float> a{m->N, m->NT, m->order, m->inc_ptr, m->sptr, m->nbr, m->soff, m->iau, m->d.isx, m->d.isy,
                         m->d.isz, m->d.qa, m->d.vol, m->dt, m->d.res, m->af, m->b, p->gamma, p->entropy_fix,
                         rhs_mode};
        hipLaunchKernelGGL(k_jacobian<float>, dim3(nb), dim3(128), 0, m->stream, a);
    } else {
        JacArgs<double> a{m->N, m->NT, m->order, m->inc_ptr, m->sptr, m->nbr, m->soff, m->iau, m->d.isx, m->d.isy,
                          m->d.isz, m->d.qa, m->d.vol, m->dt, m->d.res, m->a, m->b, p->gamma, p->entropy_fix,
                          rhs_mode};
        hipLaunchKernelGGL(k_jacobian<double>, dim3(nb), dim3(128), 0, m->stream, a);
    }
    CFD_HIP(m, hipGetLastError());
    m->jac_valid = true;
    m->jac_mixed = mixed != 0;
    m->prec_valid = false;
    m->asell_valid = false;
    return CFD_OK;
}

int launch_jacobian_fd(cfd_mesh *m, const cfd_params *p, int method, int iter, int add_time, int rhs_mode,
                       int mixed) {
    int fdm = method;
    if (fdm == CFD_JAC_FD_ALTERNATE) fdm = (iter % 2 == 0) ? CFD_JAC_FD_FORWARD : CFD_JAC_FD_BACKWARD;
    FdArgs a{};
    a.N = m->N;
    a.NT = m->NT;
    a.order = m->order;
    a.ip = m->inc_ptr;
    a.sptr = m->sptr;
    a.nbr = m->nbr;
    a.soff = m->soff;
    a.iau = m->iau;
    a.isx = m->d.isx;
    a.isy = m->d.isy;
    a.isz = m->d.isz;
    a.qa = m->d.qa;
    a.qb = m->d.qb;
    a.vol = m->d.vol;
    a.dt = m->dt;
    a.res = m->d.res;
    a.btype = m->btype;
    a.bcase = m->bcase;
    a.case_inf = m->case_inf;
    a.a = mixed ? nullptr : m->a;
    a.af = mixed ? m->af : nullptr;
    a.b = m->b;
    a.g = p->gamma;
    a.frac = p->limiter_order == 2 ? 0.5 : 0.0;
    a.coeff = fdm == CFD_JAC_FD_CENTRAL ? 0.5 : 1.0;
    a.efix = p->entropy_fix;
    a.rhs_mode = rhs_mode;
    a.add_time = add_time;
    a.order2 = p->order == 2;
    a.lim = p->limiter != CFD_LIM_NONE;
    a.plus = fdm == CFD_JAC_FD_CENTRAL || fdm == CFD_JAC_FD_FORWARD;
    a.minus = fdm == CFD_JAC_FD_CENTRAL || fdm == CFD_JAC_FD_BACKWARD;
    a.iter = iter;
    a.zpg = p->zpg_iter;
    a.free = BcFree<double>{p->inf_rho, p->inf_u, p->inf_v, p->inf_w, p->inf_p, p->outflow_p};
    a.flag = m->flag;
    const int nb = (int)((10 * (size_t)m->N + 127) / 128);
    if (p->flux == CFD_FLUX_HLLC)
        hipLaunchKernelGGL(k_jacobian_fd<CFD_FLUX_HLLC>, dim3(nb), dim3(128), 0, m->stream, a);
    else if (p->flux == CFD_FLUX_AUSMP)
        hipLaunchKernelGGL(k_jacobian_fd<CFD_FLUX_AUSMP>, dim3(nb), dim3(128), 0, m->stream, a);
    else
        hipLaunchKernelGGL(k_jacobian_fd<CFD_FLUX_ROE>, dim3(nb), dim3(128), 0, m->stream, a);
    CFD_HIP(m, hipGetLastError());
    m->jac_valid = true;
    m->jac_mixed = mixed != 0;
    m->prec_valid = false;
    m->asell_valid = false;
    return CFD_OK;
}

namespace {

// Block-SELL value arrays by storage type.
template <class TM>
struct SellVals;
template <>
struct SellVals<double> {
    static double *&a(cfd_mesh *m) { return m->asv; }
    static double *&l(cfd_mesh *m) { return m->lsv; }
    static double *&u(cfd_mesh *m) { return m->usv; }
    static double *&d(cfd_mesh *m) { return m->dsv; }
};
template <>
struct SellVals<float> {
    static float *&a(cfd_mesh *m) { return m->asvf; }
    static float *&l(cfd_mesh *m) { return m->lsvf; }
    static float *&u(cfd_mesh *m) { return m->usvf; }
    static float *&d(cfd_mesh *m) { return m->dsvf; }
};
// The BSR Jacobian by storage type.
template <class TM>
const TM *bsr_a(cfd_mesh *m);
template <>
const double *bsr_a<double>(cfd_mesh *m) { return m->a; }
template <>
const float *bsr_a<float>(cfd_mesh *m) { return m->af; }

template <class TM>
int build_prec(cfd_mesh *m, int kind, const TM *a, TM *lu, TM *dinv, bool half) {
    if (kind == CFD_PREC_ILU0_MC) {
        hipLaunchKernelGGL((k_gather_blocks<TM, TM>), dim3(grid_cap(25 * (size_t)m->nnz, 8192)), dim3(kBlock), 0,
                           m->stream, (size_t)m->nnz, m->mc_src, a, lu);
        for (int l = 0; l < m->mc_lower.nlevels; l++) {
            const int b0 = m->mc_lower.ptr[l], nr = m->mc_lower.ptr[l + 1] - b0;
            hipLaunchKernelGGL(k_ilu0_level<TM>, dim3((nr + 47) / 48), dim3(kBlock), 0, m->stream, nr,
                               m->mc_lower.rows + b0, m->mc_ia, m->mc_ja, m->mc_iau, m->mc_moff, m->mc_mqq, lu, dinv,
                               lu, dinv, m->flag);
        }
        // block-SELL copies for the sweeps, FP16 with mixed = 2
        if constexpr (std::is_same<TM, float>::value) {
            if (half) {
                hipLaunchKernelGGL(k_to_sell_half, dim3(grid_cap((size_t)m->mcl_sell.nblk, 8192)), dim3(kBlock), 0,
                                   m->stream, (size_t)m->mcl_sell.nblk, m->mcl_sell.src, lu, m->lsvh, m->lsvs);
                hipLaunchKernelGGL(k_to_sell_half, dim3(grid_cap((size_t)m->mcu_sell.nblk, 8192)), dim3(kBlock), 0,
                                   m->stream, (size_t)m->mcu_sell.nblk, m->mcu_sell.src, lu, m->usvh, m->usvs);
                hipLaunchKernelGGL(k_to_sell_half, dim3(grid_cap((size_t)m->N, 8192)), dim3(kBlock), 0, m->stream,
                                   (size_t)m->N, nullptr, dinv, m->dsvh, m->dsvs);
            }
        }
        if (m->use_sell && !half) {
            hipLaunchKernelGGL(k_to_sell<TM>, dim3(grid_cap((size_t)m->mcl_sell.nblk, 8192)), dim3(kBlock), 0, m->stream,
                               (size_t)m->mcl_sell.nblk, m->mcl_sell.src, lu, SellVals<TM>::l(m));
            hipLaunchKernelGGL(k_to_sell<TM>, dim3(grid_cap((size_t)m->mcu_sell.nblk, 8192)), dim3(kBlock), 0, m->stream,
                               (size_t)m->mcu_sell.nblk, m->mcu_sell.src, lu, SellVals<TM>::u(m));
            hipLaunchKernelGGL(k_to_sell<TM>, dim3(grid_cap((size_t)m->N, 8192)), dim3(kBlock), 0, m->stream,
                               (size_t)m->N, nullptr, dinv, SellVals<TM>::d(m));
        }
    } else if (kind == CFD_PREC_ILU0) {
        const int rc = cfd_ensure_ilu_map(m);
        if (rc) return rc;
        CFD_HIP(m, hipMemcpyAsync(lu, a, sizeof(TM) * 25 * (size_t)m->nnz, hipMemcpyDeviceToDevice, m->stream));
        for (int l = 0; l < m->lower.nlevels; l++) {
            const int b0 = m->lower.ptr[l], nr = m->lower.ptr[l + 1] - b0;
            hipLaunchKernelGGL(k_ilu0_level<TM>, dim3((nr + 47) / 48), dim3(kBlock), 0, m->stream, nr,
                               m->lower.rows + b0, m->ia, m->ja, m->iau, m->ilu_moff, m->ilu_mqq, lu, dinv, lu, dinv,
                               m->flag);
        }
    } else if (kind == CFD_PREC_BJACOBI) {
        hipLaunchKernelGGL((k_block_inv<TM, TM>), dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream,
                           m->N, m->iau, a, dinv, m->flag);
    }
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

// y = A x, or y = bsub - A x, with the block-SELL copy of the Jacobian.
template <class TM, class TX, class TY>
void spmv(cfd_mesh *m, const TX *x, TY *y, const double *bsub) {
    if (std::is_same<TM, float>::value && m->use_sell && m->asell_half)
        hipLaunchKernelGGL((k_spmv_sell_half<TX, TY>), dim3(xcd_grid((m->N + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           m->stream, m->N, m->ia, m->a_sell.off, m->a_sell.col, m->asvh, m->asvs, x, y, bsub);
    else if (m->use_sell)
        hipLaunchKernelGGL((k_spmv_sell<TM, TX, TY>), dim3(xcd_grid((m->N + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           m->stream,
                           m->N, m->ia, m->a_sell.off, m->a_sell.col, (const TM *)SellVals<TM>::a(m), x, y, bsub);
    else
        hipLaunchKernelGGL((k_spmv<TM, TX, TY>), dim3((int)((5 * (size_t)m->N + kBlock - 1) / kBlock)), dim3(kBlock),
                           0, m->stream, m->N, m->ia, m->ja, bsr_a<TM>(m), x, y, bsub);
}

template <class TV>
TV *mc_work(cfd_mesh *m);
template <>
double *mc_work<double>(cfd_mesh *m) { return m->zp; }
template <>
float *mc_work<float>(cfd_mesh *m) { return m->zpf; }

template <class TM, class TV>
void apply_prec(cfd_mesh *m, int kind, const TM *lu, const TM *dinv, const TV *r, TV *z) {
    const size_t L = 5 * (size_t)m->N;
    if (kind == CFD_PREC_ILU0_MC && !m->use_sell) {
        TV *wk = mc_work<TV>(m);
        for (int pass = 0; pass < 2; pass++) {
            const LevelSchedule &ls = pass ? m->mc_upper : m->mc_lower;
            for (int l = 0; l < ls.nlevels; l++) {
                const int b0 = ls.ptr[l], nr = ls.ptr[l + 1] - b0;
                const int nwaves = (nr + 11) / 12;
                hipLaunchKernelGGL((k_trsv_mc<TM, TV>), dim3((nwaves + 3) / 4), dim3(kBlock), 0, m->stream, nr,
                                   ls.rows + b0, m->mc_ia, m->mc_ja, m->mc_iau, m->mc_perm, lu, dinv, r, wk, z, pass);
            }
        }
    } else if (kind == CFD_PREC_ILU0_MC && m->prec_half) {
        if constexpr (std::is_same<TM, float>::value && std::is_same<TV, float>::value) {
            float *wk = mc_work<float>(m);
            for (int pass = 0; pass < 2; pass++) {
                const BlockSell &bs = pass ? m->mcu_sell : m->mcl_sell;
                const _Float16 *val = pass ? m->usvh : m->lsvh;
                const float *vsc = pass ? m->usvs : m->lsvs;
                for (int t = 0; t < m->ncolor; t++) {
                    const int c = pass ? m->ncolor - 1 - t : t;
                    const int r0 = m->mc_cstart[c], r1 = m->mc_cstart[c + 1];
                    const int g = xcd_grid((r1 - (r0 & ~63) + kBlock - 1) / kBlock);
                    hipLaunchKernelGGL(k_trsv_mc_sell_half, dim3(g), dim3(kBlock), 0, m->stream, r0, r1, m->mc_ia,
                                       m->mc_iau, m->mc_perm, bs.off, bs.col, val, vsc, m->dsvh, m->dsvs, r, wk, z,
                                       pass);
                }
            }
        }
    } else if (kind == CFD_PREC_ILU0_MC) {
        // one launch per colour: forward in colour order, backward in reverse
        TV *wk = mc_work<TV>(m);
        const TM *dv = SellVals<TM>::d(m);
        for (int pass = 0; pass < 2; pass++) {
            const BlockSell &bs = pass ? m->mcu_sell : m->mcl_sell;
            const TM *val = pass ? SellVals<TM>::u(m) : SellVals<TM>::l(m);
            for (int t = 0; t < m->ncolor; t++) {
                const int c = pass ? m->ncolor - 1 - t : t;
                const int r0 = m->mc_cstart[c], r1 = m->mc_cstart[c + 1];
                const int g = xcd_grid((r1 - (r0 & ~63) + kBlock - 1) / kBlock);
                hipLaunchKernelGGL((k_trsv_mc_sell<TM, TV>), dim3(g), dim3(kBlock), 0, m->stream, r0, r1, m->mc_ia,
                                   m->mc_iau, m->mc_perm, bs.off, bs.col, val, dv, r, wk, z, pass);
            }
        }
    } else if (kind == CFD_PREC_ILU0 && m->sync_free) {
        for (int pass = 0; pass < 2; pass++) {
            const LevelSchedule &ls = pass ? m->upper : m->lower;
            hipLaunchKernelGGL(k_epoch_bump, dim3(1), dim3(64), 0, m->stream, m->epoch);
            hipLaunchKernelGGL((k_trsv_sf<TM, TV>), dim3((ls.n12 + 47) / 48), dim3(kBlock), 0, m->stream, ls.n12,
                               ls.slots12, m->ia, m->ja, m->iau, lu, dinv, r, z, pass, m->ready, m->epoch, m->flag);
        }
    } else if (kind == CFD_PREC_ILU0) {
        for (int pass = 0; pass < 2; pass++) {
            const LevelSchedule &ls = pass ? m->upper : m->lower;
            for (int l = 0; l < ls.nlevels; l++) {
                const int b0 = ls.ptr[l], nr = ls.ptr[l + 1] - b0;
                const int nwaves = (nr + 11) / 12;
                hipLaunchKernelGGL((k_trsv_level<TM, TV>), dim3((nwaves + 3) / 4), dim3(kBlock), 0, m->stream, nr,
                                   ls.rows + b0, m->ia, m->ja, m->iau, lu, dinv, r, z, pass);
            }
        }
    } else if (kind == CFD_PREC_BJACOBI) {
        hipLaunchKernelGGL((k_bjac_apply<TM, TV>), dim3((int)((L + kBlock - 1) / kBlock)), dim3(kBlock), 0, m->stream,
                           m->N, dinv, r, z);
    } else {
        hipMemcpyAsync(z, r, sizeof(TV) * L, hipMemcpyDeviceToDevice, m->stream);
    }
}

// dots of w with V_0..V_{nv-1}; results in m->redout + off (device)
template <class TV>
void mdot(cfd_mesh *m, int nv, const TV *V, size_t ld, const TV *w, int off) {
    const size_t L = 5 * (size_t)m->N;
    const int np = dot_parts(L);
    if (L >= 4 * (size_t)np * kBlock)  // each thread walks four strides or more
        hipLaunchKernelGGL((k_mdot<TV, true>), dim3(np), dim3(kBlock), 0, m->stream, L, nv, V, ld, w, m->red);
    else
        hipLaunchKernelGGL((k_mdot<TV, false>), dim3(np), dim3(kBlock), 0, m->stream, L, nv, V, ld, w, m->red);
    hipLaunchKernelGGL(k_mdot_final, dim3(nv), dim3(64), 0, m->stream, np, nv, m->red, m->redout + off);
}

// w -= V_{0..nv-1} h (h = m->redout + hoff on the device), then the dots of
// the new w with V_0..V_{nv-1} (or, NORM, its squared norm) into
// m->redout + off: the update and mdot in one pass (the same results).
template <class TV, bool NORM>
void cgs_pass(cfd_mesh *m, int nv, const TV *V, size_t ld, int hoff, TV *w, int off) {
    const size_t L = 5 * (size_t)m->N;
    const int np = dot_parts(L);
    const double *h = m->redout + hoff;
    const bool deep = L >= 4 * (size_t)np * kBlock;  // each thread walks four strides or more
    dim3 g(np), b(kBlock);
    // NVMAX: the fewest unrolled vectors that cover nv
#define CFD_CGS(NVMAX)                                                                                          \
    do {                                                                                                         \
        if (deep) hipLaunchKernelGGL((k_cgs_pass<TV, NVMAX, NORM, true>), g, b, 0, m->stream, L, nv, V, ld, h, w, \
                                     m->red);                                                                    \
        else hipLaunchKernelGGL((k_cgs_pass<TV, NVMAX, NORM, false>), g, b, 0, m->stream, L, nv, V, ld, h, w,     \
                                m->red);                                                                         \
    } while (0)
    if (nv <= 8) CFD_CGS(8);
    else if (nv <= 16) CFD_CGS(16);
    else if (nv <= 32) CFD_CGS(32);
    else if (nv <= 41) CFD_CGS(41);
    else CFD_CGS(kMaxKrylov + 1);
#undef CFD_CGS
    const int nd = NORM ? 1 : nv;
    hipLaunchKernelGGL(k_mdot_final, dim3(nd), dim3(64), 0, m->stream, np, nd, m->red, m->redout + off);
}

// Krylov basis of one storage type for restart length mk.  Each precision
// keeps its own basis, so alternating FP64 and mixed solves on one mesh
// allocates each once; a basis too short for a larger mk is replaced (the
// old one freed, after the stream is idle and the graphs captured on it are
// dropped), so the mesh's footprint never grows past one basis per
// precision.
int ensure_krylov(cfd_mesh *m, int mk, bool mixed) {
    const size_t L = 5 * (size_t)m->N;
    if (mixed && !m->zpf) m->zpf = (float *)cfd_dalloc(m, sizeof(float) * L);
    if (!mixed && !m->zp) m->zp = (double *)cfd_dalloc(m, sizeof(double) * L);
    if ((mixed && !m->zpf) || (!mixed && !m->zp))
        return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (multicolour work vector)");
    if (!m->r) m->r = (double *)cfd_dalloc(m, sizeof(double) * L);
    if (!m->r) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (Krylov)");
    int &have = mixed ? m->kry_m32 : m->kry_m64;
    if (have >= mk) return CFD_OK;
    if (have > 0) {
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        drop_krylov_graphs(m);
    }
    if (mixed) {
        cfd_dfree(m, m->Vf);
        cfd_dfree(m, m->Zf);
        cfd_dfree(m, m->wf);
        m->Vf = (float *)cfd_dalloc(m, sizeof(float) * L * (mk + 1));
        m->Zf = (float *)cfd_dalloc(m, sizeof(float) * L * mk);
        m->wf = (float *)cfd_dalloc(m, sizeof(float) * L);
        have = 0;
        if (!m->Vf || !m->Zf || !m->wf) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (Krylov FP32)");
    } else {
        cfd_dfree(m, m->V);
        cfd_dfree(m, m->Z);
        cfd_dfree(m, m->w);
        m->V = (double *)cfd_dalloc(m, sizeof(double) * L * (mk + 1));
        m->Z = (double *)cfd_dalloc(m, sizeof(double) * L * mk);
        m->w = (double *)cfd_dalloc(m, sizeof(double) * L);
        have = 0;
        if (!m->V || !m->Z || !m->w) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (Krylov)");
    }
    have = mk;
    return CFD_OK;
}

// One Arnoldi step j of FGMRES: V_j = w / ||w|| (j > 0, norm of the previous
// step), Z_j = M^-1 V_j, w = A Z_j, two CGS passes against V_0..V_j, ||w||^2,
// and the copy of the Hessenberg column to the host.  The sequence depends
// only on j and the mesh's buffers, so it is captured once per j as a graph.
template <class TM, class TV>
void arnoldi_step(cfd_mesh *m, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z, TV *w, int j) {
    const size_t L = 5 * (size_t)m->N;
    const int gN = grid_cap(L, 4096);
    const int gS = (int)((L + kBlock - 1) / kBlock);
    TV *vj = V + L * j, *zj = Z + L * j;
    if (j > 0)
        hipLaunchKernelGGL((k_scale_dev<TV, TV>), dim3(gN), dim3(kBlock), 0, m->stream, L, w, vj,
                           m->redout + 2 * (kMaxKrylov + 1));
    apply_prec<TM, TV>(m, m->prec_kind, lu, dinv, vj, zj);
    spmv<TM, TV, TV>(m, zj, w, nullptr);
    // CGS2: h1 = V^T w, w -= V h1, h2 = V^T w, w -= V h2, ||w||^2 (each
    // update fused with the reduction after it)
    mdot<TV>(m, j + 1, V, L, w, 0);
    cgs_pass<TV, false>(m, j + 1, V, L, 0, w, kMaxKrylov + 1);
    cgs_pass<TV, true>(m, j + 1, V, L, kMaxKrylov + 1, w, 2 * (kMaxKrylov + 1));
    hipMemcpyAsync(m->hred, m->redout, sizeof(double) * (2 * (kMaxKrylov + 1) + 1), hipMemcpyDeviceToHost, m->stream);
}

// Enqueues Arnoldi step j, replaying its captured graph (captured on first use).
template <class TM, class TV>
int enqueue_step(cfd_mesh *m, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z, TV *w, int j, int mk) {
    if (!m->use_graphs) {
        arnoldi_step<TM, TV>(m, a, lu, dinv, V, Z, w, j);
        CFD_HIP(m, hipGetLastError());
        return CFD_OK;
    }
    KrylovGraphs &kg = m->kgraph;
    if (kg.key != (const void *)V || kg.prec != m->prec_kind || kg.m != mk || kg.lu != (const void *)lu) {
        drop_krylov_graphs(m);
        kg.key = V;
        kg.lu = lu;
        kg.prec = m->prec_kind;
        kg.m = mk;
        kg.exec.assign(mk, nullptr);
    }
    if (!kg.exec[j]) {
        // Relaxed mode: other threads driving other meshes may call
        // allocation or synchronous copy functions meanwhile.  If the capture
        // still fails, this mesh falls back to direct launches of the same
        // sequence.
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        hipError_t e;
        {
            std::unique_lock<std::shared_mutex> lk(cfd_capture_mutex());
            e = hipStreamBeginCapture(m->stream, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                arnoldi_step<TM, TV>(m, a, lu, dinv, V, Z, w, j);
                e = hipStreamEndCapture(m->stream, &g);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (g) hipGraphDestroy(g);
        }
        if (e != hipSuccess || !ex) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(m->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(m->stream, &g2);
                if (g2) hipGraphDestroy(g2);
            }
            (void)hipGetLastError();
            m->use_graphs = false;
            drop_krylov_graphs(m);
            arnoldi_step<TM, TV>(m, a, lu, dinv, V, Z, w, j);
            CFD_HIP(m, hipGetLastError());
            return CFD_OK;
        }
        kg.exec[j] = ex;
    }
    CFD_HIP(m, hipGraphLaunch((hipGraphExec_t)kg.exec[j], m->stream));
    return CFD_OK;
}

// Per-case dots of V_0..V_{nv-1} with w into m->bred[c][slot..] (union of cases).
template <class TV>
void mdot_seg(cfd_mesh *m, int nv, const TV *V, size_t ld, const TV *w, int slot) {
    const SegMap &sg = m->seg_dot;
    hipLaunchKernelGGL(k_mdot_seg<TV>, dim3(sg.nblocks), dim3(kBlock), 0, m->stream, sg.blk, sg.first, sg.off, sg.len,
                       nv, V, ld, w, m->bpart);
    hipLaunchKernelGGL(k_mdot_seg_final, dim3(nv, m->ncase), dim3(64), 0, m->stream, sg.first, nv, m->bpart, m->bred,
                       slot);
}

// Per case: w -= V h_c (h_c = m->bred[c][hslot..]), then the dots of the new
// w with V_0..V_{nv-1} (or, NORM, its squared norm) into m->bred[c][slot..].
template <class TV, bool NORM>
void cgs_pass_seg(cfd_mesh *m, int nv, const TV *V, size_t ld, int hslot, TV *w, int slot) {
    const SegMap &sg = m->seg_dot;
    dim3 g(sg.nblocks), b(kBlock);
#define CFD_CGS_SEG(NVMAX, NORM_)                                                                                   \
    hipLaunchKernelGGL((k_cgs_pass_seg<TV, NVMAX, NORM_>), g, b, 0, m->stream, sg.blk, sg.first, sg.off, sg.len, nv, \
                       V, ld, m->bred, hslot, w, m->bpart)
    if (nv <= 8) CFD_CGS_SEG(8, NORM);
    else if (nv <= 16) CFD_CGS_SEG(16, NORM);
    else if (nv <= 32) CFD_CGS_SEG(32, NORM);
    else if (nv <= 41) CFD_CGS_SEG(41, NORM);
    else CFD_CGS_SEG(kMaxKrylov + 1, NORM);
#undef CFD_CGS_SEG
    const int nd = NORM ? 1 : nv;
    hipLaunchKernelGGL(k_mdot_seg_final, dim3(nd, m->ncase), dim3(64), 0, m->stream, sg.first, nd, m->bpart, m->bred,
                       slot);
}

// Arnoldi step j of the batched FGMRES: the operations of arnoldi_step with
// per-case dots, coefficients and norms.
template <class TM, class TV>
void arnoldi_step_batch(cfd_mesh *m, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z, TV *w, int j) {
    const size_t L = 5 * (size_t)m->N;
    const int gN = grid_cap(L, 4096);
    const int gS = (int)((L + kBlock - 1) / kBlock);
    TV *vj = V + L * j, *zj = Z + L * j;
    if (j > 0)
        hipLaunchKernelGGL((k_scale_dev_seg<TV, TV>), dim3(gN), dim3(kBlock), 0, m->stream, L, m->ecase, w, vj,
                           m->bred, 2 * (kMaxKrylov + 1));
    apply_prec<TM, TV>(m, m->prec_kind, lu, dinv, vj, zj);
    spmv<TM, TV, TV>(m, zj, w, nullptr);
    mdot_seg<TV>(m, j + 1, V, L, w, 0);
    cgs_pass_seg<TV, false>(m, j + 1, V, L, 0, w, kMaxKrylov + 1);
    cgs_pass_seg<TV, true>(m, j + 1, V, L, kMaxKrylov + 1, w, 2 * (kMaxKrylov + 1));
    hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * (size_t)m->ncase, hipMemcpyDeviceToHost, m->stream);
}

template <class TM, class TV>
int enqueue_step_batch(cfd_mesh *m, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z, TV *w, int j, int mk) {
    if (!m->use_graphs) {
        arnoldi_step_batch<TM, TV>(m, a, lu, dinv, V, Z, w, j);
        CFD_HIP(m, hipGetLastError());
        return CFD_OK;
    }
    KrylovGraphs &kg = m->kgraph_b;
    if (kg.key != (const void *)V || kg.prec != m->prec_kind || kg.m != mk || kg.lu != (const void *)lu) {
        CfdRuntimeGuard guard;
        for (auto &g : kg.exec)
            if (g) hipGraphExecDestroy((hipGraphExec_t)g);
        kg.key = V;
        kg.lu = lu;
        kg.prec = m->prec_kind;
        kg.m = mk;
        kg.exec.assign(mk, nullptr);
    }
    if (!kg.exec[j]) {
        // Relaxed mode: other threads driving other meshes may call
        // allocation or synchronous copy functions meanwhile.  If the capture
        // still fails, this mesh falls back to direct launches of the same
        // sequence.
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        hipError_t e;
        {
            std::unique_lock<std::shared_mutex> lk(cfd_capture_mutex());
            e = hipStreamBeginCapture(m->stream, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                arnoldi_step_batch<TM, TV>(m, a, lu, dinv, V, Z, w, j);
                e = hipStreamEndCapture(m->stream, &g);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (g) hipGraphDestroy(g);
        }
        if (e != hipSuccess || !ex) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(m->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(m->stream, &g2);
                if (g2) hipGraphDestroy(g2);
            }
            (void)hipGetLastError();
            m->use_graphs = false;
            drop_krylov_graphs(m);
            arnoldi_step_batch<TM, TV>(m, a, lu, dinv, V, Z, w, j);
            CFD_HIP(m, hipGetLastError());
            return CFD_OK;
        }
        kg.exec[j] = ex;
    }
    CFD_HIP(m, hipGraphLaunch((hipGraphExec_t)kg.exec[j], m->stream));
    return CFD_OK;
}

// Batched restarted FGMRES(m): every case of the union runs the algorithm of
// fgmres() on its own rows, in lockstep; a case that has converged (or is
// not being solved) keeps riding along but its results are not used.
template <class TM, class TV>
int fgmres_batch(cfd_mesh *m, const cfd_linsolve_opts *o, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z,
                 TV *w, const std::vector<char> &solve, std::vector<int> &iters, std::vector<double> &fres,
                 std::vector<int> &status) {
    const int K = m->ncase, mk = o->krylov_m;
    const size_t L = 5 * (size_t)m->N;
    const int gN = grid_cap(L, 4096);
    const int gS = (int)((L + kBlock - 1) / kBlock);
    const size_t ldc = kMaxKrylov + 1;
    hipLaunchKernelGGL(k_fill, dim3(gN), dim3(kBlock), 0, m->stream, L, 0.0, m->x);
    CFD_HIP(m, hipMemcpyAsync(m->r, m->b, sizeof(double) * L, hipMemcpyDeviceToDevice, m->stream));
    mdot_seg<double>(m, 1, m->r, 0, m->r, 0);
    CFD_HIP(m, hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * K, hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    std::vector<double> beta(K), tol(K), res(K);
    std::vector<char> live(K), cyc(K), step(K);
    std::vector<int> kdim(K);
    std::vector<double> H((size_t)K * (mk + 1) * mk, 0.0), cs((size_t)K * mk), sn((size_t)K * mk),
        gv((size_t)K * (mk + 1)), hcol(mk + 2), y((size_t)mk);
    iters.assign(K, 0);
    fres.assign(K, 0.0);
    status.assign(K, CFD_OK);
    for (int c = 0; c < K; c++) {
        beta[c] = sqrt(m->hbred[(size_t)c * kBOut]);
        tol[c] = o->rtol * beta[c];
        res[c] = beta[c];
        live[c] = solve[c] && beta[c] > tol[c] && beta[c] > 0.0;
    }
    for (int cycle = 0; cycle <= o->max_restarts; cycle++) {
        bool any = false;
        for (int c = 0; c < K; c++) {
            cyc[c] = step[c] = live[c];
            kdim[c] = 0;
            any = any || live[c];
            m->hbcoef[(size_t)c * ldc] = live[c] ? beta[c] : 1.0;
            for (int i = 0; i <= mk; i++) gv[(size_t)c * (mk + 1) + i] = 0.0;
            gv[(size_t)c * (mk + 1)] = beta[c];
        }
        if (!any) break;
        CFD_HIP(m, hipMemcpyAsync(m->bcoef, m->hbcoef, sizeof(double) * ldc * K, hipMemcpyHostToDevice, m->stream));
        hipLaunchKernelGGL((k_scale_seg<double, TV>), dim3(gN), dim3(kBlock), 0, m->stream, L, m->ecase, m->r, V,
                           m->bcoef);
        for (int j = 0; j < mk; j++) {
            int rc = enqueue_step_batch<TM, TV>(m, a, lu, dinv, V, Z, w, j, mk);
            if (rc) return rc;
            CFD_HIP(m, hipStreamSynchronize(m->stream));
            bool stepping = false;
            for (int c = 0; c < K; c++) {
                if (!step[c]) continue;
                const double *hr = m->hbred + (size_t)c * kBOut;
                double *Hc = H.data() + (size_t)c * (mk + 1) * mk, *csc = cs.data() + (size_t)c * mk,
                       *snc = sn.data() + (size_t)c * mk, *gvc = gv.data() + (size_t)c * (mk + 1);
                for (int k = 0; k <= j; k++) hcol[k] = 0.0 + hr[k];
                for (int k = 0; k <= j; k++) hcol[k] += hr[kMaxKrylov + 1 + k];
                const double hn = sqrt(hr[2 * (kMaxKrylov + 1)]);
                hcol[j + 1] = hn;
                for (int k = 0; k < j; k++) {
                    const double t = csc[k] * hcol[k] + snc[k] * hcol[k + 1];
                    hcol[k + 1] = -snc[k] * hcol[k] + csc[k] * hcol[k + 1];
                    hcol[k] = t;
                }
                const double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
                csc[j] = hcol[j] / den;
                snc[j] = hcol[j + 1] / den;
                hcol[j] = den;
                hcol[j + 1] = 0.0;
                gvc[j + 1] = -snc[j] * gvc[j];
                gvc[j] = csc[j] * gvc[j];
                for (int k = 0; k <= j; k++) Hc[(size_t)k * mk + j] = hcol[k];
                iters[c]++;
                kdim[c] = j + 1;
                res[c] = fabs(gvc[j + 1]);
                if (!(res[c] == res[c])) {
                    // this case fails on its own (as its solve alone would);
                    // its correction is dropped and the other cases go on
                    status[c] = CFD_E_NONFINITE;
                    step[c] = cyc[c] = live[c] = false;
                    continue;
                }
                if (res[c] <= tol[c] || hn == 0.0) step[c] = false;
                stepping = stepping || step[c];
            }
            if (!stepping) break;
        }
        for (int c = 0; c < K; c++) {
            double *yc = m->hbcoef + (size_t)c * ldc;
            for (int k = 0; k <= kMaxKrylov; k++) yc[k] = 0.0;
            m->hbkdim[c] = cyc[c] ? kdim[c] : 0;
            if (!cyc[c]) continue;
            const double *Hc = H.data() + (size_t)c * (mk + 1) * mk, *gvc = gv.data() + (size_t)c * (mk + 1);
            for (int k = kdim[c] - 1; k >= 0; k--) {
                double t = gvc[k];
                for (int q = k + 1; q < kdim[c]; q++) t -= Hc[(size_t)k * mk + q] * y[q];
                y[k] = t / Hc[(size_t)k * mk + k];
            }
            for (int k = 0; k < kdim[c]; k++) yc[k] = y[k];
        }
        CFD_HIP(m, hipMemcpyAsync(m->bcoef, m->hbcoef, sizeof(double) * ldc * K, hipMemcpyHostToDevice, m->stream));
        CFD_HIP(m, hipMemcpyAsync(m->bkdim, m->hbkdim, sizeof(int) * K, hipMemcpyHostToDevice, m->stream));
        hipLaunchKernelGGL(k_xupdate_seg<TV>, dim3(gN), dim3(kBlock), 0, m->stream, L, m->ecase, Z, L, m->bcoef,
                           m->bkdim, m->x);
        any = false;
        for (int c = 0; c < K; c++) {
            if (cyc[c] && res[c] <= tol[c]) live[c] = false;
            any = any || live[c];
        }
        if (!any) break;
        // true residual for the restart
        spmv<TM, double, double>(m, m->x, m->r, m->b);
        mdot_seg<double>(m, 1, m->r, 0, m->r, 0);
        CFD_HIP(m, hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * K, hipMemcpyDeviceToHost, m->stream));
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        for (int c = 0; c < K; c++) {
            if (!live[c]) continue;
            beta[c] = sqrt(m->hbred[(size_t)c * kBOut]);
            res[c] = beta[c];
            live[c] = cycle < o->max_restarts && beta[c] > tol[c] && beta[c] > 0.0;
        }
    }
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    for (int c = 0; c < K; c++) fres[c] = res[c];
    return CFD_OK;
}

// Restarted right-preconditioned FGMRES(m) with CGS2; same algorithm as
// orc_fgmres.  TM: matrix and preconditioner storage, TV: Krylov vectors.
template <class TM, class TV>
int fgmres(cfd_mesh *m, const cfd_linsolve_opts *o, const TM *a, const TM *lu, const TM *dinv, TV *V, TV *Z, TV *w,
           cfd_linsolve_stats *st) {
    const int mk = o->krylov_m;
    const size_t L = 5 * (size_t)m->N;
    const int gN = grid_cap(L, 4096);
    const int gS = (int)((L + kBlock - 1) / kBlock);
    double *hr = m->hred;
    // x0 = 0; r0 = b
    hipLaunchKernelGGL(k_fill, dim3(gN), dim3(kBlock), 0, m->stream, L, 0.0, m->x);
    CFD_HIP(m, hipMemcpyAsync(m->r, m->b, sizeof(double) * L, hipMemcpyDeviceToDevice, m->stream));
    mdot<double>(m, 1, m->r, 0, m->r, 0);
    CFD_HIP(m, hipMemcpyAsync(hr, m->redout, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    const double bnorm = sqrt(hr[0]);
    const double tol = o->rtol * bnorm;
    double beta = bnorm, res = bnorm;
    std::vector<double> H((size_t)(mk + 1) * mk, 0.0), cs(mk), sn(mk), gv(mk + 1), hcol(mk + 2), y(mk);
    int total = 0, cycles = 0;
    st->b_norm = bnorm;
    for (int cycle = 0; cycle <= o->max_restarts && beta > tol && beta > 0.0; cycle++) {
        cycles++;
        hipLaunchKernelGGL((k_scale<double, TV>), dim3(gN), dim3(kBlock), 0, m->stream, L, m->r, V, beta);
        for (int i = 0; i <= mk; i++) gv[i] = 0.0;
        gv[0] = beta;
        int kdim = 0;
        for (int j = 0; j < mk; j++) {
            int rc = enqueue_step<TM, TV>(m, a, lu, dinv, V, Z, w, j, mk);
            if (rc) return rc;
            CFD_HIP(m, hipStreamSynchronize(m->stream));
            for (int k = 0; k <= j; k++) hcol[k] = 0.0 + hr[k];
            for (int k = 0; k <= j; k++) hcol[k] += hr[kMaxKrylov + 1 + k];
            const double hn = sqrt(hr[2 * (kMaxKrylov + 1)]);
            hcol[j + 1] = hn;
            for (int k = 0; k < j; k++) {
                const double t = cs[k] * hcol[k] + sn[k] * hcol[k + 1];
                hcol[k + 1] = -sn[k] * hcol[k] + cs[k] * hcol[k + 1];
                hcol[k] = t;
            }
            const double den = sqrt(hcol[j] * hcol[j] + hcol[j + 1] * hcol[j + 1]);
            cs[j] = hcol[j] / den;
            sn[j] = hcol[j + 1] / den;
            hcol[j] = den;
            hcol[j + 1] = 0.0;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            for (int k = 0; k <= j; k++) H[(size_t)k * mk + j] = hcol[k];
            total++;
            kdim = j + 1;
            res = fabs(gv[j + 1]);
            if (!(res == res)) return cfd_set_error(m->ctx, CFD_E_NONFINITE, "FGMRES: non-finite residual");
            // the next step starts with V_{j+1} = w / hn (k_scale_dev)
            if (res <= tol || hn == 0.0) break;
        }
        for (int k = kdim - 1; k >= 0; k--) {
            double s = gv[k];
            for (int t = k + 1; t < kdim; t++) s -= H[(size_t)k * mk + t] * y[t];
            y[k] = s / H[(size_t)k * mk + k];
        }
        for (int k = 0; k < kdim; k++) hr[k] = y[k];
        CFD_HIP(m, hipMemcpyAsync(m->redout + 3 * (kMaxKrylov + 1), hr, sizeof(double) * kdim, hipMemcpyHostToDevice,
                                  m->stream));
        hipLaunchKernelGGL(k_xupdate<TV>, dim3(gN), dim3(kBlock), 0, m->stream, L, kdim, Z, L,
                           m->redout + 3 * (kMaxKrylov + 1), m->x);
        if (res <= tol) break;
        // true residual for the restart
        spmv<TM, double, double>(m, m->x, m->r, m->b);
        mdot<double>(m, 1, m->r, 0, m->r, 0);
        CFD_HIP(m, hipMemcpyAsync(hr, m->redout, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        beta = sqrt(hr[0]);
        res = beta;
    }
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    st->iterations = total;
    st->restarts = cycles > 0 ? cycles - 1 : 0;
    st->final_res = res;
    return CFD_OK;
}

}  // namespace

namespace {
int prepare_fgmres(cfd_mesh *m, const cfd_linsolve_opts *o);
int read_solve_flags(cfd_mesh *m);
}  // namespace

void drop_krylov_graphs(cfd_mesh *m) {
    CfdRuntimeGuard guard;
    for (KrylovGraphs *kg : {&m->kgraph, &m->kgraph_b}) {
        for (auto &g : kg->exec)
            if (g) hipGraphExecDestroy((hipGraphExec_t)g);
        kg->exec.clear();
        kg->key = nullptr;
    }
}

// Forward block Gauss-Seidel sweeps of MC_Iterative_Block_LU_Jacobi_CRS
// (MATH/src/MC_Iterative_Jacobi.c:346-535) with the LU factors of the
// diagonal blocks in gslu ([N][25]) and the per-row sums of the RMS in rowbuf
// ([N]); both are written in full before they are read.
static int gs_sweeps(cfd_mesh *m, int sweeps, double *gslu, double *rowbuf, cfd_linsolve_stats *st) {
    const size_t L = 5 * (size_t)m->N;
    hipLaunchKernelGGL(k_gs_lu, dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, m->N, m->iau, m->a,
                       gslu);
    hipLaunchKernelGGL(k_fill, dim3(grid_cap(L, 4096)), dim3(kBlock), 0, m->stream, L, 0.0, m->x);
    double rms = 0.0;
    int it;
    for (it = 0; it < sweeps; it++) {
        for (int l = 0; l < m->lower.nlevels; l++) {
            const int b0 = m->lower.ptr[l], nr = m->lower.ptr[l + 1] - b0;
            hipLaunchKernelGGL(k_gs_level, dim3((nr + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, nr,
                               m->lower.rows + b0, m->ia, m->ja, m->iau, m->a, m->b, gslu, m->x, rowbuf);
        }
        // row RMS: ordered reduction of the per-row sums
        const int np = grid_cap(m->N, kDotParts);
        hipLaunchKernelGGL(k_vsum_partial, dim3(np), dim3(kBlock), 0, m->stream, (size_t)m->N, rowbuf, m->red);
        hipLaunchKernelGGL(k_mdot_final, dim3(1), dim3(64), 0, m->stream, np, 1, m->red, m->redout);
        CFD_HIP(m, hipMemcpyAsync(m->hred, m->redout, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        rms = sqrt(m->hred[0] / (double)m->N);
        if (rms < 100.0 * DBL_EPSILON) {
            it++;
            break;
        }
    }
    st->iterations = it;
    st->final_res = rms;
    return CFD_OK;
}

// Scratch of cfd_linsolve_gs_noalloc: the LU factors of the diagonal blocks
// (25 values per row) and the per-row sums (1 per row) of meshes of up to
// kNoAllocMaxNodes nodes, 6.8 MB.  Each GPU holds its own copy until the
// process ends; host code uses the address hipGetSymbolAddress returns for
// the current GPU (kept per context), never the array's name.  Every call on
// a GPU shares that GPU's copy, so calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[26 * (size_t)kNoAllocMaxNodes];

double *noalloc_scratch_address(hipError_t *err) {
    void *p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    return *err == hipSuccess ? (double *)p : nullptr;
}

int run_gs_noalloc(cfd_mesh *m, int sweeps, cfd_linsolve_stats *st) {
    *st = cfd_linsolve_stats{};
    if (!m->jac_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "cfd_linsolve_gs_noalloc: no Jacobian assembled");
    if (m->jac_mixed) return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_linsolve_gs_noalloc: needs an FP64 Jacobian");
    if (m->N > kNoAllocMaxNodes)
        return cfd_set_error(m->ctx, CFD_E_ARG,
                             "cfd_linsolve_gs_noalloc: %d nodes exceed the scratch array's %d (use cfd_linsolve)",
                             m->N, kNoAllocMaxNodes);
    if (!m->ctx->noalloc_scratch)
        return cfd_set_error(m->ctx, CFD_E_HIP, "cfd_linsolve_gs_noalloc: scratch array address unavailable");
    double *gslu = m->ctx->noalloc_scratch, *rowbuf = gslu + 25 * (size_t)kNoAllocMaxNodes;
    const int rc = gs_sweeps(m, sweeps > 0 ? sweeps : 10000, gslu, rowbuf, st);
    if (rc) return rc;
    CFD_HIP(m, hipGetLastError());
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    return CFD_OK;
}

int run_linsolve(cfd_mesh *m, const cfd_linsolve_opts *o, cfd_linsolve_stats *st) {
    if (!m->jac_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "cfd_linsolve: no Jacobian assembled");
    *st = cfd_linsolve_stats{};
    CFD_HIP(m, hipMemsetAsync(m->flag, 0, 4 * sizeof(int), m->stream));
    if (o->solver == CFD_SOLVER_BLOCK_GS || o->solver == CFD_SOLVER_LUSGS) {
        if (m->jac_mixed) return cfd_set_error(m->ctx, CFD_E_ARG, "reference sweeps need an FP64 Jacobian");
        const int sweeps = o->sweeps > 0 ? o->sweeps : (o->solver == CFD_SOLVER_BLOCK_GS ? 10000 : 100);
        if (o->solver == CFD_SOLVER_BLOCK_GS) {
            if (!m->gslu) m->gslu = (double *)cfd_dalloc(m, sizeof(double) * 25 * (size_t)m->N);
            if (!m->rowbuf) m->rowbuf = (double *)cfd_dalloc(m, sizeof(double) * (size_t)m->N);
            if (!m->gslu || !m->rowbuf) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory");
            int rc = gs_sweeps(m, sweeps, m->gslu, m->rowbuf, st);
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(k_lusgs_init, dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, m->N,
                               m->b, m->x);
            for (int it = 0; it < sweeps; it++) {
                for (int pass = 0; pass < 2; pass++) {
                    const LevelSchedule &ls = pass ? m->upper : m->lower;
                    for (int l = 0; l < ls.nlevels; l++) {
                        const int b0 = ls.ptr[l], nr = ls.ptr[l + 1] - b0;
                        hipLaunchKernelGGL(k_lusgs_level, dim3((nr + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream,
                                           nr, ls.rows + b0, m->ia, m->ja, m->iau, m->a, m->b, m->x);
                    }
                }
            }
            st->iterations = sweeps;
        }
        CFD_HIP(m, hipGetLastError());
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        return CFD_OK;
    }
    int rc = prepare_fgmres(m, o);
    if (rc) return rc;
    if (o->mixed)
        rc = fgmres<float, float>(m, o, m->af, m->luf, m->dinvf, m->Vf, m->Zf, m->wf, st);
    else
        rc = fgmres<double, double>(m, o, m->a, m->lu, m->dinv, m->V, m->Z, m->w, st);
    if (rc) return rc;
    return read_solve_flags(m);
}

int run_linsolve_batch(cfd_mesh *m, const cfd_linsolve_opts *o, const std::vector<char> &solve,
                       std::vector<int> &iters, std::vector<double> &final_res, std::vector<int> &status) {
    if (m->ncase < 1) return cfd_set_error(m->ctx, CFD_E_STATE, "batched solve on a mesh that is not a union of cases");
    if (!m->jac_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "batched linsolve: no Jacobian assembled");
    if (o->solver != CFD_SOLVER_FGMRES) return cfd_set_error(m->ctx, CFD_E_ARG, "batched solves use FGMRES");
    CFD_HIP(m, hipMemsetAsync(m->flag, 0, 4 * sizeof(int), m->stream));
    int rc = prepare_fgmres(m, o);
    if (rc) return rc;
    if (o->mixed)
        rc = fgmres_batch<float, float>(m, o, m->af, m->luf, m->dinvf, m->Vf, m->Zf, m->wf, solve, iters, final_res,
                                        status);
    else
        rc = fgmres_batch<double, double>(m, o, m->a, m->lu, m->dinv, m->V, m->Z, m->w, solve, iters, final_res,
                                          status);
    if (rc) return rc;
    return read_solve_flags(m);
}

namespace {

// Jacobian precision check, Krylov workspace and preconditioner for FGMRES.
// Allocates the block-SELL value arrays of storage type TM: A always, the
// multicolour factors for that preconditioner.
template <class TM>
int ensure_sell(cfd_mesh *m, int precond, bool half, bool half_a) {
    if (!m->use_sell) return CFD_OK;
    TM *&av = SellVals<TM>::a(m);
    if (!av) av = (TM *)cfd_dalloc(m, sizeof(TM) * 25 * (size_t)m->a_sell.nblk);
    if (!av) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (block-SELL Jacobian)");
    if (precond != CFD_PREC_ILU0_MC) return CFD_OK;
    if (half) {
        const size_t nl = std::max<int64_t>(m->mcl_sell.nblk, 64), nu = std::max<int64_t>(m->mcu_sell.nblk, 64),
                     nd = ((size_t)m->N + 63) / 64 * 64;
        if (half_a && !m->asvh) m->asvh = (_Float16 *)cfd_dalloc(m, sizeof(_Float16) * 25 * (size_t)m->a_sell.nblk);
        if (half_a && !m->asvs) m->asvs = (float *)cfd_dalloc(m, sizeof(float) * (size_t)m->a_sell.nblk);
        if (half_a && (!m->asvh || !m->asvs))
            return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (FP16 A)");
        if (!m->lsvh) m->lsvh = (_Float16 *)cfd_dalloc(m, sizeof(_Float16) * 25 * nl);
        if (!m->usvh) m->usvh = (_Float16 *)cfd_dalloc(m, sizeof(_Float16) * 25 * nu);
        if (!m->dsvh) m->dsvh = (_Float16 *)cfd_dalloc(m, sizeof(_Float16) * 25 * nd);
        if (!m->lsvs) m->lsvs = (float *)cfd_dalloc(m, sizeof(float) * nl);
        if (!m->usvs) m->usvs = (float *)cfd_dalloc(m, sizeof(float) * nu);
        if (!m->dsvs) m->dsvs = (float *)cfd_dalloc(m, sizeof(float) * nd);
        if (!m->lsvh || !m->usvh || !m->dsvh || !m->lsvs || !m->usvs || !m->dsvs)
            return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (FP16 block-SELL factors)");
        return CFD_OK;
    }
    TM *&lv = SellVals<TM>::l(m), *&uv = SellVals<TM>::u(m), *&dv = SellVals<TM>::d(m);
    if (!lv) lv = (TM *)cfd_dalloc(m, sizeof(TM) * 25 * (size_t)std::max<int64_t>(m->mcl_sell.nblk, 64));
    if (!uv) uv = (TM *)cfd_dalloc(m, sizeof(TM) * 25 * (size_t)std::max<int64_t>(m->mcu_sell.nblk, 64));
    if (!dv) dv = (TM *)cfd_dalloc(m, sizeof(TM) * 25 * (((size_t)m->N + 63) / 64 * 64));
    if (!lv || !uv || !dv) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (block-SELL factors)");
    return CFD_OK;
}

int prepare_fgmres(cfd_mesh *m, const cfd_linsolve_opts *o) {
    if (o->krylov_m < 1 || o->krylov_m > kMaxKrylov)
        return cfd_set_error(m->ctx, CFD_E_ARG, "krylov_m must be in [1, %d]", kMaxKrylov);
    if (o->precond < CFD_PREC_NONE || o->precond > CFD_PREC_ILU0_MC)
        return cfd_set_error(m->ctx, CFD_E_ARG, "unknown preconditioner %d", o->precond);
    if (o->mixed < 0 || o->mixed > 3) return cfd_set_error(m->ctx, CFD_E_ARG, "mixed must be 0, 1, 2 or 3");
    const bool mixed = o->mixed != 0;
    // FP16 copies of the factors (mixed >= 2) and of A (mixed = 3), for the
    // multicolour ILU(0) on the block-SELL kernels
    const bool sell_mc = o->precond == CFD_PREC_ILU0_MC && m->use_sell;
    const bool half = o->mixed >= 2 && sell_mc, half_a = o->mixed == 3 && sell_mc;
    if (mixed != m->jac_mixed)
        return cfd_set_error(m->ctx, CFD_E_ARG, "Jacobian precision does not match linsolve precision");
    int rc = ensure_krylov(m, o->krylov_m, mixed);
    if (rc) return rc;
    rc = mixed ? ensure_sell<float>(m, o->precond, half, half_a)
               : ensure_sell<double>(m, o->precond, false, false);
    if (rc) return rc;
    if (m->use_sell && (!m->asell_valid || m->asell_half != half_a)) {
        if (m->asell_half != half_a) drop_krylov_graphs(m);  // captured steps launch the other SpMV kernel
        m->asell_half = half_a;
        if (half_a)
            hipLaunchKernelGGL(k_to_sell_half, dim3(grid_cap((size_t)m->a_sell.nblk, 8192)), dim3(kBlock), 0,
                               m->stream, (size_t)m->a_sell.nblk, m->a_sell.src, m->af, m->asvh, m->asvs);
        else if (mixed)
            hipLaunchKernelGGL(k_to_sell<float>, dim3(grid_cap((size_t)m->a_sell.nblk, 8192)), dim3(kBlock), 0,
                               m->stream, (size_t)m->a_sell.nblk, m->a_sell.src, m->af, m->asvf);
        else
            hipLaunchKernelGGL(k_to_sell<double>, dim3(grid_cap((size_t)m->a_sell.nblk, 8192)), dim3(kBlock), 0,
                               m->stream, (size_t)m->a_sell.nblk, m->a_sell.src, m->a, m->asv);
        CFD_HIP(m, hipGetLastError());
        m->asell_valid = true;
    }
    if (!m->prec_valid || m->prec_kind != o->precond || m->prec_mixed != mixed || m->prec_half != half) {
        if (m->prec_half != half) drop_krylov_graphs(m);  // captured steps launch the other sweep kernels
        m->prec_kind = o->precond;
        m->prec_mixed = mixed;
        m->prec_half = half;
        if (mixed) {
            if ((o->precond == CFD_PREC_ILU0 || o->precond == CFD_PREC_ILU0_MC) && !m->luf)
                m->luf = (float *)cfd_dalloc(m, sizeof(float) * 25 * (size_t)m->nnz);
            if (o->precond != CFD_PREC_NONE && !m->dinvf) m->dinvf = (float *)cfd_dalloc(m, sizeof(float) * 25 * (size_t)m->N);
            rc = build_prec<float>(m, o->precond, m->af, m->luf, m->dinvf, half);
        } else {
            if ((o->precond == CFD_PREC_ILU0 || o->precond == CFD_PREC_ILU0_MC) && !m->lu)
                m->lu = (double *)cfd_dalloc(m, sizeof(double) * 25 * (size_t)m->nnz);
            if (o->precond != CFD_PREC_NONE && !m->dinv) m->dinv = (double *)cfd_dalloc(m, sizeof(double) * 25 * (size_t)m->N);
            rc = build_prec<double>(m, o->precond, m->a, m->lu, m->dinv, false);
        }
        if (rc) return rc;
        m->prec_valid = true;
    }
    return CFD_OK;
}

// Error flags raised by the preconditioner and the sweeps.
int read_solve_flags(cfd_mesh *m) {
    CFD_HIP(m, hipMemcpyAsync(m->hflag, m->flag, 4 * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    if (m->hflag[2]) return cfd_set_error(m->ctx, CFD_E_SINGULAR, "zero pivot in %d diagonal blocks", m->hflag[2]);
    if (m->hflag[3]) return cfd_set_error(m->ctx, CFD_E_HIP, "sync-free triangular sweep exceeded its wait bound");
    return CFD_OK;
}

}  // namespace
