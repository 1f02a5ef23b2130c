// Residual path kernels: characteristic BC, LSQ and Green-Gauss gradients,
// limiters, the flux residual, local time step, RMS, update and forces.
//
// Every node-centred kernel owns one node per thread and accumulates its
// contributions in the order the reference's sequential edge loops deliver
// them (interior edges by index, then boundary edges), so the FP64 results
// are bit-identical to the reference without atomics or colouring; each edge
// flux is evaluated by both of its nodes.
#include <cfloat>
#include <cstdlib>
#include <cmath>

#include "cfd_internal.h"
#include "flux.h"

// This file is also the fast-arithmetic build (residual_fast.hip defines
// CFD_FAST_MATH 1 and compiles it with FMA contraction); its launchers then
// live in res_fast instead of res_exact.
#if CFD_FAST_MATH
#define CFD_RES_NS res_fast
#else
#define CFD_RES_NS res_exact
#endif

using namespace cfd;

// The kernels live in the build's namespace too, so that profiles tell the
// two builds' kernels apart (res_fast::... and res_exact::...).
namespace CFD_RES_NS {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void flag_set(int *flag, int idx) { atomicAdd(flag + idx, 1); }

// Non-negative doubles/floats order like their bit patterns.
__device__ __forceinline__ void atomic_min_nonneg(unsigned long long *a, double v) {
    atomicMin(a, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long *a, double v) {
    atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// LSQ coefficients, Compute_Least_Square_Gradient_Coefficients,
// Euler3D/src/Gradient.cpp:137-222 (unweighted: alpha = 1)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_lsq_coef(int N, const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ qa, double *__restrict__ lsq) {
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= N) return;
    const double xi = qa[kQA * (size_t)i + 5], yi = qa[kQA * (size_t)i + 6], zi = qa[kQA * (size_t)i + 7];
    double r11 = 0.0, r12 = 0.0, r13 = 0.0, r22 = 0.0, r23 = 0.0, ro22 = 0.0, ro33 = 0.0;
    const int b = ia[i], e = ia[i + 1];
    for (int k = b; k < e; k++) {
        const size_t j = ja[k];
        double dx = 1.0 * (qa[kQA * j + 5] - xi), dy = 1.0 * (qa[kQA * j + 6] - yi), dz = 1.0 * (qa[kQA * j + 7] - zi);
        r11 += dx * dx;
        r12 += dx * dy;
        r13 += dx * dz;
    }
    for (int k = b; k < e; k++) {
        const size_t j = ja[k];
        double dx = 1.0 * (qa[kQA * j + 5] - xi), dy = 1.0 * (qa[kQA * j + 6] - yi), dz = 1.0 * (qa[kQA * j + 7] - zi);
        double t = dy - (r12 / r11) * dx;
        r22 += t * t;
        r23 += t * dz;
        ro22 += t * dy;
    }
    for (int k = b; k < e; k++) {
        const size_t j = ja[k];
        double dx = 1.0 * (qa[kQA * j + 5] - xi), dy = 1.0 * (qa[kQA * j + 6] - yi), dz = 1.0 * (qa[kQA * j + 7] - zi);
        double t = dz - (r13 / r11) * dx - (r23 / r22) * (dy - (r12 / r11) * dx);
        ro33 += t * dz;
    }
    double *o = lsq + 8 * (size_t)i;
    o[0] = r11;
    o[1] = r12;
    o[2] = r13;
    o[3] = r22;
    o[4] = r23;
    o[5] = ro22;
    o[6] = ro33;
    o[7] = 0.0;
}

// ---------------------------------------------------------------------------
// Characteristic boundary condition, Compute_Characteristic_BoundaryCondition_New,
// Euler3D/src/BC.cpp:381-613, one thread per boundary edge.
// ---------------------------------------------------------------------------
template <class T>
struct BcArgs {
    int N, E, B, iter, zpg;
    const int *bnode, *btype;
    const T *sa;
    T *qa;
    T g, inf_rho, inf_u, inf_v, inf_w, inf_p, out_p;
    int *flag;
    // union of independent cases: freestream of each case (rho, u, v, w, p,
    // outflow p) and the case of each boundary edge; NULL for one case
    const double *case_inf;
    const int *bcase;
    // cases still iterating (union of cases); a finished case keeps the
    // ghost states of its last iteration.  NULL: every edge.
    const int *case_active;
    // failure bits per case (union of cases), NULL for one case
    int *case_flag;
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_bc(BcArgs<T> a) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= a.B) return;
    if (a.case_active && !a.case_active[a.bcase[e]]) return;
    const size_t nl = a.bnode[e], nr = (size_t)a.N + e;
    T QL[5], QR[5], QB[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        QL[k] = a.qa[kQA * nl + k];
        QR[k] = a.qa[kQA * nr + k];
    }
    const T *s = a.sa + 4 * (size_t)e;
    BcFree<T> f{a.inf_rho, a.inf_u, a.inf_v, a.inf_w, a.inf_p, a.out_p};
    if (a.case_inf) f = bc_free_of_case<T>(a.case_inf, a.bcase[e]);
    const int st = bc_ghost(a.g, QL, QR, s[0], s[1], s[2], a.btype[e], a.iter, a.zpg, f, QB);
    if (st == 2) {
        flag_set(a.flag, 3);
        if (a.case_flag) atomicOr(a.case_flag + a.bcase[e], 2);
        return;
    }
    if (st == 1) {
        flag_set(a.flag, 0);
        if (a.case_flag) atomicOr(a.case_flag + a.bcase[e], 1);
    }
    T *o = a.qa + kQA * nr;
#pragma unroll
    for (int k = 0; k < 5; k++) o[k] = QB[k];
}

// ---------------------------------------------------------------------------
// Direct global->LDS loads (gfx950 global_load_lds): each lane's W bytes go
// to row r of a wave's staging slot at column lane.  Neighbour walks use them
// to bring incidence k+1 into LDS while incidence k is processed; no VGPR
// holds an in-flight load, so nothing waits on the prefetch early.  The
// compiler does not order these loads against LDS reads: a walk waits
// vmcnt(0) before reading its slot and lgkmcnt(0) before refilling it.
// ---------------------------------------------------------------------------
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) void glb_void;

// cache-policy bits of the staging loads: streaming data (indices, areas)
// is not kept in the caches, node records are
constexpr int kPolNT = 2;

template <int POL>
__device__ __forceinline__ void dma16(const void *g, void *l) {
    __builtin_amdgcn_global_load_lds((glb_void *)g, (lds_void *)l, 16, 0, POL);
}
template <int POL>
__device__ __forceinline__ void dma4(const void *g, void *l) {
    __builtin_amdgcn_global_load_lds((glb_void *)g, (lds_void *)l, 4, 0, POL);
}


// ---------------------------------------------------------------------------
// Gradients
// ---------------------------------------------------------------------------
// Compute_Least_Square_Gradient, Euler3D/src/Gradient.cpp:227-295
template <class T>
__global__ __launch_bounds__(kBlock) void k_lsq_grad(int N, const int *__restrict__ order, const int *__restrict__ ia,
                                                     const int *__restrict__ ja, const T *__restrict__ lsq,
                                                     const T *__restrict__ qa, T *__restrict__ qb) {
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= N) return;
    const T *ra = qa + kQA * (size_t)i;
    T qi[5], xi = ra[5], yi = ra[6], zi = ra[7];
#pragma unroll
    for (int k = 0; k < 5; k++) qi[k] = ra[k];
    const T *c = lsq + 8 * (size_t)i;
    const T r11 = c[0], r12 = c[1], r13 = c[2], r22 = c[3], r23 = c[4], ro22 = c[5], ro33 = c[6];
    T gs[15];
#pragma unroll
    for (int f = 0; f < 15; f++) gs[f] = T(0.0);
    const int b = ia[i], e = ia[i + 1];
    for (int k = b; k < e; k++) {
        const T *rj = qa + kQA * (size_t)ja[k];
        const T alpha = T(1.0);
        T dx = alpha * (rj[5] - xi), dy = alpha * (rj[6] - yi), dz = alpha * (rj[7] - zi);
        T Wz = (dz - (r13 / r11) * dx - (r23 / r22) * (dy - (r12 / r11) * dx)) / ro33;
        T Wy = (dy - (r12 / r11) * dx - r23 * Wz) / ro22;
        T Wx = (dx - r12 * Wy - r13 * Wz) / r11;
#pragma unroll
        for (int f = 0; f < 5; f++) {
            T d = rj[f] - qi[f];
            gs[3 * f + 0] += Wx * alpha * d;
            gs[3 * f + 1] += Wy * alpha * d;
            gs[3 * f + 2] += Wz * alpha * d;
        }
    }
    T *o = qb + kQB * (size_t)i;
#pragma unroll
    for (int f = 0; f < 15; f++) o[f] = gs[f];
}

// Green-Gauss sum over the control volume: divided in the exact build, times
// 1/v (formed once per node; the compiler hoists it) in the fast build.
template <class T>
__device__ __forceinline__ T grad_of(T sum, T v) {
#if CFD_FAST_MATH
    return sum * (T(1.0) / v);
#else
    return sum / v;
#endif
}

// Median-dual edge-based Green-Gauss:
//   grad_i = (1/V_i) [ sum_edges +-1/2 (Q_L+Q_R) S_LR + sum_bnd Q_i S_b ]
// accumulated in edge order (the reference has no 3D Green-Gauss).  One
// thread per node walks the node's SELL incidence list.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gg_grad(int N, const int *__restrict__ ip, const int *__restrict__ sptr,
                                                    const int *__restrict__ nbr, const T *__restrict__ isx,
                                                    const T *__restrict__ isy, const T *__restrict__ isz,
                                                    const T *__restrict__ vol, const T *__restrict__ qa,
                                                    T *__restrict__ qb) {
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= N) return;
    const T *ra = qa + kQA * (size_t)i;
    T qi[5];
#pragma unroll
    for (int k = 0; k < 5; k++) qi[k] = ra[k];
    T gs[15];
#pragma unroll
    for (int f = 0; f < 15; f++) gs[f] = T(0.0);
    const int deg = ip[i + 1] - ip[i];
    size_t idx = (size_t)sptr[i / kSlice] + (i % kSlice);
#pragma unroll 2
    for (int k = 0; k < deg; k++, idx += kSlice) {
        const int jr = ld_stream(nbr + idx);
        const T s0 = ld_stream(isx + idx), s1 = ld_stream(isy + idx), s2 = ld_stream(isz + idx);
        if (jr >= N) {
#pragma unroll
            for (int f = 0; f < 5; f++) {
                gs[3 * f + 0] += qi[f] * s0;
                gs[3 * f + 1] += qi[f] * s1;
                gs[3 * f + 2] += qi[f] * s2;
            }
        } else {
            const bool iL = jr >= 0;
            const T *rj = qa + kQA * (size_t)(iL ? jr : ~jr);
#pragma unroll
            for (int f = 0; f < 5; f++) {
                const T qf = T(0.5) * (qi[f] + rj[f]);
                if (iL) {
                    gs[3 * f + 0] += qf * s0;
                    gs[3 * f + 1] += qf * s1;
                    gs[3 * f + 2] += qf * s2;
                } else {
                    gs[3 * f + 0] -= qf * s0;
                    gs[3 * f + 1] -= qf * s1;
                    gs[3 * f + 2] -= qf * s2;
                }
            }
        }
    }
    const T v = vol[i];
    T *o = qb + kQB * (size_t)i;
#pragma unroll
    for (int f = 0; f < 15; f++) o[f] = grad_of(gs[f], v);
}

// ---------------------------------------------------------------------------
// Limiters, Euler3D/src/Limiters.cpp:20-304
// ---------------------------------------------------------------------------
// Relative pressure margin below which a face state counts as near-vacuum.
// When every face state of both ends of an edge has rho > 0 and p above this
// margin, the Roe-averaged pressure of the edge is positive: h - q^2/2 of the
// Roe average is at least the sqrt(rho)-weighted mean of c_k^2/(g-1) > 0, a
// margin orders of magnitude above the rounding of its evaluation.
template <class T>
__device__ __forceinline__ T vacuum_margin();
template <>
__device__ __forceinline__ double vacuum_margin<double>() { return 1.0e-6; }
template <>
__device__ __forceinline__ float vacuum_margin<float>() { return 1.0e-3f; }

template <class T>
__device__ __forceinline__ bool near_vacuum(const T *Q, T P) {
    return !(Q[0] > T(0.0)) || !(P > vacuum_margin<T>() * fabs(Q[4]));
}

// One neighbour's bound on the limiter of node c (Compute_Limiter_Venkatakrishnan,
// Limiters.cpp:185-267; Barth-Jespersen, Limiters.cpp:116-180): qj and
// (xj, yj, zj) are the neighbour's state and coordinates.
// The fast build (CFD_FAST_MATH) keeps the Venkatakrishnan minimum as a
// fraction ph[k] / pd[k] (the denominator is positive: (dP + dM/2)^2 +
// 7/4 dM^2 + eps^2 with dM != 0) and compares fractions by cross
// multiplication, so a node divides once per variable (limiter_finish)
// instead of once per neighbour; the exact build divides per neighbour as the
// reference does and leaves pd alone.
// F0: the reconstruction fraction is 0, so the face value is qc + grad . R
// and the neighbour's state is not read.
template <class T, int KIND, bool F0 = false>  // KIND 1 Venkatakrishnan, 2 Barth-Jespersen
__device__ __forceinline__ void limiter_nbr(const T *qj, T xj, T yj, T zj, const T *qc, T xc, T yc, T zc, const T *gr,
                                            const T *qmin, const T *qmax, T eps2, T frac, T *ph, T *pd) {
    const T Rx = T(0.5) * (xj - xc), Ry = T(0.5) * (yj - yc), Rz = T(0.5) * (zj - zc);
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const T gR = gr[3 * k] * Rx + gr[3 * k + 1] * Ry + gr[3 * k + 2] * Rz;
        const T qf = F0 ? qc[k] + gR : qc[k] + frac * (qj[k] - qc[k]) + (T(1.0) - frac) * gR;
        const T dM = qf - qc[k];
        // The three cases (dM > 0: towards the maximum, dM < 0: towards the
        // minimum, otherwise no bound) as selects, so the five variables'
        // divisions are independent straight-line code; the value chosen
        // is the one the branch would compute.
        const bool up = dM > T(0.0), bound = up || dM < T(0.0);
        const T dP = up ? qmax[k] - qc[k] : qmin[k] - qc[k];
#if CFD_FAST_MATH
        if (KIND == 1) {
            const T num = (dP * dP + eps2) + T(2.0) * (dM * dP);
            const T den = dP * dP + T(2.0) * dM * dM + dM * dP + eps2;
            const bool lower = bound && num * pd[k] < ph[k] * den;
            ph[k] = lower ? num : ph[k];
            pd[k] = lower ? den : pd[k];
            continue;
        }
#endif
        T t;
        if (KIND == 1) {
            t = (dP * dP + eps2) + T(2.0) * (dM * dP);
            t = t / (dP * dP + T(2.0) * dM * dM + dM * dP + eps2);
        } else {
            t = tmin(T(1.0), dP / dM);
        }
        ph[k] = tmin(ph[k], bound ? t : T(1.0));
        ph[k] = tmax(T(0.0), ph[k]);
    }
}

// The limiter value of each variable once every neighbour has bounded it
// (see limiter_nbr): the fast build's fraction, clamped at zero as the
// reference clamps after each neighbour.
template <class T, int KIND>
__device__ __forceinline__ void limiter_finish(T *ph, const T *pd) {
#if CFD_FAST_MATH
    if (KIND == 1) {
#pragma unroll
        for (int k = 0; k < 5; k++) ph[k] = tmax(T(0.0), ph[k] / pd[k]);
    }
#endif
}

// Pressure of the face state towards one neighbour with the limiter ph
// (Compute_Limiter_PressureCorrection, Limiters.cpp:273-304): a negative
// pressure zeroes the limiter; a near-vacuum one marks the node as risky.
template <class T, bool F0 = false>
__device__ __forceinline__ void pressure_nbr(const T *qj, T xj, T yj, T zj, const T *qc, T xc, T yc, T zc,
                                             const T *gr, T frac, T g, T *ph, bool &zeroed, bool &risky) {
    const T Rx = T(0.5) * (xj - xc), Ry = T(0.5) * (yj - yc), Rz = T(0.5) * (zj - zc);
    T Q[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const T gR = gr[3 * k] * Rx + gr[3 * k + 1] * Ry + gr[3 * k + 2] * Rz;
        Q[k] = F0 ? qc[k] + ph[k] * gR : qc[k] + ph[k] * (frac * (qj[k] - qc[k]) + (T(1.0) - frac) * gR);
    }
    const T P = (g - T(1.0)) * (Q[4] - (T(0.5) / Q[0]) * (Q[1] * Q[1] + Q[2] * Q[2] + Q[3] * Q[3]));
    const bool neg = P < T(0.0);
    const bool vac = !neg && near_vacuum(Q, P);
    if (neg) {
#pragma unroll
        for (int k = 0; k < 5; k++) ph[k] = T(0.0);
    }
    // both flags written on every path: a store to one flag or the other
    // made the compiler keep the pair in scratch memory
    zeroed = zeroed || neg;
    risky = risky || vac;
}

// After every neighbour: if a face zeroed the limiter, every face state is
// the node state, whose own pressure decides the risk flag.
template <class T>
__device__ __forceinline__ bool pressure_final(const T *qc, T g, bool zeroed, bool risky) {
    if (zeroed) {
        const T P = (g - T(1.0)) * (qc[4] - (T(0.5) / qc[0]) * (qc[1] * qc[1] + qc[2] * qc[2] + qc[3] * qc[3]));
        risky = near_vacuum(qc, P);
    }
    return risky;
}

// Whether no face state of node c can have a negative or near-vacuum pressure,
// from bounds alone.  Component k of a face state differs from qc[k] by
// ph[k] |frac (qj - qc) + (1 - frac) gr.R|, which is at most
// ph[k] (|frac| A_k + |1 - frac| |gr_k| rmax), where A_k bounds |qj - qc| by
// the neighbour extrema and rmax is the longest half edge.  From these, a lower
// bound of the density and of the pressure (concave in the state) and an upper
// bound of |E|; the slack covers the rounding of both this bound and the face
// states.  When it holds, the pressure pass would change nothing: every face
// has a positive density and a pressure above the vacuum margin.
template <class T, bool F0 = false>
__device__ __forceinline__ bool faces_safe(const T *qc, const T *gr, const T *qmin, const T *qmax, const T *ph,
                                           T frac, T rmax, T g) {
    const T slack = sizeof(T) == 8 ? T(1e-12) : T(1e-5);
    T e[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const T A = tmax(qmax[k] - qc[k], qc[k] - qmin[k]);
        const T G = sqrt(gr[3 * k] * gr[3 * k] + gr[3 * k + 1] * gr[3 * k + 1] + gr[3 * k + 2] * gr[3 * k + 2]);
        e[k] = F0 ? ph[k] * (G * rmax) : ph[k] * (fabs(frac) * A + fabs(T(1.0) - frac) * G * rmax);
        e[k] = e[k] * (T(1.0) + slack) + slack * fabs(qc[k]);
    }
    const T rho = qc[0] - e[0];
    const T m1 = fabs(qc[1]) + e[1], m2 = fabs(qc[2]) + e[2], m3 = fabs(qc[3]) + e[3];
    const T kin = T(0.5) * (m1 * m1 + m2 * m2 + m3 * m3) / rho;
    const T P = (g - T(1.0)) * ((qc[4] - e[4]) - kin);
    const T Eabs = fabs(qc[4]) + e[4];
    return rho > slack * fabs(qc[0]) && kin >= T(0.0) &&
           P > vacuum_margin<T>() * Eabs * (T(1.0) + slack) + slack * (g - T(1.0)) * (Eabs + kin);
}

// Limiter value of every variable of node c from its state qc, coordinates,
// gradient gr and the neighbour extrema, followed by the pressure correction
// (Compute_Limiter_Venkatakrishnan / _BarthJespersen and
// Compute_Limiter_PressureCorrection, Limiters.cpp:130-304), which is skipped
// when faces_safe shows it cannot act.  Returns whether a final face state of
// the node is near vacuum (the only case in which the Roe pressure correction
// can act).  Neither pass depends on the order of the neighbours.
template <class T, int KIND, bool F0 = false>  // KIND 1 Venkatakrishnan, 2 Barth-Jespersen
__device__ __forceinline__ bool limiter_passes(int N, int deg, size_t base, const int *__restrict__ nbr,
                                               const T *__restrict__ qa, const T *qc, T xc, T yc, T zc, const T *gr,
                                               const T *qmin, const T *qmax, T v, T g, T frac, T k3, T *ph) {
    T pd[5];
#pragma unroll
    for (int k = 0; k < 5; k++) ph[k] = pd[k] = T(1.0);
    T eps2 = k3;
    eps2 = T(6.0) * eps2 * v / T(M_PI);
    T r2 = T(0.0);  // longest half edge, squared
    // two neighbours at a time: both records are loaded before either bound
    // is applied (the minimum does not depend on the order)
    for (int n = 0; n < deg; n += 2) {
        const int j0 = nbr[base + (size_t)n * kSlice];
        const int j1 = n + 1 < deg ? nbr[base + (size_t)(n + 1) * kSlice] : N;
        T r[2][8];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int jr = u ? j1 : j0;
            if (jr >= N) continue;
            const T *rj = qa + kQA * (size_t)(jr >= 0 ? jr : ~jr);
#pragma unroll
            for (int e = F0 ? 5 : 0; e < 8; e++) r[u][e] = rj[e];
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if ((u ? j1 : j0) >= N) continue;
            limiter_nbr<T, KIND, F0>(r[u], r[u][5], r[u][6], r[u][7], qc, xc, yc, zc, gr, qmin, qmax, eps2, frac, ph, pd);
            const T Rx = T(0.5) * (r[u][5] - xc), Ry = T(0.5) * (r[u][6] - yc), Rz = T(0.5) * (r[u][7] - zc);
            r2 = tmax(r2, Rx * Rx + Ry * Ry + Rz * Rz);
        }
    }
    limiter_finish<T, KIND>(ph, pd);
    if (faces_safe<T, F0>(qc, gr, qmin, qmax, ph, frac, sqrt(r2), g)) return false;
    bool zeroed = false, risky = false;
    for (int n = 0; n < deg; n++) {
        const int jr = nbr[base + (size_t)n * kSlice];
        if (jr >= N) continue;
        const T *rj = qa + kQA * (size_t)(jr >= 0 ? jr : ~jr);
        pressure_nbr<T, F0>(rj, rj[5], rj[6], rj[7], qc, xc, yc, zc, gr, frac, g, ph, zeroed, risky);
    }
    return pressure_final<T>(qc, g, zeroed, risky);
}

template <class T, int KIND>
__global__ __launch_bounds__(kBlock) void k_limiter(int N, const int *__restrict__ ip, const int *__restrict__ sptr,
                                                    const int *__restrict__ nbr, const T *__restrict__ qa,
                                                    T *__restrict__ qb, const T *__restrict__ vol, T g, T frac, T k3,
                                                    int *risk) {
    const int c = xcd_block() * kBlock + threadIdx.x;
    if (c >= N) return;
    const T *ra = qa + kQA * (size_t)c;
    T qc[5], xc = ra[5], yc = ra[6], zc = ra[7];
#pragma unroll
    for (int k = 0; k < 5; k++) qc[k] = ra[k];
    T *rb = qb + kQB * (size_t)c;
    T gr[15];
#pragma unroll
    for (int f = 0; f < 15; f++) gr[f] = rb[f];
    T qmin[5], qmax[5];
#pragma unroll
    for (int k = 0; k < 5; k++) qmin[k] = qmax[k] = qc[k];
    // The neighbours are the interior entries of the node's incidence list;
    // min/max and the limiter minimum do not depend on their order.
    const int deg = ip[c + 1] - ip[c];
    const size_t base = (size_t)sptr[c / kSlice] + (c % kSlice);
    for (int n = 0; n < deg; n++) {
        const int jr = nbr[base + (size_t)n * kSlice];
        if (jr >= N) continue;
        const T *rj = qa + kQA * (size_t)(jr >= 0 ? jr : ~jr);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            qmin[k] = tmin(qmin[k], rj[k]);
            qmax[k] = tmax(qmax[k], rj[k]);
        }
    }
    T ph[5];
    const bool risky = limiter_passes<T, KIND>(N, deg, base, nbr, qa, qc, xc, yc, zc, gr, qmin, qmax, vol[c], g, frac,
                                               k3, ph);
    if (risky) atomicAdd(risk, 1);
#pragma unroll
    for (int k = 0; k < 5; k++) rb[kPHI + k] = ph[k];
}

// Green-Gauss gradient and limiter of one node in one pass over its
// incidences: the edge-ordered Green-Gauss sums and the neighbour extrema are
// gathered together, the gradient stays in registers for the limiter passes.
// Bit-identical to k_gg_grad followed by k_limiter.
// Staging slot of a walk over node records: the qa record (q and
// coordinates) of the next neighbour, its area vector when AREA, and the
// index of the incidence after it.
template <class T>
struct QaStage {
    static constexpr int P = 16 / (int)sizeof(T);  // values per 16-byte chunk
    static constexpr int QC = 8 / P;               // chunks of a qa record (8 values)
    static constexpr int AW = (int)sizeof(T) / 4;  // dwords per area component
    static constexpr int AUX = 1 + 3 * AW;
};

template <class T, bool AREA>
__device__ __forceinline__ void stage_qa(const T *qa, const T *isx, const T *isy, const T *isz, const int *nbr,
                                         int jr, size_t idx, bool has_next, uint4 (*slot)[64], uint32_t (*aux)[64]) {
    using S = QaStage<T>;
    const int j = jr >= 0 ? jr : ~jr;  // ghosts N + b have qa rows too
    const char *q = reinterpret_cast<const char *>(qa + kQA * (size_t)j);
#pragma unroll
    for (int c = 0; c < S::QC; c++) dma16<0>(q + 16 * c, &slot[c][0]);
    if (AREA) {
        const char *ax = reinterpret_cast<const char *>(isx + idx), *ay = reinterpret_cast<const char *>(isy + idx),
                   *az = reinterpret_cast<const char *>(isz + idx);
#pragma unroll
        for (int w = 0; w < S::AW; w++) {
            dma4<kPolNT>(ax + 4 * w, &aux[1 + w][0]);
            dma4<kPolNT>(ay + 4 * w, &aux[1 + S::AW + w][0]);
            dma4<kPolNT>(az + 4 * w, &aux[1 + 2 * S::AW + w][0]);
        }
    }
    if (has_next) dma4<kPolNT>(nbr + idx + kSlice, &aux[0][0]);
}

template <class T>
__device__ __forceinline__ T aux_word(const uint32_t (*aux)[64], int row, int lane) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t v = (uint64_t)aux[row][lane] | ((uint64_t)aux[row + 1][lane] << 32);
        return __builtin_bit_cast(T, v);
    } else {
        return __builtin_bit_cast(T, aux[row][lane]);
    }
}

// Walks one node's incidences (SELL entries base, base + 64, ...) in order
// with the next neighbour's record staged in LDS; body(jr, r, s0, s1, s2)
// receives the neighbour's record r[8] (and the area vector when AREA).
template <class T, bool AREA, class F>
__device__ __forceinline__ void staged_walk(const T *qa, const T *isx, const T *isy, const T *isz,
                                            const int *nbr, size_t base, int deg, uint4 (*slot)[64],
                                            uint32_t (*aux)[64], int lane, F body) {
    using S = QaStage<T>;
    int jr = 0;
    if (deg > 0) {
        jr = ld_stream(nbr + base);
        stage_qa<T, AREA>(qa, isx, isy, isz, nbr, jr, base, deg > 1, slot, aux);
    }
    size_t idx = base;
    for (int k = 0; k < deg; k++, idx += kSlice) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        T r[8];
#pragma unroll
        for (int e = 0; e < 8; e++) r[e] = reinterpret_cast<const T *>(&slot[e / S::P][lane])[e % S::P];
        T s0 = T(0.0), s1 = T(0.0), s2 = T(0.0);
        if (AREA) {
            s0 = aux_word<T>(aux, 1, lane);
            s1 = aux_word<T>(aux, 1 + S::AW, lane);
            s2 = aux_word<T>(aux, 1 + 2 * S::AW, lane);
        }
        const int jr_next = (int)aux[0][lane];
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        if (k + 1 < deg) stage_qa<T, AREA>(qa, isx, isy, isz, nbr, jr_next, idx + kSlice, k + 2 < deg, slot, aux);
        body(jr, r, s0, s1, s2);
        jr = jr_next;
    }
}

template <class T, int KIND, bool STAGE, bool F0>
__global__ __launch_bounds__(kBlock) void k_gg_limiter(int N, const int *__restrict__ ip, const int *__restrict__ sptr,
                                                       const int *__restrict__ nbr, const T *__restrict__ isx,
                                                       const T *__restrict__ isy, const T *__restrict__ isz,
                                                       const T *__restrict__ vol, const T *__restrict__ qa,
                                                       T *__restrict__ qb, T g, T frac, T k3, int *risk) {
    const int c = xcd_block() * kBlock + threadIdx.x;
    if (c >= N) return;
    const T *ra = qa + kQA * (size_t)c;
    T qc[5], xc = ra[5], yc = ra[6], zc = ra[7];
#pragma unroll
    for (int k = 0; k < 5; k++) qc[k] = ra[k];
    T gs[15], qmin[5], qmax[5];
#pragma unroll
    for (int f = 0; f < 15; f++) gs[f] = T(0.0);
#pragma unroll
    for (int k = 0; k < 5; k++) qmin[k] = qmax[k] = qc[k];
    const int deg = ip[c + 1] - ip[c];
    const size_t base = (size_t)sptr[c / kSlice] + (c % kSlice);
    if constexpr (STAGE) {
        // the same three walks with every neighbour record staged in LDS
        __shared__ uint4 stage[kBlock / 64][QaStage<T>::QC][64];
        __shared__ uint32_t stage_aux[kBlock / 64][QaStage<T>::AUX][64];
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
        uint4(*slot)[64] = stage[wv];
        uint32_t(*aux)[64] = stage_aux[wv];
        const int lane = threadIdx.x % 64;
        staged_walk<T, true>(qa, isx, isy, isz, nbr, base, deg, slot, aux, lane,
                             [&](int jr, const T *r, T s0, T s1, T s2) {
            // branch-free: a ghost adds the node's own value; an interior
            // neighbour adds (as L) or subtracts (as R) the face average, and
            // a - b equals a + (-b) exactly
            const bool ghost = jr >= N, neg = jr < 0;
#pragma unroll
            for (int f = 0; f < 5; f++) {
                const T qj = r[f];
                qmin[f] = ghost ? qmin[f] : tmin(qmin[f], qj);
                qmax[f] = ghost ? qmax[f] : tmax(qmax[f], qj);
                const T qf = ghost ? qc[f] : T(0.5) * (qc[f] + qj);
                const T a0 = qf * s0, a1 = qf * s1, a2 = qf * s2;
                gs[3 * f + 0] += neg ? -a0 : a0;
                gs[3 * f + 1] += neg ? -a1 : a1;
                gs[3 * f + 2] += neg ? -a2 : a2;
            }
        });
        const T v = vol[c];
        T gr[15];
        T *rb = qb + kQB * (size_t)c;
#pragma unroll
        for (int f = 0; f < 15; f++) {
            gr[f] = grad_of(gs[f], v);
            rb[f] = gr[f];
        }
        T ph[5], pd[5];
#pragma unroll
        for (int k = 0; k < 5; k++) ph[k] = pd[k] = T(1.0);
        T eps2 = k3;
        eps2 = T(6.0) * eps2 * v / T(M_PI);
        T r2 = T(0.0);  // longest half edge, squared
        staged_walk<T, false>(qa, isx, isy, isz, nbr, base, deg, slot, aux, lane,
                              [&](int jr, const T *r, T, T, T) {
            if (jr >= N) return;
            limiter_nbr<T, KIND, F0>(r, r[5], r[6], r[7], qc, xc, yc, zc, gr, qmin, qmax, eps2, frac, ph, pd);
            const T Rx = T(0.5) * (r[5] - xc), Ry = T(0.5) * (r[6] - yc), Rz = T(0.5) * (r[7] - zc);
            r2 = tmax(r2, Rx * Rx + Ry * Ry + Rz * Rz);
        });
        limiter_finish<T, KIND>(ph, pd);
        // the pressure pass, unless the bounds show it would change nothing
        // (a lane with nothing to check walks no incidences)
        bool zeroed = false, risky = false;
        const bool skip = faces_safe<T, F0>(qc, gr, qmin, qmax, ph, frac, sqrt(r2), g);
        staged_walk<T, false>(qa, isx, isy, isz, nbr, base, skip ? 0 : deg, slot, aux, lane,
                              [&](int jr, const T *r, T, T, T) {
            if (jr >= N) return;
            pressure_nbr<T, F0>(r, r[5], r[6], r[7], qc, xc, yc, zc, gr, frac, g, ph, zeroed, risky);
        });
        if (pressure_final<T>(qc, g, zeroed, risky)) atomicAdd(risk, 1);
#pragma unroll
        for (int k = 0; k < 5; k++) rb[kPHI + k] = ph[k];
        return;
    }
    size_t idx = base;
    for (int k = 0; k < deg; k++, idx += kSlice) {
        const int jr = nbr[idx];
        const T s0 = ld_stream(isx + idx), s1 = ld_stream(isy + idx), s2 = ld_stream(isz + idx);
        // branch-free as in the staged walk (ghost rows of qa exist)
        const bool ghost = jr >= N, neg = jr < 0;
        const T *rj = qa + kQA * (size_t)(neg ? ~jr : jr);
#pragma unroll
        for (int f = 0; f < 5; f++) {
            const T qj = rj[f];
            qmin[f] = ghost ? qmin[f] : tmin(qmin[f], qj);
            qmax[f] = ghost ? qmax[f] : tmax(qmax[f], qj);
            const T qf = ghost ? qc[f] : T(0.5) * (qc[f] + qj);
            const T a0 = qf * s0, a1 = qf * s1, a2 = qf * s2;
            gs[3 * f + 0] += neg ? -a0 : a0;
            gs[3 * f + 1] += neg ? -a1 : a1;
            gs[3 * f + 2] += neg ? -a2 : a2;
        }
    }
    const T v = vol[c];
    T gr[15];
    T *rb = qb + kQB * (size_t)c;
#pragma unroll
    for (int f = 0; f < 15; f++) {
        gr[f] = grad_of(gs[f], v);
        rb[f] = gr[f];
    }
    T ph[5];
    const bool risky = limiter_passes<T, KIND, F0>(N, deg, base, nbr, qa, qc, xc, yc, zc, gr, qmin, qmax, v, g, frac, k3, ph);
    if (risky) atomicAdd(risk, 1);
#pragma unroll
    for (int k = 0; k < 5; k++) rb[kPHI + k] = ph[k];
}

// Face state of node c towards neighbour n with limiter ph (Limiters.cpp:327-343).
template <class T>
__device__ __forceinline__ void face_state(const T *ra, const T *rbg, const T *ph, const T *rn, T frac, T *Q) {
    const T Rx = T(0.5) * (rn[5] - ra[5]), Ry = T(0.5) * (rn[6] - ra[6]), Rz = T(0.5) * (rn[7] - ra[7]);
#pragma unroll
    for (int k = 0; k < 5; k++)
        Q[k] = ra[k] + ph[k] * (frac * (rn[k] - ra[k]) + (T(1.0) - frac) * (rbg[3 * k] * Rx + rbg[3 * k + 1] * Ry + rbg[3 * k + 2] * Rz));
}

// Pressure of the Roe average of two face states (Compute_RoeVariables,
// Euler3D/src/Roe_Fluxes.cpp:3609-3634, and the pressure of
// Compute_Limiter_RoePressureCorrection, Euler3D/src/Limiters.cpp:345-346).
template <class T>
__device__ __forceinline__ T roe_face_pressure(T g, const T *QL, const T *QR) {
    Eos<T> L, R;
    eos_face(g, QL, T(0.0), T(0.0), T(0.0), L);
    eos_face(g, QR, T(0.0), T(0.0), T(0.0), R);
    const T rho = sqrt(R.rho * L.rho);
    const T sigma = rho / (L.rho + rho);
    const T u = L.u + sigma * (R.u - L.u);
    const T v = L.v + sigma * (R.v - L.v);
    const T w = L.w + sigma * (R.w - L.w);
    const T ht = L.ht + sigma * (R.ht - L.ht);
    const T q0 = rho, q1 = rho * u, q2 = rho * v, q3 = rho * w;
    const T q4 = (rho / g) * (ht + T(0.5) * (g - T(1.0)) * (u * u + v * v + w * w));
    return (g - T(1.0)) * (q4 - (T(0.5) / q0) * (q1 * q1 + q2 * q2 + q3 * q3));
}

// Roe pressure correction (Euler3D/src/Limiters.cpp:71-110, 311-356).  The
// reference sweeps the nodes in order and tests each node against each of
// its neighbours in node2node order, zeroing the limiters of both when the
// Roe-averaged pressure of their face states is negative, with the limiters
// as the sweep has left them.  During the sweep a node's limiter is its value
// from the limiter pass or zero, so a test can only act if the pressure is
// negative for one of the four combinations of those values.
//   k_roe_candidates marks every test that can act (one thread per node, in
//     parallel) and flags its block of 4096 reference indices;
//   k_roe_replay (one wavefront) replays the reference's sweep over just the
//     marked tests, in the reference's order, with the limiters as they
//     evolve: the reference's result exactly.
// Both return at once when the limiter pass found no near-vacuum face state,
// the normal case: the first test of the sweep cannot act then (the Roe
// average of two states of positive pressure has positive pressure, and the
// margin covers rounding), and so none can.
template <class T>
__global__ __launch_bounds__(kBlock) void k_roe_candidates(int N, const int *__restrict__ order,
                                                           const int *__restrict__ n2n_ia,
                                                           const int *__restrict__ n2n_ja, const T *__restrict__ qa,
                                                           const T *__restrict__ qb, T g, T frac, uint64_t *cand,
                                                           int *blk, int epoch, const int *risk) {
    if (*risk == 0) return;
    const int c = xcd_block() * kBlock + threadIdx.x;
    if (c >= N) return;
    const T *ra = qa + kQA * (size_t)c;
    const T *rb = qb + kQB * (size_t)c;
    const T zero[5] = {T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
    uint64_t mask = 0;
    for (int k = n2n_ia[c], n = 0; k < n2n_ia[c + 1]; k++, n++) {
        const int j = n2n_ja[k];
        const T *rj = qa + kQA * (size_t)j;
        const T *rbj = qb + kQB * (size_t)j;
        T QL[5], QR[5], ZL[5], ZR[5];
        face_state(ra, rb, rb + kPHI, rj, frac, QL);
        face_state(rj, rbj, rbj + kPHI, ra, frac, QR);
        face_state(ra, rb, zero, rj, frac, ZL);
        face_state(rj, rbj, zero, ra, frac, ZR);
        const bool can = roe_face_pressure(g, QL, QR) < T(0.0) || roe_face_pressure(g, ZL, QR) < T(0.0) ||
                         roe_face_pressure(g, QL, ZR) < T(0.0) || roe_face_pressure(g, ZL, ZR) < T(0.0);
        if (can) mask |= uint64_t(1) << (n < 63 ? n : 63);
    }
    const int a = order[c];
    cand[a] = mask;
    if (mask) blk[a >> 12] = epoch;
}

template <class T>
__global__ __launch_bounds__(64) void k_roe_replay(int N, const int *__restrict__ iperm,
                                                   const int *__restrict__ n2n_ia, const int *__restrict__ n2n_ja,
                                                   const T *__restrict__ qa, T *qb, T g, T frac,
                                                   const uint64_t *__restrict__ cand, const int *__restrict__ blk,
                                                   int epoch, const int *risk) {
    if (*risk == 0) return;
    const int lane = threadIdx.x;
    for (int b = 0; b < (N + 4095) / 4096; b++) {
        if (blk[b] != epoch) continue;
        for (int a0 = b * 4096; a0 < N && a0 < (b + 1) * 4096; a0 += 64) {
            const int a = a0 + lane;
            uint64_t act = __ballot(a < N && cand[a] != 0);
            // one lane applies the tests, in order, so each reads the limiters
            // every earlier test left
            if (lane != 0) continue;
            while (act) {
                const int l = __builtin_ctzll(act);
                act &= act - 1;
                const uint64_t mk = cand[a0 + l];
                const int c = iperm[a0 + l];
                const T *ra = qa + kQA * (size_t)c;
                T *rb = qb + kQB * (size_t)c;
                for (int k = n2n_ia[c], n = 0; k < n2n_ia[c + 1]; k++, n++) {
                    if (!((mk >> (n < 63 ? n : 63)) & 1)) continue;
                    const int j = n2n_ja[k];
                    const T *rj = qa + kQA * (size_t)j;
                    T *rbj = qb + kQB * (size_t)j;
                    T QL[5], QR[5];
                    face_state(ra, rb, rb + kPHI, rj, frac, QL);
                    face_state(rj, rbj, rbj + kPHI, ra, frac, QR);
                    if (roe_face_pressure(g, QL, QR) < T(0.0)) {
#pragma unroll
                        for (int v = 0; v < 5; v++) rb[kPHI + v] = rbj[kPHI + v] = T(0.0);
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Residual: Compute_Residual_Roe_New (Euler3D/src/FluxTest.cpp:7968-8151),
// Compute_Residual_HLLC (Euler3D/src/HLLC_Fluxes.cpp:519-573), AUSM+.
// ---------------------------------------------------------------------------
template <class T>
struct ResArgs {
    int N, NT;
    const int *ip, *sptr, *nbr;
    const T *isx, *isy, *isz;
    const T *qa, *qb;
    T *res;
    T g, frac;
    int efix;
    unsigned long long *eig;
    int *flag;
    // union of cases: failure bits per case and the case of each node
    // (internal numbering; cases hold contiguous ranges), NULL for one case
    int *case_flag;
    const int *ecase;
};

// Neighbour data one incidence needs: q and coordinates, and for second
// order the gradient and limiter.  The residual loop stages incidence k+1 in
// LDS with direct global->LDS loads while the flux of incidence k is
// computed: the neighbour's record (16 bytes per lane per load), the area
// vector and the index of incidence k+2 (4 bytes per lane per load).  No VGPR
// holds an in-flight load, so nothing in the loop waits on the prefetch
// before the next incidence begins.  Each wave owns one staging slot; a
// wave-wide load of width W to row r puts lane l's bytes at row r, column l.
template <class T, int ORDER>
struct NbrStage {
    static constexpr int P = 16 / (int)sizeof(T);       // values per 16-byte chunk
    static constexpr int NQ = ORDER == 2 ? 8 : 5;       // q (and x, y, z)
    static constexpr int QC = (NQ + P - 1) / P;         // chunks of the qa record
    static constexpr int BC = ORDER == 2 ? 20 / P : 0;  // chunks of gradient + limiter
    static constexpr int CH = QC + BC;
    static constexpr int AW = (int)sizeof(T) / 4;       // dwords per area component
    static constexpr int AUX = 1 + 3 * AW;              // next index, area vector
};

template <class T, int ORDER>
struct NbrRec {
    T q[NbrStage<T, ORDER>::NQ];
    T b[ORDER == 2 ? 20 : 1];
};

// Stages incidence `idx` with neighbour jr: the neighbour's record, the area
// vector and, if has_next, the neighbour index of the incidence after it.
// qa rows hold 8 values and qb rows 24, so every chunk lies inside its row;
// ghosts (j >= N) have no qb row and load only q.
template <class T, int ORDER>
__device__ __forceinline__ void stage_inc(const ResArgs<T> &a, int jr, size_t idx, bool has_next,
                                          uint4 (*slot)[64], uint32_t (*aux)[64]) {
    using S = NbrStage<T, ORDER>;
    const int j = jr >= 0 ? jr : ~jr;
    const char *q = reinterpret_cast<const char *>(a.qa + kQA * (size_t)j);
#pragma unroll
    for (int c = 0; c < S::QC; c++) dma16<0>(q + 16 * c, &slot[c][0]);
    if (ORDER == 2 && j < a.N) {
        const char *g = reinterpret_cast<const char *>(a.qb + kQB * (size_t)j);
#pragma unroll
        for (int c = 0; c < S::BC; c++) dma16<0>(g + 16 * c, &slot[S::QC + c][0]);
    }
    const char *ax = reinterpret_cast<const char *>(a.isx + idx), *ay = reinterpret_cast<const char *>(a.isy + idx),
               *az = reinterpret_cast<const char *>(a.isz + idx);
#pragma unroll
    for (int w = 0; w < S::AW; w++) {
        dma4<kPolNT>(ax + 4 * w, &aux[1 + w][0]);
        dma4<kPolNT>(ay + 4 * w, &aux[1 + S::AW + w][0]);
        dma4<kPolNT>(az + 4 * w, &aux[1 + 2 * S::AW + w][0]);
    }
    if (has_next) dma4<kPolNT>(a.nbr + idx + kSlice, &aux[0][0]);
}

template <class T, int ORDER, bool LIM>
__device__ __forceinline__ void read_nbr(const uint4 (*slot)[64], int lane, NbrRec<T, ORDER> &r) {
    using S = NbrStage<T, ORDER>;
#pragma unroll
    for (int e = 0; e < S::NQ; e++) r.q[e] = reinterpret_cast<const T *>(&slot[e / S::P][lane])[e % S::P];
    if (ORDER == 2) {
#pragma unroll
        for (int e = 0; e < 20; e++) {
            const T v = reinterpret_cast<const T *>(&slot[S::QC + e / S::P][lane])[e % S::P];
            r.b[ORDER == 2 ? e : 0] = (e >= 15 && !LIM) ? T(1.0) : v;
        }
    }
}

template <class T>
__device__ __forceinline__ T aux_val(const uint32_t (*aux)[64], int row, int lane) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t v = (uint64_t)aux[row][lane] | ((uint64_t)aux[row + 1][lane] << 32);
        return __builtin_bit_cast(T, v);
    } else {
        return __builtin_bit_cast(T, aux[row][lane]);
    }
}

// Neighbour record views for NodeResidual::add.  q(k): state and
// coordinates; b(f): 15 gradient components, then 5 limiter values.
template <class T, bool LIM>
struct GlobalRec {  // read where used, so the compiler schedules the gathers
    const T *rj, *rbj;
    __device__ __forceinline__ T q(int k) const { return rj[k]; }
    __device__ __forceinline__ T b(int f) const { return f < 15 ? rbj[f] : (LIM ? rbj[kPHI + f - 15] : T(1.0)); }
};
template <class T, int ORDER>
struct RegRec {  // already staged into registers
    const NbrRec<T, ORDER> &r;
    __device__ __forceinline__ T q(int k) const { return r.q[k]; }
    __device__ __forceinline__ T b(int f) const { return r.b[ORDER == 2 ? f : 0]; }
};

// One node's residual: its own state, gradient and limiter, and the sums of
// the fluxes of its incidences, added in the order they are visited.
template <class T, int FLUX, int ORDER, bool LIM, bool STATS>
struct NodeResidual {
    T qi[5], xi, yi, zi, gi[15], pi[5];
    T ac[5], ad[5], dt;
    T emin[3], emax[3];
    int bad;

    __device__ __forceinline__ void init(const ResArgs<T> &a, int i) {
        const T *ra = a.qa + kQA * (size_t)i;
#pragma unroll
        for (int k = 0; k < 5; k++) qi[k] = ra[k];
        xi = yi = zi = T(0.0);
        if (ORDER == 2) {
            xi = ra[5];
            yi = ra[6];
            zi = ra[7];
            const T *rb = a.qb + kQB * (size_t)i;
#pragma unroll
            for (int f = 0; f < 15; f++) gi[f] = rb[f];
#pragma unroll
            for (int k = 0; k < 5; k++) pi[k] = LIM ? rb[kPHI + k] : T(1.0);
        }
#pragma unroll
        for (int k = 0; k < 5; k++) ac[k] = ad[k] = T(0.0);
        dt = T(0.0);
#pragma unroll
        for (int t = 0; t < 3; t++) {
            emin[t] = T(DBL_MAX);
            emax[t] = T(0.0);
        }
        bad = 0;
    }

    // Adds the flux of the incidence with neighbour jr (j: this node is L;
    // ~j: this node is R; N+b: boundary edge b) and area vector s.
    template <class Rec>
    __device__ __forceinline__ void add(const ResArgs<T> &a, int jr, const Rec &cur, T s0, T s1, T s2) {
        const bool iL = jr >= 0;  // ghosts (jr >= N) always pair with a physical L
        const int j = iL ? jr : ~jr;
        T QL[5], QR[5];
        if (ORDER == 2 && j < a.N) {
            // Compute_SecondOrderReconstructQ, Euler3D/src/HigherOrderReconstructQ.cpp:20-71
            const T xj = cur.q(5), yj = cur.q(6), zj = cur.q(7);
            const T RxL = iL ? T(0.5) * (xj - xi) : T(0.5) * (xi - xj);
            const T RyL = iL ? T(0.5) * (yj - yi) : T(0.5) * (yi - yj);
            const T RzL = iL ? T(0.5) * (zj - zi) : T(0.5) * (zi - zj);
            const T RxR = -RxL, RyR = -RyL, RzR = -RzL;
#pragma unroll
            for (int k2 = 0; k2 < 5; k2++) {
                const T qj = cur.q(k2);
                const T pj = cur.b(15 + k2);
                const T gjx = cur.b(3 * k2), gjy = cur.b(3 * k2 + 1), gjz = cur.b(3 * k2 + 2);
                const T qL = iL ? qi[k2] : qj, qR = iL ? qj : qi[k2];
                const T pL = iL ? pi[k2] : pj, pR = iL ? pj : pi[k2];
                const T gLx = iL ? gi[3 * k2] : gjx, gLy = iL ? gi[3 * k2 + 1] : gjy, gLz = iL ? gi[3 * k2 + 2] : gjz;
                const T gRx = iL ? gjx : gi[3 * k2], gRy = iL ? gjy : gi[3 * k2 + 1], gRz = iL ? gjz : gi[3 * k2 + 2];
                QL[k2] = qL + pL * (a.frac * (qR - qL) + (T(1.0) - a.frac) * (gLx * RxL + gLy * RyL + gLz * RzL));
                QR[k2] = qR + pR * (a.frac * (qL - qR) + (T(1.0) - a.frac) * (gRx * RxR + gRy * RyR + gRz * RzR));
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < 5; k2++) {
                const T qj = cur.q(k2);
                QL[k2] = iL ? qi[k2] : qj;
                QR[k2] = iL ? qj : qi[k2];
            }
        }
        T cf[5], df[5], lam;
        if (FLUX == 0) {
            T E0, E1, E2;
            lam = roe_flux(a.g, a.efix, s0, s1, s2, QL, QR, cf, df, E0, E1, E2);
            if (STATS) {
                emin[0] = tmin(emin[0], E0);
                emax[0] = tmax(emax[0], E0);
                emin[1] = tmin(emin[1], E1);
                emax[1] = tmax(emax[1], E1);
                emin[2] = tmin(emin[2], E2);
                emax[2] = tmax(emax[2], E2);
            }
        } else if (FLUX == 1) {
            lam = hllc_flux(a.g, s0, s1, s2, QL, QR, cf, bad);
        } else {
            lam = ausmp_flux(a.g, s0, s1, s2, QL, QR, cf);
        }
        if (iL) {
#pragma unroll
            for (int k2 = 0; k2 < 5; k2++) {
                ac[k2] += cf[k2];
                if (FLUX == 0) ad[k2] += df[k2];
            }
        } else {
#pragma unroll
            for (int k2 = 0; k2 < 5; k2++) {
                ac[k2] -= cf[k2];
                if (FLUX == 0) ad[k2] -= df[k2];
            }
        }
        dt += lam;
    }

    __device__ __forceinline__ void store(const ResArgs<T> &a, int i, int deg) {
        T *o = a.res + kRES * (size_t)i;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            o[k] = ac[k];
            o[5 + k] = ad[k];
        }
        o[10] = dt;
        o[11] = T(0.0);
        if (bad) {
            flag_set(a.flag, 1);
            if (a.case_flag) atomicOr(a.case_flag + a.ecase[i], 4);
        }
        if (STATS && deg > 0) {
            for (int t = 0; t < 3; t++) {
                atomic_min_nonneg(a.eig + 2 * t, (double)emin[t]);
                atomic_max_nonneg(a.eig + 2 * t + 1, (double)emax[t]);
            }
        }
    }
};

// One thread per node, walking its incidence list in the reference's
// accumulation order.  STAGE selects how neighbour records arrive: straight
// into registers (the compiler schedules the gathers), or staged one
// incidence ahead through LDS.  Which is faster depends on the flux: the Roe
// flux has the most arithmetic to overlap with direct gathers, HLLC and AUSM+
// gain from the staging (measured on the 160^3 cube, FP64 and FP32).
template <class T, int FLUX, int ORDER, bool LIM, bool STATS, bool STAGE>
__global__ __launch_bounds__(kBlock) void k_residual(ResArgs<T> a) {
    using S = NbrStage<T, ORDER>;
    const int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.N) return;
    NodeResidual<T, FLUX, ORDER, LIM, STATS> nr;
    nr.init(a, i);
    const int deg = a.ip[i + 1] - a.ip[i];
    size_t idx = (size_t)a.sptr[i / kSlice] + (i % kSlice);
    if constexpr (!STAGE) {
        for (int k = 0; k < deg; k++, idx += kSlice) {
            const int jr = ld_stream(a.nbr + idx);
            const T s0 = ld_stream(a.isx + idx), s1 = ld_stream(a.isy + idx), s2 = ld_stream(a.isz + idx);
            const int j = jr >= 0 ? jr : ~jr;
            const GlobalRec<T, LIM> cur{a.qa + kQA * (size_t)j, a.qb + kQB * (size_t)j};  // qb row used only if j < N
            nr.add(a, jr, cur, s0, s1, s2);
        }
    } else {
        __shared__ uint4 stage[kBlock / 64][S::CH][64];
        __shared__ uint32_t stage_aux[kBlock / 64][S::AUX][64];
        const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
        uint4(*slot)[64] = stage[wv];
        uint32_t(*aux)[64] = stage_aux[wv];
        const int lane = threadIdx.x % 64;
        int jr = 0;
        if (deg > 0) {
            jr = ld_stream(a.nbr + idx);
            stage_inc<T, ORDER>(a, jr, idx, deg > 1, slot, aux);
        }
        for (int k = 0; k < deg; k++, idx += kSlice) {
            // incidence k has landed (LDS-DMA completion counts in vmcnt,
            // which the compiler does not track for these loads); it is read
            // out, and the reads complete before incidence k+1 is staged over it
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            NbrRec<T, ORDER> cur;
            read_nbr<T, ORDER, LIM>(slot, lane, cur);
            const T s0 = aux_val<T>(aux, 1, lane), s1 = aux_val<T>(aux, 1 + S::AW, lane),
                    s2 = aux_val<T>(aux, 1 + 2 * S::AW, lane);
            const int jr_next = (int)aux[0][lane];
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            if (k + 1 < deg) stage_inc<T, ORDER>(a, jr_next, idx + kSlice, k + 2 < deg, slot, aux);
            nr.add(a, jr, RegRec<T, ORDER>{cur}, s0, s1, s2);
            jr = jr_next;
        }
    }
    nr.store(a, i, deg);
}

// ---------------------------------------------------------------------------
// Flux-once residual (second order).  Workgroup p owns a patch of at most
// kBlock consecutive internal nodes (a compact piece of the Morton order;
// patches restart at each case of a union) and its work edges: each interior edge with both ends in the patch
// once, each edge to another patch (both patches evaluate it, each for its
// own end) and each boundary edge of a patch node.  Chunks of kBlock work
// edges are evaluated one per thread into LDS; then every node adds the
// chunk's fluxes of its incidences, whose slots it lists in ascending order
// (held in LDS).  The work edges are in the reference's edge order, so each
// node adds its fluxes in the order its incidence list gives (that of the
// reference's edge loop), and a flux evaluated once for both ends has the
// bits of the node-centred kernel's two evaluations (the R end's face state
// q_R - phi (g . R) equals q_R + phi (g . (-R)) exactly): the exact build's
// results are bit-identical to the node-centred kernel's.  Interior fluxes are
// evaluated once per patch instead of once per end, and every flux stays in
// the patch's LDS.
// ---------------------------------------------------------------------------
template <class T>
struct PatchArgs {
    int N, npatch;
    const int *pstart;  // [npatch + 1] first node of each patch (at most kBlock nodes each)
    const int *ip, *sptr;
    // SELL (as the incidence lists): patch-local work-edge slot of each
    // incidence of a node, ascending; bit 31 set where the node is the R end
    const int *slot;
    const int *wptr;    // [npatch + 1] first work edge of each patch
    const int *wl, *wr; // L and R end of each work edge (R = N + b: ghost of boundary edge b)
    const T *ws;        // [W][3] area vector, oriented L -> R (outward for boundary edges)
    const T *qa, *qb;
    T *res;
    T g, frac;
    int efix;
    unsigned long long *eig;
    int *flag, *case_flag;
    const int *ecase;
};

// Flux of work edge w: conv, diss and the time-step term area * max |lambda|.
// Second-order face states on interior edges (Compute_SecondOrderReconstructQ,
// Euler3D/src/HigherOrderReconstructQ.cpp:20-71), first order on boundary edges.
template <class T, int FLUX, int ORDER, bool LIM, bool STATS>
__device__ __forceinline__ T edge_flux(const PatchArgs<T> &a, int w, T *cf, T *df, T *emin, T *emax, int &bad) {
    const int l = a.wl[w], r = a.wr[w];
    const T *sv = a.ws + 3 * (size_t)w;
    const T s0 = sv[0], s1 = sv[1], s2 = sv[2];
    const T *ql = a.qa + kQA * (size_t)l, *qr = a.qa + kQA * (size_t)r;
    T QL[5], QR[5];
    if (ORDER == 2 && r < a.N) {
        const T *bl = a.qb + kQB * (size_t)l, *br = a.qb + kQB * (size_t)r;
        const T Rx = T(0.5) * (qr[5] - ql[5]), Ry = T(0.5) * (qr[6] - ql[6]), Rz = T(0.5) * (qr[7] - ql[7]);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const T pl = LIM ? bl[kPHI + k] : T(1.0), pr = LIM ? br[kPHI + k] : T(1.0);
            const T dq = qr[k] - ql[k];
            QL[k] = ql[k] + pl * (a.frac * dq + (T(1.0) - a.frac) * (bl[3 * k] * Rx + bl[3 * k + 1] * Ry + bl[3 * k + 2] * Rz));
            QR[k] = qr[k] - pr * (a.frac * dq + (T(1.0) - a.frac) * (br[3 * k] * Rx + br[3 * k + 1] * Ry + br[3 * k + 2] * Rz));
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) {
            QL[k] = ql[k];
            QR[k] = qr[k];
        }
    }
    if (FLUX == 0) {
        T E0, E1, E2;
        const T lam = roe_flux(a.g, a.efix, s0, s1, s2, QL, QR, cf, df, E0, E1, E2);
        if (STATS) {
            emin[0] = tmin(emin[0], E0);
            emax[0] = tmax(emax[0], E0);
            emin[1] = tmin(emin[1], E1);
            emax[1] = tmax(emax[1], E1);
            emin[2] = tmin(emin[2], E2);
            emax[2] = tmax(emax[2], E2);
        }
        return lam;
    }
#pragma unroll
    for (int k = 0; k < 5; k++) df[k] = T(0.0);
    if (FLUX == 1) return hllc_flux(a.g, s0, s1, s2, QL, QR, cf, bad);
    return ausmp_flux(a.g, s0, s1, s2, QL, QR, cf);
}

template <class T, int FLUX, int ORDER, bool LIM, bool STATS>
__global__ __launch_bounds__(kBlock) void k_residual_patch(PatchArgs<T> a) {
    constexpr int NV = FLUX == 0 ? 11 : 6;  // conv, [diss,] area * max |lambda|
    constexpr int kSlots = 24;             // slots of a node's list held in LDS (longer lists read the rest)
    __shared__ T buf[NV][kBlock];
    __shared__ int sslot[kSlots][kBlock];
    const int p = xcd_block();
    if (p >= a.npatch) return;  // padding blocks of the XCD-rounded grid
    const int t = threadIdx.x;
    const int i = a.pstart[p] + t;
    const bool own = i < a.pstart[p + 1];
    const int w0 = a.wptr[p], W = a.wptr[p + 1] - w0;
    int deg = 0;
    size_t idx = 0;
    if (own) {
        deg = a.ip[i + 1] - a.ip[i];
        idx = (size_t)a.sptr[i / kSlice] + (i % kSlice);
        // the node's slot list, loaded at once (independent loads); only this
        // thread reads its column, so no barrier is needed
        for (int k = 0; k < deg && k < kSlots; k++) sslot[k][t] = ld_stream(a.slot + idx + (size_t)k * kSlice);
    }
    auto slot_at = [&](int k) {
        return k >= deg ? INT_MAX : (k < kSlots ? sslot[k][t] : ld_stream(a.slot + idx + (size_t)k * kSlice));
    };
    T ac[5], ad[5], dt = T(0.0), emin[3], emax[3];
#pragma unroll
    for (int k = 0; k < 5; k++) ac[k] = ad[k] = T(0.0);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        emin[k] = T(DBL_MAX);
        emax[k] = T(0.0);
    }
    int bad = 0, k = 0, nxt = slot_at(0);
    for (int c0 = 0; c0 < W; c0 += kBlock) {
        if (c0 + t < W) {
            T cf[5], df[5];
            int wbad = 0;
            const T lam = edge_flux<T, FLUX, ORDER, LIM, STATS>(a, w0 + c0 + t, cf, df, emin, emax, wbad);
            if (wbad) {  // no HLLC wave configuration: flag the mesh and the edge's case
                bad = 1;
                if (a.case_flag) atomicOr(a.case_flag + a.ecase[a.wl[w0 + c0 + t]], 4);
            }
#pragma unroll
            for (int v = 0; v < 5; v++) {
                buf[v][t] = cf[v];
                if (FLUX == 0) buf[5 + v][t] = df[v];
            }
            buf[NV - 1][t] = lam;
        }
        __syncthreads();
        // this node's incidences among the chunk's work edges, in slot order
        while ((nxt & 0x7fffffff) < c0 + kBlock) {
            const int s = (nxt & 0x7fffffff) - c0;
            const bool neg = nxt < 0;
#pragma unroll
            for (int v = 0; v < 5; v++) {
                const T f = buf[v][s];
                ac[v] += neg ? -f : f;
                if (FLUX == 0) {
                    const T d = buf[5 + v][s];
                    ad[v] += neg ? -d : d;
                }
            }
            dt += buf[NV - 1][s];
            nxt = slot_at(++k);
        }
        __syncthreads();  // every node has read the chunk before the next one overwrites it
    }
    if (own) {
        T *o = a.res + kRES * (size_t)i;
#pragma unroll
        for (int v = 0; v < 5; v++) {
            o[v] = ac[v];
            o[5 + v] = ad[v];
        }
        o[10] = dt;
        o[11] = T(0.0);
    }
    if (bad) flag_set(a.flag, 1);
    if (STATS && W > t) {
        for (int v = 0; v < 3; v++) {
            atomic_min_nonneg(a.eig + 2 * v, (double)emin[v]);
            atomic_max_nonneg(a.eig + 2 * v + 1, (double)emax[v]);
        }
    }
}

// ---------------------------------------------------------------------------
// Local time step, RMS, update, forces
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_delta_t(int N, const double *vol, const double *res, double cfl,
                                                    double *dt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < N) dt[i] = vol[i] * cfl / res[kRES * (size_t)i + 10];
}

// Deterministic two-stage reductions: fixed per-thread strides, fixed tree.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out) {
    __shared__ double sh[NV][kBlock];
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
#pragma unroll
            for (int k = 0; k < NV; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] = sh[k][0];
    }
}

__global__ __launch_bounds__(kBlock) void k_rms_partial(int N, const double *res, double *part) {
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const double r = res[kRES * (size_t)i + k] + res[kRES * (size_t)i + 5 + k];
            v[k] += r * r;
        }
    }
    block_reduce_store<5>(v, part + 5 * blockIdx.x);
}

template <int NV>
__global__ __launch_bounds__(kBlock) void k_reduce_final(int nparts, const double *part, double *out) {
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += part[NV * p + k];
    block_reduce_store<NV>(v, out);
}

__global__ __launch_bounds__(kBlock) void k_update(int N, double relax, const int *order, const double *x, double *qa) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const size_t r = order[i];  // X is in the reference ordering
#pragma unroll
    for (int k = 0; k < 5; k++) qa[kQA * (size_t)i + k] += relax * x[5 * r + k];
}

__global__ __launch_bounds__(kBlock) void k_forces_partial(int N, int E, int B, const int *bnode, const int *btype,
                                                           const int *btag, int wall_tag, const double *sa,
                                                           const double *qa, double g, double *part) {
    double v[4] = {0, 0, 0, 0};
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < B; e += gridDim.x * kBlock) {
        const bool use = wall_tag > 0 ? (btag && btag[e] == wall_tag) : (btype[e] == 0);
        if (!use) continue;
        const double *r = qa + kQA * (size_t)bnode[e];
        Eos<double> s;
        eos_face(g, r, 0.0, 0.0, 0.0, s);
        const double *S = sa + 4 * (size_t)e;
        v[0] += s.p * S[0];
        v[1] += s.p * S[1];
        v[2] += s.p * S[2];
        v[3] += s.p * sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    }
    block_reduce_store<4>(v, part + 4 * blockIdx.x);
}

// Fills n ints; replaces hipMemsetAsync so the launch sequence can be
// captured in a graph.
__global__ void k_fill_int(int n, int v, int *p) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

template <class TD, class TS>
__global__ void k_convert(size_t n, const TS *s, TD *d) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) d[i] = (TD)s[i];
}

int grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    return (int)(g < 4096 ? (g ? g : 1) : 4096);
}

}  // namespace
}  // namespace CFD_RES_NS

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace CFD_RES_NS {

#if !CFD_FAST_MATH
int launch_lsq_coef(cfd_mesh *m) {
    const int nb = (m->N + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_lsq_coef, dim3(xcd_grid(nb)), dim3(kBlock), 0, m->stream, m->N, m->n2n_ia, m->n2n_ja, m->d.qa,
                       m->d.lsq);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

template <class T>
int launch_bc(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, int iter) {
    if (m->B == 0) return CFD_OK;
    BcArgs<T> a{m->N, m->E, m->B, iter, p->zpg_iter, m->bnode, m->btype, b.sa, b.qa, (T)p->gamma, (T)p->inf_rho,
                (T)p->inf_u, (T)p->inf_v, (T)p->inf_w, (T)p->inf_p, (T)p->outflow_p, m->flag, m->case_inf,
                m->bcase, m->ncase ? m->case_active : nullptr, m->ncase ? m->case_flag : nullptr};
    hipLaunchKernelGGL(k_bc<T>, dim3((m->B + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, a);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

#endif

template <class T>
int launch_gradient(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p) {
    const int nb = xcd_grid((m->N + kBlock - 1) / kBlock);
    if (p->gradient == CFD_GRAD_GG)
        hipLaunchKernelGGL(k_gg_grad<T>, dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->inc_ptr, m->sptr, m->nbr,
                           b.isx, b.isy, b.isz, b.vol, b.qa, b.qb);
    else
        hipLaunchKernelGGL(k_lsq_grad<T>, dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->order, m->n2n_ia, m->n2n_ja,
                           b.lsq, b.qa, b.qb);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

// Roe pressure correction after the limiter (Limiters.cpp:71-110).
template <class T>
static int launch_roe_correction(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, int nb, T g, T frac) {
    if (p->flux != CFD_FLUX_ROE) return CFD_OK;
    const int epoch = ++m->mark_epoch;
    hipLaunchKernelGGL(k_roe_candidates<T>, dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->order, m->n2n_ia,
                       m->n2n_ja, b.qa, b.qb, g, frac, m->roe_cand, m->roe_blk, epoch, m->risk);
    hipLaunchKernelGGL(k_roe_replay<T>, dim3(1), dim3(64), 0, m->stream, m->N, m->iperm, m->n2n_ia, m->n2n_ja, b.qa,
                       b.qb, g, frac, m->roe_cand, m->roe_blk, epoch, m->risk);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

template <class T>
int launch_limiter(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p) {
    const int nb = xcd_grid((m->N + kBlock - 1) / kBlock);
    const T frac = (p->limiter_order == 2) ? T(0.5) : T(0.0);
    const T k3 = (T)pow(p->venkat_k, 3.0);  // host pow, as the reference's glibc pow
    const T g = (T)p->gamma;
    if (p->limiter != CFD_LIM_VENKAT && p->limiter != CFD_LIM_BJ) return CFD_OK;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, m->stream, 1, 0, m->risk);
    if (p->limiter == CFD_LIM_VENKAT)
        hipLaunchKernelGGL((k_limiter<T, 1>), dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->inc_ptr, m->sptr,
                           m->nbr, b.qa, b.qb, b.vol, g, frac, k3, m->risk);
    else
        hipLaunchKernelGGL((k_limiter<T, 2>), dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->inc_ptr, m->sptr,
                           m->nbr, b.qa, b.qb, b.vol, g, frac, k3, m->risk);
    CFD_HIP(m, hipGetLastError());
    return launch_roe_correction<T>(m, b, p, nb, g, frac);
}

// Gradient then limiter; one fused kernel for the Green-Gauss gradient.
template <class T>
int launch_gradient_limiter(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p) {
    if (p->gradient != CFD_GRAD_GG || (p->limiter != CFD_LIM_VENKAT && p->limiter != CFD_LIM_BJ)) {
        int rc = launch_gradient<T>(m, b, p);
        return rc ? rc : launch_limiter<T>(m, b, p);
    }
    const int nb = xcd_grid((m->N + kBlock - 1) / kBlock);
    const T frac = (p->limiter_order == 2) ? T(0.5) : T(0.0);
    const T k3 = (T)pow(p->venkat_k, 3.0);
    const T g = (T)p->gamma;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, m->stream, 1, 0, m->risk);
    // neighbour records staged in LDS for FP32, direct gathers for FP64 (the
    // faster of the two for each precision on the 160^3 cube), unless the
    // autotuner chose for this mesh; CFD_STAGE_GL=0/1 (read at each call)
    // forces either
    const char *const env = getenv("CFD_STAGE_GL");
    const int tuned = m->tune_gl_stage[sizeof(T) == 4 ? 1 : 0];
    const bool stage = env && (env[0] == '0' || env[0] == '1') ? env[0] == '1'
                       : tuned >= 0                            ? tuned == 1
                                                               : sizeof(T) == 4;
    // with the fraction at 0 its terms drop out: 0 (q_j - q_c) is a zero, and
    // q_c + 0 + 1 g.R equals q_c + g.R, so the exact build keeps its bits
    // (for finite states; zeros may differ only in sign where every use
    // compares them with zero or squares them)
    const bool f0 = frac == T(0.0);
#define CFD_GL_LAUNCH(KIND, STAGE, F0)                                                                             \
    hipLaunchKernelGGL((k_gg_limiter<T, KIND, STAGE, F0>), dim3(nb), dim3(kBlock), 0, m->stream, m->N, m->inc_ptr, \
                       m->sptr, m->nbr, b.isx, b.isy, b.isz, b.vol, b.qa, b.qb, g, frac, k3, m->risk)
#define CFD_GL_STAGE(KIND, F0)          \
    do {                                \
        if (stage) CFD_GL_LAUNCH(KIND, true, F0); \
        else CFD_GL_LAUNCH(KIND, false, F0);      \
    } while (0)
    if (p->limiter == CFD_LIM_VENKAT) {
        if (f0) CFD_GL_STAGE(1, true);
        else CFD_GL_STAGE(1, false);
    } else {
        if (f0) CFD_GL_STAGE(2, true);
        else CFD_GL_STAGE(2, false);
    }
#undef CFD_GL_STAGE
#undef CFD_GL_LAUNCH
    CFD_HIP(m, hipGetLastError());
    return launch_roe_correction<T>(m, b, p, nb, g, frac);
}

template <class T, int FLUX>
static void launch_res_flux(cfd_mesh *m, const ResArgs<T> &a, int order, bool lim, bool stats, int nb) {
    constexpr bool kStage = FLUX != 0;  // staged gathers for HLLC and AUSM+
    dim3 g(nb), bl(kBlock);
    if (order == 2) {
        if (lim) {
            if (stats) hipLaunchKernelGGL((k_residual<T, FLUX, 2, true, true, kStage>), g, bl, 0, m->stream, a);
            else hipLaunchKernelGGL((k_residual<T, FLUX, 2, true, false, kStage>), g, bl, 0, m->stream, a);
        } else {
            if (stats) hipLaunchKernelGGL((k_residual<T, FLUX, 2, false, true, kStage>), g, bl, 0, m->stream, a);
            else hipLaunchKernelGGL((k_residual<T, FLUX, 2, false, false, kStage>), g, bl, 0, m->stream, a);
        }
    } else {
        if (stats) hipLaunchKernelGGL((k_residual<T, FLUX, 1, false, true, kStage>), g, bl, 0, m->stream, a);
        else hipLaunchKernelGGL((k_residual<T, FLUX, 1, false, false, kStage>), g, bl, 0, m->stream, a);
    }
}

template <class T, int FLUX>
static void launch_patch_flux(cfd_mesh *m, const PatchArgs<T> &a, int order, bool lim, bool stats) {
    dim3 g(xcd_grid(a.npatch)), bl(kBlock);
    if (order == 2) {
        if (lim) {
            if (stats) hipLaunchKernelGGL((k_residual_patch<T, FLUX, 2, true, true>), g, bl, 0, m->stream, a);
            else hipLaunchKernelGGL((k_residual_patch<T, FLUX, 2, true, false>), g, bl, 0, m->stream, a);
        } else {
            if (stats) hipLaunchKernelGGL((k_residual_patch<T, FLUX, 2, false, true>), g, bl, 0, m->stream, a);
            else hipLaunchKernelGGL((k_residual_patch<T, FLUX, 2, false, false>), g, bl, 0, m->stream, a);
        }
    } else {
        if (stats) hipLaunchKernelGGL((k_residual_patch<T, FLUX, 1, false, true>), g, bl, 0, m->stream, a);
        else hipLaunchKernelGGL((k_residual_patch<T, FLUX, 1, false, false>), g, bl, 0, m->stream, a);
    }
}

// The fast build's flux-once residual over work-edge patches (built on the
// first call).
template <class T>
static int launch_patch_residual(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, bool stats) {
    // the fast build evaluates its work edges in node order; the exact build
    // in the reference's edge order, so each node's sum keeps its order
#if CFD_FAST_MATH
    const int order = kPatchNodeOrder;
#else
    const int order = kPatchEdgeOrder;
#endif
    int rc = cfd_ensure_patches(m, sizeof(T) == 4, order);
    if (rc) return rc;
    const PatchSet &ps = m->patch[order];
    const T *ws;
    if constexpr (sizeof(T) == 4) ws = ps.ws_f;
    else ws = ps.ws;
    PatchArgs<T> a{m->N, ps.npatch, ps.pn_start, m->inc_ptr, m->sptr, ps.pn_slot, ps.pw_ptr, ps.pw_l, ps.pw_r, ws, b.qa, b.qb,
                   b.res, (T)p->gamma, (p->limiter_order == 2) ? T(0.5) : T(0.0), p->entropy_fix, m->eigbits,
                   m->flag, m->ncase ? m->case_flag : nullptr, m->ecase};
    const bool lim = p->limiter != CFD_LIM_NONE;
    const bool st = stats && p->flux == CFD_FLUX_ROE;
    if (p->flux == CFD_FLUX_ROE) launch_patch_flux<T, 0>(m, a, p->order, lim, st);
    else if (p->flux == CFD_FLUX_HLLC) launch_patch_flux<T, 1>(m, a, p->order, lim, false);
    else launch_patch_flux<T, 2>(m, a, p->order, lim, false);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

template <class T>
static int launch_node_residual(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, bool stats) {
    ResArgs<T> a{m->N, m->NT, m->inc_ptr, m->sptr, m->nbr, b.isx, b.isy, b.isz, b.qa, b.qb, b.res, (T)p->gamma,
                 (p->limiter_order == 2) ? T(0.5) : T(0.0), p->entropy_fix, m->eigbits, m->flag,
                 m->ncase ? m->case_flag : nullptr, m->ecase};
    const int nb = xcd_grid((m->N + kBlock - 1) / kBlock);
    const bool lim = p->limiter != CFD_LIM_NONE;
    const bool st = stats && p->flux == CFD_FLUX_ROE;
    if (p->flux == CFD_FLUX_ROE) launch_res_flux<T, 0>(m, a, p->order, lim, st, nb);
    else if (p->flux == CFD_FLUX_HLLC) launch_res_flux<T, 1>(m, a, p->order, lim, false, nb);
    else launch_res_flux<T, 2>(m, a, p->order, lim, false, nb);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

// The fast build runs the flux-once patch kernel for second order, where each
// edge's reconstruction and flux are worth computing once, and the
// node-centred kernel for first order, where the patch kernel's extra traffic
// (work-edge lists, LDS partial sums) costs more than the second flux
// evaluation.  The exact build runs the node-centred kernel unless asked
// (cfd_residual_uses_patches; CFD_FLUX_ONCE=0/1 forces either); both kernels
// give the same bits there.
template <class T>
int launch_residual(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, bool stats) {
    if (cfd_residual_uses_patches(p)) return launch_patch_residual<T>(m, b, p, stats);
    return launch_node_residual<T>(m, b, p, stats);
}

#if !CFD_FAST_MATH
int launch_delta_t(cfd_mesh *m, double cfl) {
    hipLaunchKernelGGL(k_delta_t, dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, m->N, m->d.vol,
                       m->d.res, cfl, m->dt);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

int launch_rms(cfd_mesh *m, double *sums5) {
    const int nparts = grid_for(m->N) > 1024 ? 1024 : grid_for(m->N);
    hipLaunchKernelGGL(k_rms_partial, dim3(nparts), dim3(kBlock), 0, m->stream, m->N, m->d.res, m->red);
    hipLaunchKernelGGL(k_reduce_final<5>, dim3(1), dim3(kBlock), 0, m->stream, nparts, m->red, m->redout);
    CFD_HIP(m, hipGetLastError());
    CFD_HIP(m, hipMemcpyAsync(m->hred, m->redout, 5 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    for (int k = 0; k < 5; k++) sums5[k] = m->hred[k];
    return CFD_OK;
}

int launch_update(cfd_mesh *m, double relax) {
    hipLaunchKernelGGL(k_update, dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, m->N, relax, m->order,
                       m->x, m->d.qa);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

int launch_forces(cfd_mesh *m, const cfd_params *p, int wall_tag, double *out4) {
    const int nparts = m->B ? (grid_for(m->B) > 256 ? 256 : grid_for(m->B)) : 1;
    hipLaunchKernelGGL(k_forces_partial, dim3(nparts), dim3(kBlock), 0, m->stream, m->N, m->E, m->B, m->bnode,
                       m->btype, m->btag, wall_tag, m->d.sa, m->d.qa, p->gamma, m->red);
    hipLaunchKernelGGL(k_reduce_final<4>, dim3(1), dim3(kBlock), 0, m->stream, nparts, m->red, m->redout);
    CFD_HIP(m, hipGetLastError());
    CFD_HIP(m, hipMemcpyAsync(m->hred, m->redout, 4 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    for (int k = 0; k < 4; k++) out4[k] = m->hred[k];
    return CFD_OK;
}

int launch_convert_state_to_fp32(cfd_mesh *m) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for((size_t)kQA * m->NT)), dim3(kBlock), 0, m->stream,
                       (size_t)kQA * m->NT, m->d.qa, m->f.qa);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for((size_t)kQB * m->N)), dim3(kBlock), 0, m->stream,
                       (size_t)kQB * m->N, m->d.qb, m->f.qb);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

int launch_residual_to_fp64(cfd_mesh *m) {
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for((size_t)kRES * m->N)), dim3(kBlock), 0, m->stream,
                       (size_t)kRES * m->N, m->f.res, m->d.res);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

int launch_convert_geometry_fp32(cfd_mesh *m) {
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for((size_t)4 * (m->E + m->B))), dim3(kBlock), 0,
                       m->stream, (size_t)4 * m->B, m->d.sa, m->f.sa);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for((size_t)m->N)), dim3(kBlock), 0, m->stream,
                       (size_t)m->N, m->d.vol, m->f.vol);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(m->nsell)), dim3(kBlock), 0, m->stream, m->nsell,
                       m->d.isx, m->f.isx);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(m->nsell)), dim3(kBlock), 0, m->stream, m->nsell,
                       m->d.isy, m->f.isy);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(m->nsell)), dim3(kBlock), 0, m->stream, m->nsell,
                       m->d.isz, m->f.isz);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for((size_t)8 * m->N)), dim3(kBlock), 0, m->stream,
                       (size_t)8 * m->N, m->d.lsq, m->f.lsq);
    CFD_HIP(m, hipGetLastError());
    return CFD_OK;
}

#endif

#if !CFD_FAST_MATH
template int launch_bc<double>(cfd_mesh *, ResidualBufs<double> &, const cfd_params *, int);
template int launch_bc<float>(cfd_mesh *, ResidualBufs<float> &, const cfd_params *, int);
#endif
template int launch_gradient<double>(cfd_mesh *, ResidualBufs<double> &, const cfd_params *);
template int launch_gradient<float>(cfd_mesh *, ResidualBufs<float> &, const cfd_params *);
template int launch_limiter<double>(cfd_mesh *, ResidualBufs<double> &, const cfd_params *);
template int launch_limiter<float>(cfd_mesh *, ResidualBufs<float> &, const cfd_params *);
template int launch_gradient_limiter<double>(cfd_mesh *, ResidualBufs<double> &, const cfd_params *);
template int launch_gradient_limiter<float>(cfd_mesh *, ResidualBufs<float> &, const cfd_params *);
template int launch_residual<double>(cfd_mesh *, ResidualBufs<double> &, const cfd_params *, bool);
template int launch_residual<float>(cfd_mesh *, ResidualBufs<float> &, const cfd_params *, bool);

}  // namespace CFD_RES_NS
