// Device-side thermodynamics, numerical fluxes and reconstruction.
//
// Conservative variables, ideal gas, GENERIC non-dimensionalisation: the
// path of Euler3D/src/Material.cpp:216-240, Euler3D/src/FluxTest.cpp:3026-3200
// (Roe), Euler3D/src/HLLC_Fluxes.cpp:143-326 (HLLC).  Expressions keep the
// reference's operation order; the library is compiled with
// -ffp-contract=off, so FP64 results are bit-identical to the reference built
// without FMA contraction.  Templated on the arithmetic type for the FP32
// variant of the residual path.
#pragma once

#include <hip/hip_runtime.h>

// CFD_FAST_MATH 1 (residual_fast.hip, compiled with FMA contraction): each
// shared denominator is inverted once and multiplied, and the 5x5 products
// skip the transformation matrix's structural zeros.  The functions of each
// build live in their own inline namespace, so both builds link together.
#ifndef CFD_FAST_MATH
#define CFD_FAST_MATH 0
#endif

namespace cfd {
#if CFD_FAST_MATH
inline namespace fast_math {
#else
inline namespace exact_math {
#endif

template <class T>
struct Eos {
    T rho, u, v, w, et, q2, p, ht, c, ubar;
};

template <class T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }  // MIN, UTILS/include/Trim_Utils.h:95
template <class T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }  // MAX, UTILS/include/Trim_Utils.h:102

// Compute_EOS_Variables_Face, Euler3D/src/Material.cpp:226-238
template <class T>
__device__ __forceinline__ void eos_face(T g, const T *Q, T nx, T ny, T nz, Eos<T> &e) {
    e.rho = Q[0];
#if CFD_FAST_MATH
    const T ir = T(1.0) / e.rho;
    e.u = Q[1] * ir;
    e.v = Q[2] * ir;
    e.w = Q[3] * ir;
    e.et = Q[4] * ir;
    e.q2 = e.u * e.u + e.v * e.v + e.w * e.w;
    e.p = (g - T(1.0)) * (Q[4] - T(0.5) * e.rho * e.q2);
    e.ht = (Q[4] + e.p) * ir;
    e.c = sqrt(g * e.p * ir);
#else
    e.u = Q[1] / e.rho;
    e.v = Q[2] / e.rho;
    e.w = Q[3] / e.rho;
    e.et = Q[4] / e.rho;
    e.q2 = e.u * e.u + e.v * e.v + e.w * e.w;
    e.p = (g - T(1.0)) * (e.rho * e.et - T(0.5) * e.rho * e.q2);
    e.ht = (e.rho * e.et + e.p) / e.rho;
    e.c = sqrt((g * e.p) / e.rho);
#endif
    e.ubar = e.u * nx + e.v * ny + e.w * nz;
}

// Vector3D::magnitude and Vector3D::normalize, MATH/src/Vector3D.cpp:216-229
template <class T>
__device__ __forceinline__ T area_normal(T sx, T sy, T sz, T &nx, T &ny, T &nz) {
    T mag = sqrt(sx * sx + sy * sy + sz * sz);
    nx = sx;
    ny = sy;
    nz = sz;
    if (mag > T(1.0e-20)) {
#if CFD_FAST_MATH
        const T im = T(1.0) / mag;
        nx = sx * im;
        ny = sy * im;
        nz = sz * im;
#else
        nx = sx / mag;
        ny = sy / mag;
        nz = sz / mag;
#endif
    }
    return mag;
}

// Get_SpeedSound(u,v,w,ht), Euler3D/src/Material.cpp:939-975
template <class T>
__device__ __forceinline__ T roe_speed_sound(T g, T u, T v, T w, T ht) {
    T phi = T(0.5) * (g - T(1.0)) * (u * u + v * v + w * w);
    return sqrt((g - T(1.0)) * ht - phi);
}

// Roe_EntropyFix, Euler3D/src/Roe_EntropyFix.cpp:19-59: E0 = |ubar|,
// E1 = |ubar+c|, E2 = |ubar-c| with Kermani's smoothing when fix == 1.
template <class T>
__device__ __forceinline__ void roe_eigen(int fix, T ubL, T cL, T ubR, T cR, T ub, T c, T &E0, T &E1, T &E2) {
    E0 = fabs(ub);
    if (fix == 1) {
        T lL = ubL + cL, lR = ubR + cR, l = ub + c;
        T eps = T(4.0) * tmax(T(0.0), tmax((l - lL), (lR - l)));
        E1 = (fabs(l) < eps) ? (l * l + eps * eps) / (T(2.0) * eps) : fabs(l);
        lL = ubL - cL;
        lR = ubR - cR;
        l = ub - c;
        eps = T(4.0) * tmax(T(0.0), tmax((l - lL), (lR - l)));
        E2 = (fabs(l) < eps) ? (l * l + eps * eps) / (T(2.0) * eps) : fabs(l);
    } else {
        E1 = fabs(ub + c);
        E2 = fabs(ub - c);
    }
}

// Compute_Flux_Roe_Optimized, Euler3D/src/FluxTest.cpp:3026-3200.
// Returns area * max eigenvalue; E receives the three eigenvalue magnitudes.
template <class T>
__device__ __forceinline__ T roe_flux(T g, int efix, T sx, T sy, T sz, const T *QL, const T *QR, T *conv, T *diss,
                                      T &E0, T &E1, T &E2) {
    T nx, ny, nz;
    const T area = area_normal(sx, sy, sz, nx, ny, nz);
    Eos<T> L, R;
    eos_face(g, QL, nx, ny, nz, L);
    eos_face(g, QR, nx, ny, nz, R);
    T fL0 = L.rho * L.ubar, fR0 = R.rho * R.ubar;
    T fL[5] = {fL0, L.u * fL0 + L.p * nx, L.v * fL0 + L.p * ny, L.w * fL0 + L.p * nz, L.ht * fL0};
    T fR[5] = {fR0, R.u * fR0 + R.p * nx, R.v * fR0 + R.p * ny, R.w * fR0 + R.p * nz, R.ht * fR0};
    const T rho = sqrt(R.rho * L.rho);
    const T sigma = rho / (L.rho + rho);
    const T u = L.u + sigma * (R.u - L.u);
    const T v = L.v + sigma * (R.v - L.v);
    const T w = L.w + sigma * (R.w - L.w);
    const T ht = L.ht + sigma * (R.ht - L.ht);
    const T c = roe_speed_sound(g, u, v, w, ht);
    const T ubar = u * nx + v * ny + w * nz;
    roe_eigen(efix, L.ubar, L.c, R.ubar, R.c, ubar, c, E0, E1, E2);
    const T maxlambda = tmax(E0, tmax(E1, E2));
    const T phi = T(0.5) * (g - T(1.0)) * (u * u + v * v + w * w);
    const T dQ0 = R.rho - L.rho, dQ1 = R.u - L.u, dQ2 = R.v - L.v, dQ3 = R.w - L.w, dQ4 = R.p - L.p;
    const T dU = dQ1 * nx + dQ2 * ny + dQ3 * nz;
#if CFD_FAST_MATH
    {
        // the same dissipation with 1/c, 1/(rho c) and 1/(g - 1) formed once,
        // and M dw over M's nonzero entries only
        const T ic = T(1.0) / c, irc = T(1.0) / (rho * c), igm1 = T(1.0) / (g - T(1.0));
        const T e12 = (E1 + E2) - T(2.0) * E0, d12 = E1 - E2;
        const T dw0 = E0 * dQ0 + T(0.5) * e12 * dQ4 * ic * ic + T(0.5) * rho * d12 * dU * ic;
        const T a = T(0.5) * dU * e12, b = T(0.5) * d12 * dQ4 * irc;
        const T dw1 = E0 * dQ1 + nx * a + nx * b;
        const T dw2 = E0 * dQ2 + ny * a + ny * b;
        const T dw3 = E0 * dQ3 + nz * a + nz * b;
        const T dw4 = T(0.5) * dQ4 * (E1 + E2) + T(0.5) * rho * c * dU * d12;
        const T fa[5] = {dw0, u * dw0 + rho * dw1, v * dw0 + rho * dw2, w * dw0 + rho * dw3,
                         phi * igm1 * dw0 + rho * (u * dw1 + v * dw2 + w * dw3) + igm1 * dw4};
#pragma unroll
        for (int i = 0; i < 5; i++) {
            conv[i] = T(0.5) * (fL[i] + fR[i]) * area;
            diss[i] = T(-0.5) * fa[i] * area;
        }
        return area * maxlambda;
    }
#endif
    // M = Compute_Transformation_Conservative_To_RUP, Euler3D/src/Material.cpp:1172-1235
    const T M[5][5] = {{T(1.0), T(0.0), T(0.0), T(0.0), T(0.0)},
                       {u, rho, T(0.0), T(0.0), T(0.0)},
                       {v, T(0.0), rho, T(0.0), T(0.0)},
                       {w, T(0.0), T(0.0), rho, T(0.0)},
                       {phi / (g - T(1.0)), rho * u, rho * v, rho * w, T(1.0) / (g - T(1.0))}};
    T dw[5];
    dw[0] = E0 * dQ0 + (E1 + E2 - T(2.0) * E0) * dQ4 / (T(2.0) * c * c) + rho * (E1 - E2) * dU / (T(2.0) * c);
    dw[1] = E0 * dQ1 + T(0.5) * nx * dU * ((E1 + E2) - T(2.0) * E0) + nx * (E1 - E2) * dQ4 / (T(2.0) * rho * c);
    dw[2] = E0 * dQ2 + T(0.5) * ny * dU * ((E1 + E2) - T(2.0) * E0) + ny * (E1 - E2) * dQ4 / (T(2.0) * rho * c);
    dw[3] = E0 * dQ3 + T(0.5) * nz * dU * ((E1 + E2) - T(2.0) * E0) + nz * (E1 - E2) * dQ4 / (T(2.0) * rho * c);
    dw[4] = T(0.5) * dQ4 * (E1 + E2) + T(0.5) * rho * c * dU * (E1 - E2);
#pragma unroll
    for (int i = 0; i < 5; i++) {
        T fa = T(0.0);  // MC_Matrix_Mul_Vector, MATH/src/MC_Iterative_Jacobi.c:174-187
#pragma unroll
        for (int j = 0; j < 5; j++) fa += M[i][j] * dw[j];
        conv[i] = T(0.5) * (fL[i] + fR[i]) * area;
        diss[i] = T(-0.5) * fa * area;
    }
    return area * maxlambda;
}

// Compute_HLLCFlux, Euler3D/src/HLLC_Fluxes.cpp:143-326.  Sets bad = 1 when no
// wave configuration applies (reference: error()).
// Freestream and outflow pressure seen by one boundary edge.
template <class T>
struct BcFree {
    T rho, u, v, w, p, out_p;
};

// Freestream of case c of a union of cases: (rho, u, v, w, p, outflow p) at
// case_inf[8c .. 8c+5].
template <class T>
__device__ __forceinline__ BcFree<T> bc_free_of_case(const double *case_inf, int c) {
    const double *f = case_inf + 8 * (size_t)c;
    return BcFree<T>{(T)f[0], (T)f[1], (T)f[2], (T)f[3], (T)f[4], (T)f[5]};
}

// Characteristic ghost state of one boundary edge:
// Compute_Characteristic_BoundaryCondition_New, Euler3D/src/BC.cpp:381-613
// (conservative variables, gauge pressure 0; the older per-edge
// Compute_Characteristic_BoundaryCondition, Euler3D/src/BC.cpp:33-380, gives
// the same state).  QL: the node, QR: the current ghost, s: outward area
// vector, type: 0 wall, 1 far field, 2 inflow, 3 outflow, 4-7 subsonic /
// supersonic inflow / outflow.  Returns 0, or 1 if the ghost pressure is
// negative (QB still written), or 2 if no boundary instance applies (QB not
// written).
template <class T>
__device__ __forceinline__ int bc_ghost(T g, const T *QL, const T *QR, T sx, T sy, T sz, int type, int iter, int zpg,
                                        const BcFree<T> &f, T *QB) {
    T nx, ny, nz, Qa[5];
    area_normal(sx, sy, sz, nx, ny, nz);
    Eos<T> I, Z;
    eos_face(g, QL, nx, ny, nz, I);
#pragma unroll
    for (int k = 0; k < 5; k++) Qa[k] = T(0.5) * (QL[k] + QR[k]);
    eos_face(g, Qa, nx, ny, nz, Z);
    const T l1 = Z.ubar, l4 = l1 + Z.c, l5 = l1 - Z.c;
    const T infP = f.p;
    T pb = infP, rb = f.rho, ub = f.u, vb = f.v, wb = f.w;
    int inst = type;
    bool bad = false;
    switch (type) {
    case 0: break;  // Euler wall
    case 1:         // far field
        if ((l1 <= T(0.0)) && (l4 <= T(0.0)) && (l5 <= T(0.0))) inst = 10;
        else if ((l1 >= T(0.0)) && (l4 >= T(0.0)) && (l5 >= T(0.0))) inst = 11;
        else if ((l1 >= T(0.0)) && (l5 <= T(0.0)) && (l4 >= T(0.0))) inst = 9;
        else if ((l1 <= T(0.0)) && (l4 >= T(0.0)) && (l5 <= T(0.0))) inst = 8;
        else bad = true;
        break;
    case 2:  // inflow
        if ((l1 <= T(0.0)) && (l4 >= T(0.0)) && (l5 <= T(0.0))) inst = 4;
        else if ((l1 <= T(0.0)) && (l4 <= T(0.0)) && (l5 <= T(0.0))) inst = 6;
        else inst = (fabs(l1) < Z.c) ? 4 : 6;
        break;
    case 3:  // outflow
        if ((l1 >= T(0.0)) && (l5 <= T(0.0)) && (l4 >= T(0.0))) inst = 5;
        else if ((l1 >= T(0.0)) && (l4 >= T(0.0)) && (l5 >= T(0.0))) inst = 7;
        else inst = (fabs(l1) < Z.c) ? 5 : 7;
        break;
    case 4: case 5: case 6: case 7: break;
    default: bad = true;
    }
    switch (inst) {
    case 0:
        pb = I.p + Z.rho * Z.c * I.ubar;
        if (iter < zpg && pb < T(0.0)) pb = I.p;
        rb = I.rho + (pb - I.p) / (Z.c * Z.c);
        ub = I.u - nx * (pb - I.p) / (Z.rho * Z.c);
        vb = I.v - ny * (pb - I.p) / (Z.rho * Z.c);
        wb = I.w - nz * (pb - I.p) / (Z.rho * Z.c);
        break;
    case 5:
    case 9:
        pb = (inst == 5) ? f.out_p : infP;
        rb = I.rho + (pb - I.p) / (Z.c * Z.c);
        ub = I.u - nx * (pb - I.p) / (Z.rho * Z.c);
        vb = I.v - ny * (pb - I.p) / (Z.rho * Z.c);
        wb = I.w - nz * (pb - I.p) / (Z.rho * Z.c);
        break;
    case 7:
    case 11:
        pb = I.p;
        rb = I.rho;
        ub = I.u;
        vb = I.v;
        wb = I.w;
        break;
    case 8: {
        const T temp = nx * (f.u - I.u) + ny * (f.v - I.v) + nz * (f.w - I.w);
        pb = T(0.5) * (infP + I.p - Z.rho * Z.c * temp);
        rb = f.rho + (pb - infP) / (Z.c * Z.c);
        ub = f.u + nx * (pb - infP) / (Z.rho * Z.c);
        vb = f.v + ny * (pb - infP) / (Z.rho * Z.c);
        wb = f.w + nz * (pb - infP) / (Z.rho * Z.c);
        break;
    }
    default: break;  // 4, 6, 10: freestream
    }
    if (bad) return 2;
    // Get_TotalEnergy, Euler3D/src/Material.cpp:1012-1040
    T et = ub * ub + vb * vb + wb * wb;
    et = pb / ((g - T(1.0)) * rb) + T(0.5) * et;
    QB[0] = rb;
    QB[1] = rb * ub;
    QB[2] = rb * vb;
    QB[3] = rb * wb;
    QB[4] = rb * et;
    return pb < T(0.0) ? 1 : 0;
}

template <class T>
__device__ __forceinline__ T hllc_flux(T g, T sx, T sy, T sz, const T *QL, const T *QR, T *F, int &bad) {
    T nx, ny, nz;
    const T area = area_normal(sx, sy, sz, nx, ny, nz);
    Eos<T> L, R;
    eos_face(g, QL, nx, ny, nz, L);
    eos_face(g, QR, nx, ny, nz, R);
    const T rho = sqrt(R.rho * L.rho);
    const T sigma = rho / (L.rho + rho);
    const T u = L.u + sigma * (R.u - L.u);
    const T v = L.v + sigma * (R.v - L.v);
    const T w = L.w + sigma * (R.w - L.w);
    const T ht = L.ht + sigma * (R.ht - L.ht);
    const T phi = T(0.5) * (g - T(1.0)) * (u * u + v * v + w * w);
    T c = (g - T(1.0)) * ht - phi;
    c = sqrt(c);
    const T ubar = u * nx + v * ny + w * nz;
    const T maxlambda = tmax(fabs(ubar), tmax(fabs(ubar + c), fabs(ubar - c)));
    const T SL = tmin(L.ubar - L.c, ubar - c);
    const T SR = tmax(R.ubar + R.c, ubar + c);
    const T SM = (R.rho * R.ubar * (SR - R.ubar) - L.rho * L.ubar * (SL - L.ubar) + L.p - R.p) /
                 (R.rho * (SR - R.ubar) - L.rho * (SL - L.ubar));
    if (SL > T(0.0)) {
        F[0] = L.rho * L.ubar;
        F[1] = L.u * F[0] + L.p * nx;
        F[2] = L.v * F[0] + L.p * ny;
        F[3] = L.w * F[0] + L.p * nz;
        F[4] = L.ht * F[0];
    } else if (SR < T(0.0)) {
        F[0] = R.rho * R.ubar;
        F[1] = R.u * F[0] + R.p * nx;
        F[2] = R.v * F[0] + R.p * ny;
        F[3] = R.w * F[0] + R.p * nz;
        F[4] = R.ht * F[0];
    } else {
        // star region of the side facing the contact
        const bool left = (SL <= T(0.0)) && (SM > T(0.0));
        const bool right = (SM <= T(0.0)) && (SR > T(0.0));
        if (!left && !right) {
            bad = 1;
            for (int i = 0; i < 5; i++) F[i] = T(NAN);
        } else {
            const Eos<T> &K = left ? L : R;
            const T S = left ? SL : SR;
            const T pstar = K.p + K.rho * (K.ubar - S) * (K.ubar - SM);
            const T omega = T(1.0) / (S - SM);
            const T cnst1 = S - K.ubar;
            const T cnst2 = pstar - K.p;
            const T rs = omega * (cnst1 * K.rho);
            const T rus = omega * (cnst1 * K.rho * K.u + cnst2 * nx);
            const T rvs = omega * (cnst1 * K.rho * K.v + cnst2 * ny);
            const T rws = omega * (cnst1 * K.rho * K.w + cnst2 * nz);
            const T res = omega * (cnst1 * K.rho * K.et - K.p * K.ubar + pstar * SM);
            F[0] = rs * SM;
            F[1] = rus * SM + pstar * nx;
            F[2] = rvs * SM + pstar * ny;
            F[3] = rws * SM + pstar * nz;
            F[4] = (res + pstar) * SM;
        }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) F[i] *= area;
    return area * maxlambda;
}

// AUSM+ of M.-S. Liou, J. Comput. Phys. 129:364-382 (1996): split Mach
// numbers (beta = 1/8), split pressures (alpha = 3/16), interface sound speed
// from the critical speed c*^2 = 2(g-1)/(g+1) H.  The reference's AUSM files
// are empty stubs (Euler3D/src/AUSM_Fluxes.cpp:24-54).  Time-step eigenvalue
// as the HLLC routine (Roe average).
template <class T>
__device__ __forceinline__ T ausmp_mach_p(T M) {
    return fabs(M) >= T(1.0) ? T(0.5) * (M + fabs(M))
                             : T(0.25) * (M + T(1.0)) * (M + T(1.0)) + T(0.125) * (M * M - T(1.0)) * (M * M - T(1.0));
}
template <class T>
__device__ __forceinline__ T ausmp_mach_m(T M) {
    return fabs(M) >= T(1.0) ? T(0.5) * (M - fabs(M))
                             : T(-0.25) * (M - T(1.0)) * (M - T(1.0)) - T(0.125) * (M * M - T(1.0)) * (M * M - T(1.0));
}
template <class T>
__device__ __forceinline__ T ausmp_pres_p(T M) {
    return fabs(M) >= T(1.0) ? (M > T(0.0) ? T(1.0) : T(0.0))
                             : T(0.25) * (M + T(1.0)) * (M + T(1.0)) * (T(2.0) - M) +
                                   T(0.1875) * M * (M * M - T(1.0)) * (M * M - T(1.0));
}
template <class T>
__device__ __forceinline__ T ausmp_pres_m(T M) {
    return fabs(M) >= T(1.0) ? (M < T(0.0) ? T(1.0) : T(0.0))
                             : T(0.25) * (M - T(1.0)) * (M - T(1.0)) * (T(2.0) + M) -
                                   T(0.1875) * M * (M * M - T(1.0)) * (M * M - T(1.0));
}
template <class T>
__device__ __forceinline__ T ausmp_flux(T g, T sx, T sy, T sz, const T *QL, const T *QR, T *F) {
    T nx, ny, nz;
    const T area = area_normal(sx, sy, sz, nx, ny, nz);
    Eos<T> L, R;
    eos_face(g, QL, nx, ny, nz, L);
    eos_face(g, QR, nx, ny, nz, R);
    const T cs2L = T(2.0) * (g - T(1.0)) / (g + T(1.0)) * L.ht;
    const T cs2R = T(2.0) * (g - T(1.0)) / (g + T(1.0)) * R.ht;
    const T csL = sqrt(cs2L), csR = sqrt(cs2R);
    const T chL = cs2L / tmax(csL, L.ubar);
    const T chR = cs2R / tmax(csR, -R.ubar);
    const T c12 = tmin(chL, chR);
#if CFD_FAST_MATH
    const T ic12 = T(1.0) / c12;
    const T ML = L.ubar * ic12, MR = R.ubar * ic12;
#else
    const T ML = L.ubar / c12, MR = R.ubar / c12;
#endif
    const T m12 = ausmp_mach_p(ML) + ausmp_mach_m(MR);
    const T p12 = ausmp_pres_p(ML) * L.p + ausmp_pres_m(MR) * R.p;
    const T mp = T(0.5) * (m12 + fabs(m12)), mm = T(0.5) * (m12 - fabs(m12));
    const T mL = c12 * mp * L.rho, mR = c12 * mm * R.rho;
    F[0] = mL + mR;
    F[1] = mL * L.u + mR * R.u + p12 * nx;
    F[2] = mL * L.v + mR * R.v + p12 * ny;
    F[3] = mL * L.w + mR * R.w + p12 * nz;
    F[4] = mL * L.ht + mR * R.ht;
#pragma unroll
    for (int i = 0; i < 5; i++) F[i] *= area;
    const T rho = sqrt(R.rho * L.rho);
    const T sigma = rho / (L.rho + rho);
    const T u = L.u + sigma * (R.u - L.u);
    const T v = L.v + sigma * (R.v - L.v);
    const T w = L.w + sigma * (R.w - L.w);
    const T ht = L.ht + sigma * (R.ht - L.ht);
    const T c = roe_speed_sound(g, u, v, w, ht);
    const T ubar = u * nx + v * ny + w * nz;
    return area * tmax(fabs(ubar), tmax(fabs(ubar + c), fabs(ubar - c)));
}

}  // inline namespace
}  // namespace cfd
