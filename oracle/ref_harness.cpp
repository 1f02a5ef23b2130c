// Test-infrastructure driver for the compiled reference (Euler3D) hot path.
//
// Linked against objects compiled from the unmodified sources under
// Trunk/Euler3D/src, Trunk/MATH/src and Trunk/UTILS/src (see oracle/Makefile);
// used only to write golden vectors that pin the oracle.  It drives the
// reference exactly as Euler3D/main.cpp:94-158 and
// Euler3D/src/Solver_Steady_Implicit.cpp:45-140 do, setting the parameter
// globals directly instead of parsing a deck, and dumps every array of one
// pseudo-time iteration as .npy files:
//   mesh metrics, state after the boundary condition, LSQ gradients,
//   limiter, residual (convective and dissipative), time-step accumulator,
//   local time step, block-CRS Jacobian and RHS, block GS and LU-SGS solutions.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Commons.h"
#include "Gradient.h"
#include "MC.h"
#include "Material.h"
#include "MeshIO.h"
#include "RestartIO.h"
#include "Roe_Fluxes.h"
#include "Solver.h"
#include "SolverParameters.h"
#include "Trim_Utils.h"

static std::string g_out;

static void npy_write(const char *name, const char *descr, const void *data, size_t n, size_t elem,
                      size_t cols = 0) {
    std::string path = g_out + "/" + name + ".npy";
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    char shape[64];
    if (cols)
        snprintf(shape, sizeof(shape), "(%zu, %zu)", n / cols, cols);
    else
        snprintf(shape, sizeof(shape), "(%zu,)", n);
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr += std::string(pad, ' ') + "\n";
    unsigned short hl = (unsigned short)hdr.size();
    fwrite("\x93NUMPY\x01\x00", 1, 8, fp);
    fwrite(&hl, 2, 1, fp);
    fwrite(hdr.data(), 1, hdr.size(), fp);
    if (n) fwrite(data, elem, n, fp);
    fclose(fp);
}
static void dump_d(const char *name, const std::vector<double> &v, size_t cols = 0) {
    npy_write(name, "<f8", v.data(), v.size(), 8, cols);
}
static void dump_i(const char *name, const std::vector<int> &v, size_t cols = 0) {
    npy_write(name, "<i4", v.data(), v.size(), 4, cols);
}

static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
static inline double rsym(uint64_t seed, uint64_t idx, uint64_t k) {
    uint64_t r = splitmix64(seed ^ splitmix64(idx * 16ULL + k));
    return 2.0 * ((double)(r >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
}

static std::vector<double> gather5(double *a, double *b, double *c, double *d, double *e, int n) {
    std::vector<double> v(5 * (size_t)n);
    double *src[5] = {a, b, c, d, e};
    for (int k = 0; k < 5; k++)
        for (int i = 0; i < n; i++) v[(size_t)k * n + i] = src[k][i];
    return v;
}

int main(int argc, char **argv) {
    if (argc < 15) {
        fprintf(stderr,
                "usage: %s mesh.ugrid outdir flux(roe|hllc) order(1|2) limiter(none|venkat|bj) "
                "mach alpha_deg seed amp gs_sweeps iter entropy_fix venkat_k limiter_order(1|2)\n",
                argv[0]);
        return 1;
    }
    g_out = argv[2];
    const std::string flux = argv[3];
    const int order = atoi(argv[4]);
    const std::string lim = argv[5];
    const double mach = atof(argv[6]);
    const double alpha = atof(argv[7]);
    const uint64_t seed = strtoull(argv[8], nullptr, 10);
    const double amp = atof(argv[9]);
    const int sweeps = atoi(argv[10]);
    const int iter = atoi(argv[11]);
    const int efix = atoi(argv[12]);
    const double vk = atof(argv[13]);
    const int lorder = atoi(argv[14]);

    Commons_Init();
    Solver_Parameters_Init();
    MeshReorder = 0;
    VariableType = VARIABLE_CONSERVATIVE;
    SolverMethod = SOLVER_METHOD_STEADY;
    SolverScheme = SOLVER_SCHEME_IMPLICIT;
    FluxScheme = (flux == "hllc") ? FLUX_SCHEME_HLLC : FLUX_SCHEME_ROE;
    PrecondMethod = PRECOND_METHOD_NONE;
    TimeIntegrationType = TIME_INTEGRATION_TYPE_LOCAL;
    TimeIntegrationMethod = TIME_INTEGRATION_METHOD_IMPLICIT_EULER;
    BCMethod = BC_METHOD_CHARCTERISTIC;
    JacobianMethod = JACOBIAN_METHOD_APPROX;
    JacobianUpdate = 1;
    SolverOrder = (order == 2) ? SOLVER_ORDER_SECOND : SOLVER_ORDER_FIRST;
    SolverNIteration = 1;
    LinearSolverNIteration = sweeps;
    Relaxation = 1.0;
    ResidualSmoothMethod = RESIDUAL_SMOOTH_METHOD_NONE;
    LimiterMethod = (lim == "venkat") ? LIMITER_METHOD_VENKATAKRISHNAN
                    : (lim == "bj")   ? LIMITER_METHOD_BERTH_JESPERSEN
                                      : LIMITER_METHOD_NONE;
    LimiterSmooth = 0;
    LimiterOrder = (lorder == 2) ? LIMITER_ORDER_SECOND : LIMITER_ORDER_FIRST;
    Venkat_KThreshold = vk;
    EntropyFix = efix;
    CFL_MIN = 1.0;
    CFL_MAX = 1000.0;
    CFL_Ramp = 50;
    Mach_Ramp = 0;
    ZPGIteration = 0;
    NonDimensionalMethod = NONDIMENSIONAL_METHOD_GENERIC;
    Gamma = 1.4;
    Ref_Mach = mach;
    Mach_MAX = mach;
    Mach_MIN = mach;
    Ref_Alpha = alpha;
    Ref_Beta = 0.0;
    Material_Set_Properties();

    UGrid_Reader(argv[1]);
    Create_Connectivity_Maps(MeshReorder);
    Calculate_Area_Volume();
    Trim_Connectivity_Memory();
    bndType = new int[nBC];
    for (int b = 0; b < nBC; b++) bndType[b] = (b < 2) ? BC_TYPE_FAR_FIELD : BC_TYPE_EULER_WALL;
    Solver_Init();
    Solver_Set_Initial_Conditions();
    if (FluxScheme == FLUX_SCHEME_ROE)
        Roe_Init_New();
    else
        HLLC_Init();

    const int N = nNode, NB = nBNode, E = nEdge, B = nBEdge;
    {
        std::vector<double> xyz(coordXYZ, coordXYZ + 3 * (size_t)N);
        dump_d("xyz", xyz, 3);
        std::vector<int> en(2 * (size_t)E), bn(2 * (size_t)B), bt(B), btag(B);
        std::vector<double> es(3 * (size_t)E), bs(3 * (size_t)B);
        for (int e = 0; e < E; e++) {
            en[2 * e] = intEdge[e].node[0];
            en[2 * e + 1] = intEdge[e].node[1];
            for (int d = 0; d < 3; d++) es[3 * e + d] = intEdge[e].areav.vec[d];
        }
        for (int e = 0; e < B; e++) {
            bn[2 * e] = bndEdge[e].node[0];
            bn[2 * e + 1] = bndEdge[e].node[1];
            bt[e] = bndEdge[e].type;
            btag[e] = bndEdge[e].tag;
            for (int d = 0; d < 3; d++) bs[3 * e + d] = bndEdge[e].areav.vec[d];
        }
        dump_i("edge", en, 2);
        dump_d("edge_area", es, 3);
        dump_i("bedge", bn, 2);
        dump_d("bedge_area", bs, 3);
        dump_i("bedge_type", bt);
        dump_i("bedge_tag", btag);
        dump_d("vol", std::vector<double>(cVolume, cVolume + N));
        dump_i("n2n_ia", std::vector<int>(crs_IA_Node2Node, crs_IA_Node2Node + N + 1));
        dump_i("n2n_ja", std::vector<int>(crs_JA_Node2Node, crs_JA_Node2Node + crs_IA_Node2Node[N]));
        std::vector<double> inf = {Inf_Rho, Inf_U, Inf_V, Inf_W, Inf_Pressure, Inf_Et, Inf_SpeedSound,
                                   Inf_Mach, Gamma, NonDim_R, Outflow_Pressure, Gauge_Pressure};
        dump_d("inf", inf);
    }

    // Perturbed, non-converged state on every node including ghosts.
    const double gm1 = Gamma - 1.0;
    for (int i = 0; i < N + NB; i++) {
        double rho = Inf_Rho * (1.0 + amp * rsym(seed, i, 0));
        double u = Inf_U + amp * Inf_SpeedSound * rsym(seed, i, 1);
        double v = Inf_V + amp * Inf_SpeedSound * rsym(seed, i, 2);
        double w = Inf_W + amp * Inf_SpeedSound * rsym(seed, i, 3);
        double p = Inf_Pressure * (1.0 + amp * rsym(seed, i, 4));
        double et = p / (gm1 * rho) + 0.5 * (u * u + v * v + w * w);
        Q1[i] = rho;
        Q2[i] = rho * u;
        Q3[i] = rho * v;
        Q4[i] = rho * w;
        Q5[i] = rho * et;
    }
    dump_d("q_in", gather5(Q1, Q2, Q3, Q4, Q5, N + NB), N + NB);

    // One pseudo-time iteration, Solver_Steady_Implicit.cpp:45-140.
    for (int i = 0; i < N; i++) {
        Res1[i] = Res2[i] = Res3[i] = Res4[i] = Res5[i] = 0.0;
        Res1_Diss[i] = Res2_Diss[i] = Res3_Diss[i] = Res4_Diss[i] = Res5_Diss[i] = 0.0;
        DeltaT[i] = 0.0;
    }
    MinEigenLamda1 = DBL_MAX;
    MaxEigenLamda1 = DBL_MIN;
    MinEigenLamda4 = DBL_MAX;
    MaxEigenLamda4 = DBL_MIN;
    MinEigenLamda5 = DBL_MAX;
    MaxEigenLamda5 = DBL_MIN;
    MinDeltaT = DBL_MAX;
    MaxDeltaT = DBL_MIN;
    ComputeFarFieldCondition(iter);

    Apply_Boundary_Condition(iter);
    dump_d("q_bc", gather5(Q1, Q2, Q3, Q4, Q5, N + NB), N + NB);
    // restart and solution files of this state (Euler3D/src/RestartIO.cpp:45-109, MeshIO.cpp:219-399)
    RestartIteration = iter;
    Restart_Writer((g_out + "/restart.srt").c_str(), 0);
    VTK_Writer((g_out + "/solution.vtk").c_str(), 0);
    if (SolverOrder == SOLVER_ORDER_SECOND) {
        Compute_Least_Square_Gradient(0);
        std::vector<double> g(15 * (size_t)N);
        double *src[15] = {Q1x, Q1y, Q1z, Q2x, Q2y, Q2z, Q3x, Q3y, Q3z, Q4x, Q4y, Q4z, Q5x, Q5y, Q5z};
        for (int k = 0; k < 15; k++)
            for (int i = 0; i < N; i++) g[(size_t)k * N + i] = src[k][i];
        dump_d("grad", g, N);
        if (LimiterMethod != LIMITER_METHOD_NONE) {
            Compute_Limiter();
            dump_d("phi", gather5(Limiter_Phi1, Limiter_Phi2, Limiter_Phi3, Limiter_Phi4, Limiter_Phi5, N), N);
        }
    }
    Compute_Residual(TRUE);
    dump_d("res_conv", gather5(Res1, Res2, Res3, Res4, Res5, N), N);
    dump_d("res_diss", gather5(Res1_Diss, Res2_Diss, Res3_Diss, Res4_Diss, Res5_Diss, N), N);
    dump_d("dt_acc", std::vector<double>(DeltaT, DeltaT + N));
    dump_d("eig", std::vector<double>{MinEigenLamda1, MaxEigenLamda1, MinEigenLamda4, MaxEigenLamda4,
                                      MinEigenLamda5, MaxEigenLamda5});
    Compute_DeltaT(iter);
    dump_d("dt", std::vector<double>(DeltaT, DeltaT + N));
    dump_d("cfl", std::vector<double>{CFL, MinDeltaT, MaxDeltaT});

    if (FluxScheme == FLUX_SCHEME_ROE) {
        // Euler3D/src/Jacobian.cpp:18-47 only dispatches Roe.
        Compute_Jacobian(TRUE, iter);
        const int DIM = SolverBlockMatrix.DIM;
        std::vector<double> A(25 * (size_t)DIM), Bv(5 * (size_t)N);
        for (int b = 0; b < DIM; b++)
            for (int r = 0; r < 5; r++)
                for (int c = 0; c < 5; c++) A[25 * (size_t)b + 5 * r + c] = SolverBlockMatrix.A[b][r][c];
        for (int i = 0; i < N; i++)
            for (int r = 0; r < 5; r++) Bv[5 * (size_t)i + r] = SolverBlockMatrix.B[i][r];
        dump_d("jac_a", A, 25);
        dump_d("jac_b", Bv, 5);
        dump_i("bsr_ia", std::vector<int>(SolverBlockMatrix.IA, SolverBlockMatrix.IA + N + 1));
        dump_i("bsr_ja", std::vector<int>(SolverBlockMatrix.JA, SolverBlockMatrix.JA + DIM));
        dump_i("bsr_iau", std::vector<int>(SolverBlockMatrix.IAU, SolverBlockMatrix.IAU + N));
        double rms = MC_Iterative_Block_LU_Jacobi_CRS(sweeps, 0, SolverBlockMatrix);
        std::vector<double> X(5 * (size_t)N);
        for (int i = 0; i < N; i++)
            for (int r = 0; r < 5; r++) X[5 * (size_t)i + r] = SolverBlockMatrix.X[i][r];
        dump_d("gs_x", X, 5);
        dump_d("gs_rms", std::vector<double>{rms});
        MC_Iterative_Block_LUSGS_CRS(sweeps, SolverBlockMatrix);
        for (int i = 0; i < N; i++)
            for (int r = 0; r < 5; r++) X[5 * (size_t)i + r] = SolverBlockMatrix.X[i][r];
        dump_d("lusgs_x", X, 5);
        // finite-difference Jacobians of Compute_RoeFlux (Euler3D/src/Roe_Jacobian.cpp:236-776),
        // every method of Compute_Jacobian_Roe (Euler3D/src/Roe_Jacobian.cpp:780-804)
        const struct {
            int method;
            const char *name;
        } fd[] = {{JACOBIAN_METHOD_CENTRAL, "central"}, {JACOBIAN_METHOD_FORWARD, "forward"},
                  {JACOBIAN_METHOD_BACKWARD, "backward"}, {JACOBIAN_METHOD_ALTERNATE, "alternate"}};
        for (const auto &f : fd) {
            JacobianMethod = f.method;
            Compute_Jacobian_FiniteDifference_Roe(TRUE, iter);
            for (int b = 0; b < DIM; b++)
                for (int r = 0; r < 5; r++)
                    for (int c = 0; c < 5; c++) A[25 * (size_t)b + 5 * r + c] = SolverBlockMatrix.A[b][r][c];
            for (int i = 0; i < N; i++)
                for (int r = 0; r < 5; r++) Bv[5 * (size_t)i + r] = SolverBlockMatrix.B[i][r];
            dump_d((std::string("jacfd_a_") + f.name).c_str(), A, 25);
            dump_d((std::string("jacfd_b_") + f.name).c_str(), Bv, 5);
        }
        // the ghost states after the perturbations (the reference restores them)
        dump_d("q_after_fd", gather5(Q1, Q2, Q3, Q4, Q5, N + NB), N + NB);
        JacobianMethod = JACOBIAN_METHOD_APPROX;
    }
    printf("ref_harness: N=%d E=%d B=%d done\n", N, E, B);
    return 0;
}
