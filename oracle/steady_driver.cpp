// Test-infrastructure driver: the reference's own steady implicit loop,
// Solver_Steady_Implicit (Euler3D/src/Solver_Steady_Implicit.cpp:33-168), run
// on a UGRID mesh for a fixed number of pseudo-time iterations.
//
// Built twice by oracle/Makefile from the same source:
//   _ref/steady_ref   linked with the reference's objects as they are (CPU);
//   _ref/steady_shim  linked with include/ref_shim/cfd_reference_shim.cpp and
//                     libcfd_hip.so, the reference objects' definitions of the
//                     replaced functions weakened (-DCFD_SHIM), so the same
//                     reference loop calls the MI355X path through the shim.
// Both set the parameter globals directly (as Euler3D/main.cpp:94-158 would
// from a deck), start from the freestream (Solver_Set_Initial_Conditions),
// run the loop in `outdir` (which receives the reference's Residual.res) and
// write the final state Q1..Q5 over the physical and ghost nodes as
// outdir/q_final.npy.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Commons.h"
#include "Gradient.h"
#include "HLLC_Fluxes.h"
#include "MC.h"
#include "Material.h"
#include "MeshIO.h"
#include "Roe_Fluxes.h"
#include "Solver.h"
#include "SolverParameters.h"
#include "Trim_Utils.h"

int Solver_Steady_Implicit(void);  // Euler3D/src/Solver_Steady_Implicit.cpp:33
#ifdef CFD_SHIM
void CFD_Shim_Init(int device);
void CFD_Shim_Finalize(void);
#endif

static void npy_write_q(const std::string &path, int n) {
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    char shape[64];
    snprintf(shape, sizeof(shape), "(5, %d)", n);
    std::string hdr = std::string("{'descr': '<f8', 'fortran_order': False, 'shape': ") + shape + ", }";
    const size_t total = 10 + hdr.size() + 1;
    hdr += std::string((64 - total % 64) % 64, ' ') + "\n";
    const unsigned short hl = (unsigned short)hdr.size();
    fwrite("\x93NUMPY\x01\x00", 1, 8, fp);
    fwrite(&hl, 2, 1, fp);
    fwrite(hdr.data(), 1, hdr.size(), fp);
    double *q[5] = {Q1, Q2, Q3, Q4, Q5};
    for (int k = 0; k < 5; k++) fwrite(q[k], sizeof(double), (size_t)n, fp);
    fclose(fp);
}

int main(int argc, char **argv) {
    if (argc < 11) {
        fprintf(stderr,
                "usage: %s mesh.ugrid outdir flux(roe|hllc) order(1|2) limiter(none|venkat|bj) mach alpha_deg "
                "iterations gs_sweeps jacobian(approx|central|forward|backward|alternate)\n",
                argv[0]);
        return 1;
    }
    const std::string mesh = argv[1], out = argv[2], flux = argv[3], lim = argv[5], jac = argv[10];
    const int order = atoi(argv[4]);
    const double mach = atof(argv[6]), alpha = atof(argv[7]);
    const int iters = atoi(argv[8]), sweeps = atoi(argv[9]);

    Commons_Init();
    Solver_Parameters_Init();
    MeshReorder = 0;
    VariableType = VARIABLE_CONSERVATIVE;
    SolverMethod = SOLVER_METHOD_STEADY;
    SolverScheme = SOLVER_SCHEME_IMPLICIT;
    FluxScheme = flux == "hllc" ? FLUX_SCHEME_HLLC : FLUX_SCHEME_ROE;
    PrecondMethod = PRECOND_METHOD_NONE;
    TimeIntegrationType = TIME_INTEGRATION_TYPE_LOCAL;
    TimeIntegrationMethod = TIME_INTEGRATION_METHOD_IMPLICIT_EULER;
    BCMethod = BC_METHOD_CHARCTERISTIC;
    JacobianMethod = jac == "central"     ? JACOBIAN_METHOD_CENTRAL
                     : jac == "forward"   ? JACOBIAN_METHOD_FORWARD
                     : jac == "backward"  ? JACOBIAN_METHOD_BACKWARD
                     : jac == "alternate" ? JACOBIAN_METHOD_ALTERNATE
                                          : JACOBIAN_METHOD_APPROX;
    JacobianUpdate = 1;
    SolverOrder = order == 2 ? SOLVER_ORDER_SECOND : SOLVER_ORDER_FIRST;
    SolverNIteration = iters;
    FirstOrderNIteration = 0;
    LinearSolverNIteration = sweeps;
    Relaxation = 1.0;
    ResidualSmoothMethod = RESIDUAL_SMOOTH_METHOD_NONE;
    LimiterMethod = lim == "venkat" ? LIMITER_METHOD_VENKATAKRISHNAN
                    : lim == "bj"   ? LIMITER_METHOD_BERTH_JESPERSEN
                                    : LIMITER_METHOD_NONE;
    LimiterSmooth = 0;
    LimiterOrder = LIMITER_ORDER_FIRST;
    LimiterStartSolverIteration = 0;
    LimiterEndSolverIteration = iters + 1;
    Venkat_KThreshold = 5.0;
    EntropyFix = 0;
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

    UGrid_Reader(mesh.c_str());
    Create_Connectivity_Maps(MeshReorder);
    Calculate_Area_Volume();
    Trim_Connectivity_Memory();
    // boundary conditions by surface tag: x = 0 and x = 3 far field, the rest Euler walls
    bndType = new int[nBC];
    for (int b = 0; b < nBC; b++) bndType[b] = b < 2 ? BC_TYPE_FAR_FIELD : BC_TYPE_EULER_WALL;
    if (chdir(out.c_str()) != 0) {
        fprintf(stderr, "cannot enter %s\n", out.c_str());
        return 2;
    }
    Solver_Init();  // opens Residual.res in the current directory
    Solver_Set_Initial_Conditions();
    if (FluxScheme == FLUX_SCHEME_ROE)
        Roe_Init_New();
    else
        HLLC_Init();
#ifdef CFD_SHIM
    CFD_Shim_Init(0);
#endif
    Solver_Steady_Implicit();
    RMS_Writer_Finalize();
    npy_write_q("q_final.npy", nNode + nBNode);
#ifdef CFD_SHIM
    CFD_Shim_Finalize();
#endif
    printf("steady_driver: N=%d E=%d B=%d iterations=%d done\n", nNode, nEdge, nBEdge, iters);
    return 0;
}
