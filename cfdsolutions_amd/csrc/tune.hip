// Autotuner: for one device mesh and one set of parameters, times each
// candidate of the kernel variants that compute the same bits, and keeps the
// fastest for the mesh's later calls.  The choices:
//   - the fused Green-Gauss gradient + limiter: neighbour records gathered
//     directly, or staged in LDS with direct global->LDS loads;
//   - the Krylov loop's layout: BSR (five lanes per block row) or the
//     block-SELL copies (one row per lane), for the SpMV and, with the
//     multicolour ILU(0), for the colour sweeps;
//   - with the natural-order ILU(0): single-launch sync-free triangular
//     sweeps, or one launch per level.
// Every candidate gives the same results bit for bit (each pair is tested
// against the other), so tuning changes times only, never iterates.  Without
// tuning the mesh keeps fixed rules (direct gathers in FP64 and staged in
// FP32, block-SELL from 131,072 rows, sync-free sweeps); a CFD_STAGE_GL,
// CFD_SELL or CFD_SYNC_FREE setting in the environment fixes that choice and
// is not tuned.
#include "cfd_internal.h"

#include <chrono>
#include <cstdlib>

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int sync_stream(cfd_mesh *m) {
    CfdRuntimeGuard g;
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    return CFD_OK;
}

// Best time (ms) of `reps` runs of fn after one warm-up run (which may
// allocate workspaces), host clock around a synchronised stream.
template <class F>
int best_of(cfd_mesh *m, int reps, F fn, double *ms) {
    int rc = fn();
    if (!rc) rc = sync_stream(m);
    double best = 1e300;
    for (int r = 0; !rc && r < reps; r++) {
        const double t0 = now_ms();
        rc = fn();
        if (!rc) rc = sync_stream(m);
        best = std::min(best, now_ms() - t0);
    }
    *ms = best;
    return rc;
}

// The Krylov layout or sweep kind changed: factors, SELL copies and captured
// Arnoldi steps belong to the old one.
void krylov_changed(cfd_mesh *m) {
    m->asell_valid = false;
    m->prec_valid = false;
    drop_krylov_graphs(m);
}

bool env_set(const char *name) {
    const char *e = getenv(name);
    return e && (e[0] == '0' || e[0] == '1');
}

}  // namespace

int autotune_mesh(cfd_mesh *m, const cfd_params *p, const cfd_linsolve_opts *lin, cfd_tune_report *rep,
                  const std::function<int()> &linsolve) {
    *rep = cfd_tune_report{};
    rep->gl_stage = rep->sell = rep->sync_free = -1;
    int rc = CFD_OK;
    // fused gradient + limiter (Green-Gauss with a limiter, second order)
    if (p->order == 2 && p->gradient == CFD_GRAD_GG && p->limiter != CFD_LIM_NONE && !env_set("CFD_STAGE_GL")) {
        int8_t &stage = m->tune_gl_stage[p->precision == CFD_FP32 ? 1 : 0];
        for (int s = 0; s < 2 && !rc; s++) {
            stage = (int8_t)s;
            rc = best_of(m, 3, [&] { return cfd_gradient_limiter(m, p); }, &rep->gl_ms[s]);
        }
        if (rc) {
            stage = -1;
            return rc;
        }
        stage = rep->gl_ms[1] < rep->gl_ms[0] ? 1 : 0;
        rep->gl_stage = stage;
    }
    // Krylov layout and sweep kind, timed on whole linear solves of the
    // assembled system (factorisation and SELL copies included, as in each
    // Newton step)
    if (lin && lin->solver == CFD_SOLVER_FGMRES && m->jac_valid) {
        const bool sell_free = !env_set("CFD_SELL");
        const bool sf_free = lin->precond == CFD_PREC_ILU0 && !env_set("CFD_SYNC_FREE");
        const bool sell0 = m->use_sell, sf0 = m->sync_free;
        double best = 1e300;
        bool best_sell = sell0, best_sf = sf0;
        for (int a = 0; a < 2 && !rc; a++) {
            if (!sell_free && (a == 1) != sell0) continue;
            for (int b = 0; b < 2 && !rc; b++) {
                if (!sf_free && (b == 0) != sf0) continue;
                m->use_sell = a == 1;
                m->sync_free = b == 0;
                krylov_changed(m);
                double t = 0.0;
                rc = best_of(m, 1, linsolve, &t);
                rep->krylov_ms[a][b] = t;
                if (!rc && t < best) {
                    best = t;
                    best_sell = m->use_sell;
                    best_sf = m->sync_free;
                }
            }
        }
        m->use_sell = rc ? sell0 : best_sell;
        m->sync_free = rc ? sf0 : best_sf;
        krylov_changed(m);
        if (rc) return rc;
        rep->sell = m->use_sell;
        rep->sync_free = m->sync_free;
    }
    return CFD_OK;
}

extern "C" int cfd_autotune(cfd_mesh *m, const cfd_params *p, const cfd_linsolve_opts *lin, cfd_tune_report *rep) {
    if (!m || !p) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    cfd_tune_report tmp;
    cfd_linsolve_stats st;
    return autotune_mesh(m, p, lin, rep ? rep : &tmp, [&] { return run_linsolve(m, lin, &st); });
}
