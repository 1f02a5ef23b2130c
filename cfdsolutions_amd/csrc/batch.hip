// batch.hip - steady solves of many small cases at once.
//
// The cases are held as one union mesh: their nodes, edges and boundary
// edges are concatenated (each case keeps a contiguous range in every
// numbering), so the BSR system is block-diagonal and one launch of every
// residual, Jacobian, ILU(0) and Krylov kernel covers all cases.  Each case
// keeps its own freestream (boundary condition table), its own residual norm,
// Krylov dots, Givens rotations, convergence and forces.  Every per-node and
// per-row operation is the single-mesh one, and every per-case reduction
// uses the partition a launch on the case alone would use, so each case's
// iterates are bit-identical to solving it by itself.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cfd_internal.h"
#include "flux.h"

namespace {

constexpr int kBlock = 256;

int grid_for(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    return (int)(g < 4096 ? (g ? g : 1) : 4096);
}

template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *out) {
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

// Per-case sums of squares of Conv+Diss: k_rms_partial on each case's node
// range with the case's own block count.
__global__ __launch_bounds__(kBlock) void k_rms_seg(const int2 *blk, const int *first, const int *off, const int *len,
                                                    const double *res, double *part) {
    const int2 cb = blk[blockIdx.x];
    const int P = first[cb.x + 1] - first[cb.x], o = off[cb.x], n = len[cb.x];
    double v[5] = {0, 0, 0, 0, 0};
    for (int i = cb.y * kBlock + threadIdx.x; i < n; i += P * kBlock) {
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const double r = res[kRES * (size_t)(o + i) + k] + res[kRES * (size_t)(o + i) + 5 + k];
            v[k] += r * r;
        }
    }
    block_reduce<5>(v, part + 5 * blockIdx.x);
}

// Pressure force on the wall edges of each case: k_forces_partial on each
// case's boundary-edge range.
__global__ __launch_bounds__(kBlock) void k_forces_seg(const int2 *blk, const int *first, const int *off,
                                                       const int *len, const int *bnode, const int *btype,
                                                       const int *btag, int wall_tag, const double *sa,
                                                       const double *qa, double g, double *part) {
    const int2 cb = blk[blockIdx.x];
    const int P = first[cb.x + 1] - first[cb.x], o = off[cb.x], n = len[cb.x];
    double v[4] = {0, 0, 0, 0};
    for (int t = cb.y * kBlock + threadIdx.x; t < n; t += P * kBlock) {
        const int e = o + t;
        const bool use = wall_tag > 0 ? (btag && btag[e] == wall_tag) : (btype[e] == 0);
        if (!use) continue;
        const double *r = qa + kQA * (size_t)bnode[e];
        cfd::Eos<double> s;
        cfd::eos_face(g, r, 0.0, 0.0, 0.0, s);
        const double *S = sa + 4 * (size_t)e;
        v[0] += s.p * S[0];
        v[1] += s.p * S[1];
        v[2] += s.p * S[2];
        v[3] += s.p * sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    }
    block_reduce<4>(v, part + 4 * blockIdx.x);
}

// Sum of a case's partials in the order of k_reduce_final (thread-strided,
// then a block tree).
template <int NV>
__global__ __launch_bounds__(kBlock) void k_reduce_seg(const int *first, const double *part, double *out) {
    const int c = blockIdx.x, p0 = first[c], np = first[c + 1] - p0;
    double v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = 0.0;
    for (int p = threadIdx.x; p < np; p += kBlock)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] += part[NV * (size_t)(p0 + p) + k];
    block_reduce<NV>(v, out + (size_t)c * kBOut);
}

// Q += relax * X on the nodes of the cases still iterating.
__global__ __launch_bounds__(kBlock) void k_update_seg(int N, double relax, const int *order, const int *ecase,
                                                       const int *active, const double *x, double *qa) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const size_t r = order[i];
    if (!active[ecase[r]]) return;
#pragma unroll
    for (int k = 0; k < 5; k++) qa[kQA * (size_t)i + k] += relax * x[5 * r + k];
}

// Sets the nodes [n0, n0 + nn) and the ghosts [g0, g0 + ng) to the state q
// (a failed case of a union is parked at its freestream, so its rows stay
// finite and raise no more flags while it rides along).
__global__ __launch_bounds__(kBlock) void k_park_case(int n0, int nn, int g0, int ng, double q0, double q1, double q2,
                                                      double q3, double q4, double *qa) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= nn + ng) return;
    const size_t i = t < nn ? (size_t)(n0 + t) : (size_t)(g0 + t - nn);
    double *r = qa + kQA * i;
    r[0] = q0;
    r[1] = q1;
    r[2] = q2;
    r[3] = q3;
    r[4] = q4;
}

// Segment map: case c gets nb[c] blocks over [off[c], off[c] + len[c]).
int build_segmap(cfd_mesh *m, SegMap &sg, const std::vector<int> &off, const std::vector<int> &len,
                 const std::vector<int> &nb) {
    const int K = (int)off.size();
    sg.h_first.assign(K + 1, 0);
    for (int c = 0; c < K; c++) sg.h_first[c + 1] = sg.h_first[c] + nb[c];
    sg.nblocks = sg.h_first[K];
    std::vector<int2> blk(sg.nblocks);
    for (int c = 0; c < K; c++)
        for (int b = 0; b < nb[c]; b++) blk[sg.h_first[c] + b] = make_int2(c, b);
    sg.blk = (int2 *)cfd_dalloc(m, sizeof(int2) * blk.size());
    sg.first = (int *)cfd_dalloc(m, sizeof(int) * (K + 1));
    sg.off = (int *)cfd_dalloc(m, sizeof(int) * K);
    sg.len = (int *)cfd_dalloc(m, sizeof(int) * K);
    if (!sg.blk || !sg.first || !sg.off || !sg.len) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory");
    CFD_HIP(m, cfd_memcpy(sg.blk, blk.data(), sizeof(int2) * blk.size(), hipMemcpyHostToDevice));
    CFD_HIP(m, cfd_memcpy(sg.first, sg.h_first.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice));
    CFD_HIP(m, cfd_memcpy(sg.off, off.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    CFD_HIP(m, cfd_memcpy(sg.len, len.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    return CFD_OK;
}

}  // namespace

bool cfd_same_scheme(const cfd_params &a, const cfd_params &b) {
    return a.gamma == b.gamma && a.flux == b.flux && a.order == b.order && a.limiter == b.limiter &&
           a.limiter_order == b.limiter_order && a.entropy_fix == b.entropy_fix && a.gradient == b.gradient &&
           a.zpg_iter == b.zpg_iter && a.precision == b.precision && a.arith == b.arith &&
           a.venkat_k == b.venkat_k && a.cfl_min == b.cfl_min && a.cfl_max == b.cfl_max && a.cfl_ramp == b.cfl_ramp;
}

bool cfd_same_solve(const cfd_solve_opts &a, const cfd_solve_opts &b) {
    return a.max_iter == b.max_iter && a.tol == b.tol && a.first_order_iters == b.first_order_iters &&
           a.relaxation == b.relaxation && a.wall_tag == b.wall_tag && a.lin.solver == b.lin.solver &&
           a.lin.precond == b.lin.precond && a.lin.krylov_m == b.lin.krylov_m &&
           a.lin.max_restarts == b.lin.max_restarts && a.lin.rtol == b.lin.rtol && a.lin.mixed == b.lin.mixed &&
           a.jacobian == b.jacobian;
}

namespace {

// One union mesh over a subset of the cases of a batch.
struct Union {
    cfd_mesh *m = nullptr;
    std::vector<int> cases;       // batch index of each union case
    std::vector<int> noff, boff;  // node and boundary-edge offsets of each union case
    int *d_active = nullptr;      // device [K] 1 while a union case still iterates
};

// Builds the union mesh of the batch cases `cases`.  Each case starts from
// q_init[c] ([5][n_c + b_c], nodes then ghosts) when that is not empty, else
// from its freestream.
int build_union(cfd_ctx *ctx, const std::vector<int> &cases, const cfd_mesh_desc *meshes, const cfd_params *params,
                const std::vector<std::vector<double>> &q_init, bool allow_graphs, Union &u) {
    const int K = (int)cases.size();
    auto tph = std::chrono::steady_clock::now();
    auto phase = [&](const char *what) {
        if (!cfd_progress()) return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "cfd_batch_solve: %-28s %8.3f s\n", what, std::chrono::duration<double>(t - tph).count());
        tph = t;
    };
    u = Union{};
    u.cases = cases;
    std::vector<int> &noff = u.noff, &boff = u.boff, eoff(K + 1, 0);
    noff.assign(K + 1, 0);
    boff.assign(K + 1, 0);
    for (int c = 0; c < K; c++) {
        const cfd_mesh_desc &d = meshes[cases[c]];
        noff[c + 1] = noff[c] + d.n_node;
        eoff[c + 1] = eoff[c] + d.n_edge;
        boff[c + 1] = boff[c] + d.n_bedge;
    }
    const int N = noff[K], E = eoff[K], B = boff[K];
    // every entry is written below, case by case in parallel
    HostArray<int> edge(2 * (size_t)E), bedge(2 * (size_t)B), btype(B), btag(B), ia((size_t)N + 1),
        ja(2 * (size_t)E), node_case(N), bcase(B);
    HostArray<double> earea(3 * (size_t)E), barea(3 * (size_t)B), xyz(3 * (size_t)N), vol(N);
    ia[0] = 0;
    // each case fills its own ranges; its n2n lists start at entry 2 * eoff[c]
    cfd_parallel_for(
        K,
        [&](int64_t c0, int64_t c1) {
            for (int c = (int)c0; c < (int)c1; c++) {
                const cfd_mesh_desc &d = meshes[cases[c]];
                const int no = noff[c], eo = eoff[c], bo = boff[c];
                for (int e = 0; e < d.n_edge; e++) {
                    edge[2 * (size_t)(eo + e)] = d.edge[2 * e] + no;
                    edge[2 * (size_t)(eo + e) + 1] = d.edge[2 * e + 1] + no;
                }
                memcpy(&earea[3 * (size_t)eo], d.edge_area, sizeof(double) * 3 * d.n_edge);
                for (int b = 0; b < d.n_bedge; b++) {
                    bedge[2 * (size_t)(bo + b)] = d.bedge[2 * b] + no;
                    bedge[2 * (size_t)(bo + b) + 1] = N + bo + b;
                    btype[bo + b] = d.bedge_type[b];
                    btag[bo + b] = d.bedge_tag ? d.bedge_tag[b] : 0;
                    bcase[bo + b] = c;
                }
                memcpy(&barea[3 * (size_t)bo], d.bedge_area, sizeof(double) * 3 * d.n_bedge);
                memcpy(&xyz[3 * (size_t)no], d.xyz, sizeof(double) * 3 * d.n_node);
                memcpy(&vol[no], d.vol, sizeof(double) * d.n_node);
                for (int i = 0; i < d.n_node; i++) {
                    ia[no + i + 1] = 2 * eo + d.n2n_ia[i + 1];
                    node_case[no + i] = c;
                }
                for (int k = 0; k < d.n2n_ia[d.n_node]; k++) ja[2 * (size_t)eo + k] = d.n2n_ja[k] + no;
            }
        },
        1);
    cfd_mesh_desc ud{N, E, B, edge.data(), earea.data(), bedge.data(), barea.data(), btype.data(), btag.data(),
                     xyz.data(), vol.data(), ia.data(), ja.data()};
    phase("union assembly");
    int rc = cfd_mesh_create_cases(ctx, &ud, &params[cases[0]], node_case.data(), &u.m);
    if (rc) return rc;
    phase("union mesh");
    cfd_mesh *m = u.m;
    if (!allow_graphs) m->use_graphs = false;
    m->ncase = K;
    m->case_node = noff;
    m->case_bedge = boff;
    std::vector<double> inf((size_t)K * 8, 0.0);
    for (int c = 0; c < K; c++) {
        const cfd_params &p = params[cases[c]];
        double *f = &inf[(size_t)c * 8];
        f[0] = p.inf_rho;
        f[1] = p.inf_u;
        f[2] = p.inf_v;
        f[3] = p.inf_w;
        f[4] = p.inf_p;
        f[5] = p.outflow_p;
    }
    std::vector<int> zoff(K), zlen(K), nb(K), eoffv(K), elen(K), nbd(K), bo(K), blen(K), nbf(K);
    for (int c = 0; c < K; c++) {
        const int nc = noff[c + 1] - noff[c], bc = boff[c + 1] - boff[c];
        zoff[c] = noff[c];
        zlen[c] = nc;
        nb[c] = std::min(grid_for(nc), 1024);  // launch_rms
        eoffv[c] = 5 * noff[c];
        elen[c] = 5 * nc;
        nbd[c] = dot_parts(5 * (size_t)nc);  // mdot
        bo[c] = boff[c];
        blen[c] = bc;
        nbf[c] = bc ? std::min(grid_for(bc), 256) : 1;  // launch_forces
    }
    m->case_inf = (double *)cfd_dalloc(m, sizeof(double) * inf.size());
    m->bcase = (int *)cfd_dalloc(m, sizeof(int) * std::max(1, B));
    m->ecase = (int *)cfd_dalloc(m, sizeof(int) * N);
    m->bred = (double *)cfd_dalloc(m, sizeof(double) * kBOut * K);
    m->bcoef = (double *)cfd_dalloc(m, sizeof(double) * (kMaxKrylov + 1) * K);
    m->bkdim = (int *)cfd_dalloc(m, sizeof(int) * K);
    u.d_active = (int *)cfd_dalloc(m, sizeof(int) * K);
    m->case_flag = (int *)cfd_dalloc(m, sizeof(int) * K);
    if (!m->case_inf || !m->bcase || !m->ecase || !m->bred || !m->bcoef || !m->bkdim || !u.d_active ||
        !m->case_flag)
        return cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of device memory");
    if (cfd_host_alloc((void **)&m->hbred, sizeof(double) * kBOut * K) != hipSuccess ||
        cfd_host_alloc((void **)&m->hbcoef, sizeof(double) * (kMaxKrylov + 1) * K) != hipSuccess ||
        cfd_host_alloc((void **)&m->hbkdim, sizeof(int) * K) != hipSuccess)
        return cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of pinned host memory");
    const std::vector<int> ones(K, 1), zeros(K, 0);
    if (cfd_memcpy(m->case_inf, inf.data(), sizeof(double) * inf.size(), hipMemcpyHostToDevice) != hipSuccess ||
        (B && cfd_memcpy(m->bcase, bcase.data(), sizeof(int) * B, hipMemcpyHostToDevice) != hipSuccess) ||
        cfd_memcpy(m->ecase, node_case.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess ||
        cfd_memcpy(u.d_active, ones.data(), sizeof(int) * K, hipMemcpyHostToDevice) != hipSuccess ||
        cfd_memcpy(m->case_flag, zeros.data(), sizeof(int) * K, hipMemcpyHostToDevice) != hipSuccess)
        return cfd_set_error(ctx, CFD_E_HIP, "hipMemcpy failed");
    m->case_active = u.d_active;
    rc = build_segmap(m, m->seg_rms, zoff, zlen, nb);
    if (!rc) rc = build_segmap(m, m->seg_dot, eoffv, elen, nbd);
    if (!rc) rc = build_segmap(m, m->seg_forces, bo, blen, nbf);
    if (rc) return rc;
    const size_t np = std::max({(size_t)m->seg_dot.nblocks * (kMaxKrylov + 1), (size_t)m->seg_rms.nblocks * 5,
                                (size_t)m->seg_forces.nblocks * 4});
    m->bpart = (double *)cfd_dalloc(m, sizeof(double) * np);
    if (!m->bpart) return cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of device memory");
    // initial state of every case, nodes and ghosts (every entry written below)
    HostArray<double> q(5 * (size_t)(N + B));
    for (int c = 0; c < K; c++) {
        const size_t nc = (size_t)(noff[c + 1] - noff[c]), bc = (size_t)(boff[c + 1] - boff[c]), tc = nc + bc;
        const std::vector<double> &qi = q_init[cases[c]];
        for (int k = 0; k < 5; k++) {
            double *dn = q.data() + (size_t)k * (N + B) + noff[c], *dg = q.data() + (size_t)k * (N + B) + N + boff[c];
            if (!qi.empty()) {
                memcpy(dn, qi.data() + k * tc, sizeof(double) * nc);
                memcpy(dg, qi.data() + k * tc + nc, sizeof(double) * bc);
                continue;
            }
            const cfd_params &pc = params[cases[c]];
            double et = pc.inf_u * pc.inf_u + pc.inf_v * pc.inf_v + pc.inf_w * pc.inf_w;
            et = pc.inf_p / ((pc.gamma - 1.0) * pc.inf_rho) + 0.5 * et;
            const double v[5] = {pc.inf_rho, pc.inf_rho * pc.inf_u, pc.inf_rho * pc.inf_v, pc.inf_rho * pc.inf_w,
                                 pc.inf_rho * et};
            std::fill(dn, dn + nc, v[k]);
            std::fill(dg, dg + bc, v[k]);
        }
    }
    rc = cfd_field_set(m, CFD_FIELD_Q, q.data());
    phase("segments and state");
    return rc;
}

// Final results of the union cases in `take` (union indices): forces, time
// and, where asked or where keep[c] is set, the state [5][n_c + b_c] of each.
int harvest(Union &u, const std::vector<int> &take, const cfd_solve_opts *opts, const cfd_params &p, double secs,
            cfd_solve_result *results, double *const *q_out, const std::vector<char> &keep,
            std::vector<std::vector<double>> &q_keep) {
    cfd_mesh *m = u.m;
    const int K = (int)u.cases.size();
    hipLaunchKernelGGL(k_forces_seg, dim3(m->seg_forces.nblocks), dim3(kBlock), 0, m->stream, m->seg_forces.blk,
                       m->seg_forces.first, m->seg_forces.off, m->seg_forces.len, m->bnode, m->btype, m->btag,
                       opts[0].wall_tag, m->d.sa, m->d.qa, p.gamma, m->bpart);
    hipLaunchKernelGGL(k_reduce_seg<4>, dim3(K), dim3(kBlock), 0, m->stream, m->seg_forces.first, m->bpart, m->bred);
    if (hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * K, hipMemcpyDeviceToHost, m->stream) !=
            hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return cfd_set_error(m->ctx, CFD_E_HIP, "cfd_batch_solve: reading forces failed");
    bool need_q = false;
    for (int lc : take) need_q = need_q || (q_out && q_out[u.cases[lc]] && !results[u.cases[lc]].status);
    for (int lc = 0; lc < K; lc++) need_q = need_q || keep[u.cases[lc]];
    for (int lc : take) {
        const int c = u.cases[lc];
        cfd_solve_result &r = results[c];
        r.seconds = secs;
        if (r.status) continue;  // a failed solve reports no forces (as cfd_solve_steady)
        const double *v = m->hbred + (size_t)lc * kBOut;
        const double a = opts[c].alpha_deg * M_PI / 180.0;
        r.force[0] = v[0];
        r.force[1] = v[1];
        r.force[2] = v[2];
        r.drag = v[0] * cos(a) + v[1] * sin(a);
        r.lift = -v[0] * sin(a) + v[1] * cos(a);
        r.pressure_integral = v[3];
    }
    if (!need_q) return CFD_OK;
    const int N = m->N, B = m->B;
    const size_t NT = (size_t)N + B;
    std::vector<double> all(5 * NT);
    const int rc = cfd_field_get(m, CFD_FIELD_Q, all.data());
    if (rc) return rc;
    auto extract = [&](int lc, double *dst) {
        const size_t nc = (size_t)(u.noff[lc + 1] - u.noff[lc]), bc = (size_t)(u.boff[lc + 1] - u.boff[lc]),
                     tc = nc + bc;
        for (int k = 0; k < 5; k++) {
            const double *src = all.data() + k * NT;
            memcpy(dst + k * tc, src + u.noff[lc], sizeof(double) * nc);
            memcpy(dst + k * tc + nc, src + N + u.boff[lc], sizeof(double) * bc);
        }
    };
    for (int lc : take)
        if (q_out && q_out[u.cases[lc]] && !results[u.cases[lc]].status) extract(lc, q_out[u.cases[lc]]);
    for (int lc = 0; lc < K; lc++) {
        const int c = u.cases[lc];
        if (!keep[c]) continue;
        q_keep[c].resize(5 * (size_t)(u.noff[lc + 1] - u.noff[lc] + u.boff[lc + 1] - u.boff[lc]));
        extract(lc, q_keep[c].data());
    }
    return CFD_OK;
}

}  // namespace

int cfd_batch_solve_into(cfd_ctx *ctx, int32_t K, const cfd_mesh_desc *meshes, const cfd_params *params,
                         const cfd_solve_opts *opts, cfd_solve_result *results, double *const *q_out,
                         int64_t *device_bytes, bool allow_graphs, cfd_case_done_fn done, void *done_arg) {
    if (!ctx || K < 1 || !meshes || !params || !opts || !results)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: invalid arguments");
    for (int c = 1; c < K; c++)
        if (!cfd_same_scheme(params[0], params[c]) || !cfd_same_solve(opts[0], opts[c]))
            return cfd_set_error(ctx, CFD_E_ARG,
                                 "cfd_batch_solve: case %d differs from case 0 in scheme or solver options", c);
    const cfd_solve_opts &o = opts[0];
    if (o.lin.solver != CFD_SOLVER_FGMRES)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: batched solves use FGMRES");
    const auto t0 = std::chrono::steady_clock::now();
    auto elapsed = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    int64_t tn = 0, te = 0, tb = 0;
    for (int c = 0; c < K; c++) {
        const cfd_mesh_desc &d = meshes[c];
        if (d.n_node <= 0 || d.n_edge < 0 || d.n_bedge < 0 || !d.n2n_ia || !d.n2n_ja || d.n2n_ia[0] != 0 ||
            d.n2n_ia[d.n_node] != 2 * d.n_edge)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: case %d has an invalid mesh", c);
        for (int b = 0; b < d.n_bedge; b++)
            if (d.bedge[2 * b + 1] != d.n_node + b)
                return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: case %d boundary edge %d ghost index", c, b);
        tn += d.n_node;
        te += d.n_edge;
        tb += d.n_bedge;
        if (tn + tb > INT32_MAX / 2 || 2 * te + tb > INT32_MAX / 2)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: union mesh too large for int32 indices");
    }
    // Compaction: once the cases still iterating hold at most half of the
    // union's nodes, their states move to a new union of just those cases
    // and the finished ones are harvested.  A case's arithmetic does not
    // depend on the other cases of its union, so its iterates are unchanged.
    // CFD_BATCH_COMPACT=0 never compacts.
    const char *ce = getenv("CFD_BATCH_COMPACT");
    const int64_t compact_min = ce && ce[0] == '0' ? INT64_MAX : 0;
    for (int c = 0; c < K; c++) results[c] = cfd_solve_result{};
    std::vector<std::vector<double>> q_keep(K);
    std::vector<char> active(K, 1), keep(K, 0);
    std::vector<int> all(K);
    for (int c = 0; c < K; c++) all[c] = c;
    Union u;
    int rc = build_union(ctx, all, meshes, params, q_keep, allow_graphs, u);
    if (device_bytes) *device_bytes = u.m ? u.m->bytes : 0;
    std::vector<char> solve;
    std::vector<int> iters, act_i, lstatus;
    std::vector<double> fres;
    cfd_params p = params[0];
    p.precision = CFD_FP64;
    // Failure isolation.  A case that fails ends with the status and
    // iteration count its solve alone would report (cfd_solve_steady), is
    // parked at its freestream so its rows stay finite and raise no more
    // flags, and is dropped at the next compaction; the other cases go on
    // unaffected, since no case's arithmetic depends on another's.
    auto park = [&](int lc, int status, int it) -> int {
        const int c = u.cases[lc];
        results[c].status = status;
        results[c].iterations = it;
        active[c] = 0;
        const cfd_params &pc = params[c];
        double et = pc.inf_u * pc.inf_u + pc.inf_v * pc.inf_v + pc.inf_w * pc.inf_w;
        et = pc.inf_p / ((pc.gamma - 1.0) * pc.inf_rho) + 0.5 * et;
        const int nn = u.noff[lc + 1] - u.noff[lc], ng = u.boff[lc + 1] - u.boff[lc];
        hipLaunchKernelGGL(k_park_case, dim3((nn + ng + kBlock - 1) / kBlock), dim3(kBlock), 0, u.m->stream,
                           u.noff[lc], nn, u.m->N + u.boff[lc], ng, pc.inf_rho, pc.inf_rho * pc.inf_u,
                           pc.inf_rho * pc.inf_v, pc.inf_rho * pc.inf_w, pc.inf_rho * et, u.m->d.qa);
        u.m->q_version++;
        return hipGetLastError() == hipSuccess ? CFD_OK
                                               : cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: parking a case failed");
    };
    // cfd_bc and the residual checks report a failure for the whole union;
    // the per-case failure bits name the cases it belongs to, which then fail
    // on their own with the code a solve of that case alone returns.  If no
    // case is named, the failure stands for the whole batch.
    auto isolate = [&](int rc0, int it) -> int {
        cfd_mesh *m = u.m;
        const int Ku = (int)u.cases.size();
        std::vector<int> bits(Ku, 0);
        if (hipMemcpyAsync(bits.data(), m->case_flag, sizeof(int) * Ku, hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess ||
            hipMemsetAsync(m->case_flag, 0, sizeof(int) * Ku, m->stream) != hipSuccess)
            return cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: reading case flags failed");
        bool named = false;
        for (int lc = 0; lc < Ku; lc++) {
            if (!bits[lc]) continue;
            named = true;
            if (!active[u.cases[lc]]) continue;
            const int st = (bits[lc] & 2) ? CFD_E_ARG : (bits[lc] & 1) ? CFD_E_NEGATIVE_PRESSURE : CFD_E_HLLC;
            const int r2 = park(lc, st, it);
            if (r2) return r2;
        }
        return named ? CFD_OK : rc0;
    };
    // backward-Euler Newton iterations in lockstep (cfd_solve_steady per case)
    for (int it = 0; !rc && it < o.max_iter; it++) {
        if (it > 0 && u.m->N >= compact_min) {
            // the cases that finished at the last residual check hold their
            // final state (nodes and ghosts) from here on
            int64_t live = 0;
            std::vector<int> take, take_l, rest;
            for (int lc = 0; lc < (int)u.cases.size(); lc++) {
                const int c = u.cases[lc];
                if (active[c]) {
                    live += u.noff[lc + 1] - u.noff[lc];
                    rest.push_back(c);
                } else {
                    take.push_back(c);
                    take_l.push_back(lc);
                }
            }
            if (2 * live <= u.m->N) {
                for (int c : rest) keep[c] = 1;
                rc = harvest(u, take_l, opts, p, elapsed(), results, q_out, keep, q_keep);
                if (rc) break;
                // the compacted union keeps the first union's tuned variants
                // (tuning again at every compaction cost more than it gained)
                const bool tuned = u.m->tuned, use_sell = u.m->use_sell, sync_free = u.m->sync_free;
                const int8_t gl0 = u.m->tune_gl_stage[0], gl1 = u.m->tune_gl_stage[1];
                cfd_mesh_destroy(u.m);
                u.m = nullptr;
                for (int c : take)
                    if (done) done(done_arg, c);
                rc = build_union(ctx, rest, meshes, params, q_keep, allow_graphs, u);
                for (int c : rest) {
                    keep[c] = 0;
                    std::vector<double>().swap(q_keep[c]);
                }
                if (rc) break;
                if (tuned) {
                    u.m->tuned = true;
                    u.m->use_sell = use_sell;
                    u.m->sync_free = sync_free;
                    u.m->tune_gl_stage[0] = gl0;
                    u.m->tune_gl_stage[1] = gl1;
                }
                if (cfd_progress())
                    fprintf(stderr, "cfd_batch_solve: iteration %d, union compacted to %d cases (%d nodes), %.2f s\n",
                            it, (int)rest.size(), u.m->N, elapsed());
            }
        }
        cfd_mesh *m = u.m;
        const int Ku = (int)u.cases.size();
        cfd_params pi = p;
        if (pi.order == 2 && it < o.first_order_iters) pi.order = 1;
        rc = cfd_bc(m, &pi, it);
        if (rc == CFD_E_NEGATIVE_PRESSURE || rc == CFD_E_ARG) rc = isolate(rc, it);
        if (rc) break;
        if (pi.order == 2) {
            rc = pi.limiter != CFD_LIM_NONE ? cfd_gradient_limiter(m, &pi) : cfd_gradient(m, &pi);
            if (rc) break;
        }
        rc = cfd_residual(m, &pi, 1, nullptr);
        if (rc) break;
        double dummy;
        rc = cfd_residual_rms(m, &dummy);  // flag checks of the residual
        if (rc == CFD_E_HLLC) rc = isolate(rc, it);
        if (rc) break;
        hipLaunchKernelGGL(k_rms_seg, dim3(m->seg_rms.nblocks), dim3(kBlock), 0, m->stream, m->seg_rms.blk,
                           m->seg_rms.first, m->seg_rms.off, m->seg_rms.len, m->d.res, m->bpart);
        hipLaunchKernelGGL(k_reduce_seg<5>, dim3(Ku), dim3(kBlock), 0, m->stream, m->seg_rms.first, m->bpart, m->bred);
        if (hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * Ku, hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
            rc = cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: reading residual norms failed");
            break;
        }
        int64_t live_nodes = 0;
        for (int lc = 0; lc < Ku; lc++) {
            const int c = u.cases[lc];
            if (!active[c]) continue;
            const double *sv = m->hbred + (size_t)lc * kBOut;
            const double t = sv[0] + sv[1] + sv[2] + sv[3] + sv[4];
            const double rms = sqrt(t / (5.0 * (double)(u.noff[lc + 1] - u.noff[lc])));
            cfd_solve_result &r = results[c];
            if (it == 0) r.first_rms = rms;
            r.final_rms = rms;
            r.iterations = it;
            if (!(rms == rms) || std::isinf(rms)) {
                rc = park(lc, CFD_E_NONFINITE, it);
                if (rc) break;
            } else if (rms < o.tol) {
                r.converged = 1;
                active[c] = 0;
            }
            if (active[c]) live_nodes += u.noff[lc + 1] - u.noff[lc];
        }
        if (rc || !live_nodes) break;
        const int Kn = Ku;
        rc = cfd_delta_t(m, &pi, it, nullptr);
        if (!rc) rc = solve_jacobian(m, pi, &o, it);
        if (rc) break;
        solve.assign(Kn, 0);
        for (int lc = 0; lc < Kn; lc++) solve[lc] = active[u.cases[lc]];
        if (o.autotune && !m->tuned) {  // the first union mesh tunes once
            cfd_tune_report tr;
            rc = autotune_mesh(m, &pi, &o.lin, &tr,
                               [&] { return run_linsolve_batch(m, &o.lin, solve, iters, fres, lstatus); });
            if (rc) break;
            m->tuned = true;
        }
        rc = run_linsolve_batch(m, &o.lin, solve, iters, fres, lstatus);
        if (rc) break;
        for (int lc = 0; !rc && lc < Kn; lc++)
            if (lstatus[lc] && active[u.cases[lc]]) rc = park(lc, lstatus[lc], it);
        if (rc) break;
        act_i.assign(Kn, 0);
        int n_active = 0, kmax = 0;
        for (int lc = 0; lc < Kn; lc++) {
            const int c = u.cases[lc];
            act_i[lc] = active[c];
            if (active[c]) {
                results[c].krylov_total += iters[lc];
                n_active++;
                kmax = std::max(kmax, iters[lc]);
            }
        }
        if (hipMemcpyAsync(u.d_active, act_i.data(), sizeof(int) * Kn, hipMemcpyHostToDevice, m->stream) !=
            hipSuccess) {
            rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpyAsync failed");
            break;
        }
        hipLaunchKernelGGL(k_update_seg, dim3((m->N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, m->N,
                           o.relaxation, m->order, m->ecase, u.d_active, m->x, m->d.qa);
        if (hipStreamSynchronize(m->stream) != hipSuccess) {  // act_i is a host buffer of the copy above
            rc = cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: update failed");
            break;
        }
        m->q_version++;
        for (int lc = 0; lc < Kn; lc++)
            if (active[u.cases[lc]]) results[u.cases[lc]].iterations = it + 1;
        if (cfd_progress())
            fprintf(stderr, "cfd_batch_solve: iteration %d, %d of %d cases active, krylov %d, %.2f s\n", it,
                    n_active, K, kmax, elapsed());
    }
    // every case of the last union: forces, time and state
    if (!rc && u.m) {
        std::vector<int> take_l(u.cases.size());
        for (size_t lc = 0; lc < u.cases.size(); lc++) take_l[lc] = (int)lc;
        std::fill(keep.begin(), keep.end(), 0);
        rc = harvest(u, take_l, opts, p, elapsed(), results, q_out, keep, q_keep);
    }
    if (u.m) cfd_mesh_destroy(u.m);
    if (rc) {
        for (int c = 0; c < K; c++)
            if (!results[c].status) results[c].status = rc;
        return rc;
    }
    if (done)
        for (int c : u.cases) done(done_arg, c);
    return CFD_OK;
}

extern "C" int cfd_batch_solve(cfd_ctx *ctx, int32_t K, const cfd_mesh_desc *meshes, const cfd_params *params,
                               const cfd_solve_opts *opts, cfd_solve_result *results) {
    return cfd_batch_solve_into(ctx, K, meshes, params, opts, results, nullptr, nullptr, true, nullptr, nullptr);
}
