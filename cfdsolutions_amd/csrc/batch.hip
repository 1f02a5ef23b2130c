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
    CFD_HIP(m, hipMemcpy(sg.blk, blk.data(), sizeof(int2) * blk.size(), hipMemcpyHostToDevice));
    CFD_HIP(m, hipMemcpy(sg.first, sg.h_first.data(), sizeof(int) * (K + 1), hipMemcpyHostToDevice));
    CFD_HIP(m, hipMemcpy(sg.off, off.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    CFD_HIP(m, hipMemcpy(sg.len, len.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    return CFD_OK;
}

}  // namespace

bool cfd_same_scheme(const cfd_params &a, const cfd_params &b) {
    return a.gamma == b.gamma && a.flux == b.flux && a.order == b.order && a.limiter == b.limiter &&
           a.limiter_order == b.limiter_order && a.entropy_fix == b.entropy_fix && a.gradient == b.gradient &&
           a.zpg_iter == b.zpg_iter && a.precision == b.precision && a.venkat_k == b.venkat_k &&
           a.cfl_min == b.cfl_min && a.cfl_max == b.cfl_max && a.cfl_ramp == b.cfl_ramp;
}

bool cfd_same_solve(const cfd_solve_opts &a, const cfd_solve_opts &b) {
    return a.max_iter == b.max_iter && a.tol == b.tol && a.first_order_iters == b.first_order_iters &&
           a.relaxation == b.relaxation && a.wall_tag == b.wall_tag && a.lin.solver == b.lin.solver &&
           a.lin.precond == b.lin.precond && a.lin.krylov_m == b.lin.krylov_m &&
           a.lin.max_restarts == b.lin.max_restarts && a.lin.rtol == b.lin.rtol && a.lin.mixed == b.lin.mixed;
}

int cfd_batch_solve_into(cfd_ctx *ctx, int32_t K, const cfd_mesh_desc *meshes, const cfd_params *params,
                         const cfd_solve_opts *opts, cfd_solve_result *results, double *const *q_out,
                         int64_t *device_bytes, bool allow_graphs) {
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
    // union mesh
    std::vector<int> noff(K + 1, 0), eoff(K + 1, 0), boff(K + 1, 0);
    for (int c = 0; c < K; c++) {
        const cfd_mesh_desc &d = meshes[c];
        if (d.n_node <= 0 || d.n_edge < 0 || d.n_bedge < 0)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: case %d has an invalid mesh", c);
        noff[c + 1] = noff[c] + d.n_node;
        eoff[c + 1] = eoff[c] + d.n_edge;
        boff[c + 1] = boff[c] + d.n_bedge;
        if ((int64_t)noff[c + 1] + boff[c + 1] > INT32_MAX / 2 || 2LL * eoff[c + 1] + boff[c + 1] > INT32_MAX / 2)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: union mesh too large for int32 indices");
    }
    const int N = noff[K], E = eoff[K], B = boff[K];
    std::vector<int> edge(2 * (size_t)E), bedge(2 * (size_t)B), btype(B), btag(B), ia(N + 1, 0), ja(2 * (size_t)E),
        node_case(N), bcase(B);
    std::vector<double> earea(3 * (size_t)E), barea(3 * (size_t)B), xyz(3 * (size_t)N), vol(N);
    for (int c = 0; c < K; c++) {
        const cfd_mesh_desc &d = meshes[c];
        const int no = noff[c], eo = eoff[c], bo = boff[c];
        for (int e = 0; e < d.n_edge; e++) {
            edge[2 * (size_t)(eo + e)] = d.edge[2 * e] + no;
            edge[2 * (size_t)(eo + e) + 1] = d.edge[2 * e + 1] + no;
        }
        memcpy(&earea[3 * (size_t)eo], d.edge_area, sizeof(double) * 3 * d.n_edge);
        for (int b = 0; b < d.n_bedge; b++) {
            if (d.bedge[2 * b + 1] != d.n_node + b)
                return cfd_set_error(ctx, CFD_E_ARG, "cfd_batch_solve: case %d boundary edge %d ghost index", c, b);
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
            ia[no + i + 1] = ia[no] + d.n2n_ia[i + 1];
            node_case[no + i] = c;
        }
        for (int k = 0; k < d.n2n_ia[d.n_node]; k++) ja[(size_t)ia[no] + k] = d.n2n_ja[k] + no;
    }
    cfd_mesh_desc ud{N, E, B, edge.data(), earea.data(), bedge.data(), barea.data(), btype.data(), btag.data(),
                     xyz.data(), vol.data(), ia.data(), ja.data()};
    cfd_mesh *m = nullptr;
    int rc = cfd_mesh_create_cases(ctx, &ud, &params[0], node_case.data(), &m);
    if (rc) return rc;
    if (!allow_graphs) m->use_graphs = false;
    // batch bookkeeping
    m->ncase = K;
    m->case_node = noff;
    m->case_bedge = boff;
    std::vector<double> inf((size_t)K * 8, 0.0);
    for (int c = 0; c < K; c++) {
        const cfd_params &p = params[c];
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
        nb[c] = std::min(grid_for(nc), 1024);                                              // launch_rms
        eoffv[c] = 5 * noff[c];
        elen[c] = 5 * nc;
        nbd[c] = (int)std::max<size_t>(1, std::min<size_t>((5 * (size_t)nc + kBlock - 1) / kBlock, kDotParts));  // mdot
        bo[c] = boff[c];
        blen[c] = bc;
        nbf[c] = bc ? std::min(grid_for(bc), 256) : 1;                                     // launch_forces
    }
    std::vector<char> active(K, 1), solve(K, 1);
    std::vector<int> iters, act_i(K, 1);
    std::vector<double> fres;
    cfd_params p = params[0];
    p.precision = CFD_FP64;
    if (!rc) {
        m->case_inf = (double *)cfd_dalloc(m, sizeof(double) * inf.size());
        m->bcase = (int *)cfd_dalloc(m, sizeof(int) * std::max(1, B));
        m->ecase = (int *)cfd_dalloc(m, sizeof(int) * N);
        m->bred = (double *)cfd_dalloc(m, sizeof(double) * kBOut * K);
        m->bcoef = (double *)cfd_dalloc(m, sizeof(double) * (kMaxKrylov + 1) * K);
        m->bkdim = (int *)cfd_dalloc(m, sizeof(int) * K);
        if (!m->case_inf || !m->bcase || !m->ecase || !m->bred || !m->bcoef || !m->bkdim)
            rc = cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of device memory");
    }
    if (!rc && (hipHostMalloc((void **)&m->hbred, sizeof(double) * kBOut * K, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc((void **)&m->hbcoef, sizeof(double) * (kMaxKrylov + 1) * K, hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc((void **)&m->hbkdim, sizeof(int) * K, hipHostMallocDefault) != hipSuccess))
        rc = cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of pinned host memory");
    if (!rc && hipMemcpy(m->case_inf, inf.data(), sizeof(double) * inf.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpy failed");
    if (!rc && B && hipMemcpy(m->bcase, bcase.data(), sizeof(int) * B, hipMemcpyHostToDevice) != hipSuccess)
        rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpy failed");
    if (!rc && hipMemcpy(m->ecase, node_case.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess)
        rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpy failed");
    if (!rc) rc = build_segmap(m, m->seg_rms, zoff, zlen, nb);
    if (!rc) rc = build_segmap(m, m->seg_dot, eoffv, elen, nbd);
    if (!rc) rc = build_segmap(m, m->seg_forces, bo, blen, nbf);
    if (!rc) {
        const size_t np = std::max({(size_t)m->seg_dot.nblocks * (kMaxKrylov + 1), (size_t)m->seg_rms.nblocks * 5,
                                    (size_t)m->seg_forces.nblocks * 4});
        m->bpart = (double *)cfd_dalloc(m, sizeof(double) * np);
        if (!m->bpart) rc = cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of device memory");
    }
    // initial state: each case at its own freestream (nodes and ghosts)
    if (!rc) {
        std::vector<double> q(5 * (size_t)(N + B));
        auto put = [&](size_t i, int c) {
            const cfd_params &pc = params[c];
            double et = pc.inf_u * pc.inf_u + pc.inf_v * pc.inf_v + pc.inf_w * pc.inf_w;
            et = pc.inf_p / ((pc.gamma - 1.0) * pc.inf_rho) + 0.5 * et;
            const double v[5] = {pc.inf_rho, pc.inf_rho * pc.inf_u, pc.inf_rho * pc.inf_v, pc.inf_rho * pc.inf_w,
                                 pc.inf_rho * et};
            for (int k = 0; k < 5; k++) q[(size_t)k * (N + B) + i] = v[k];
        };
        for (int i = 0; i < N; i++) put(i, node_case[i]);
        for (int b = 0; b < B; b++) put((size_t)N + b, bcase[b]);
        rc = cfd_field_set(m, CFD_FIELD_Q, q.data());
    }
    for (int c = 0; c < K; c++) results[c] = cfd_solve_result{};
    // cases still iterating: the update and the boundary condition skip the
    // others, so a finished case keeps its last iteration's state and ghosts
    int *d_active = nullptr;
    if (!rc) {
        d_active = (int *)cfd_dalloc(m, sizeof(int) * K);
        if (!d_active) rc = cfd_set_error(ctx, CFD_E_OOM, "cfd_batch_solve: out of device memory");
        else if (hipMemcpy(d_active, act_i.data(), sizeof(int) * K, hipMemcpyHostToDevice) != hipSuccess)
            rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpy failed");
        m->case_active = d_active;
    }
    // backward-Euler Newton iterations in lockstep (cfd_solve_steady per case)
    for (int it = 0; !rc && it < o.max_iter; it++) {
        cfd_params pi = p;
        if (pi.order == 2 && it < o.first_order_iters) pi.order = 1;
        rc = cfd_bc(m, &pi, it);
        if (rc) break;
        if (pi.order == 2) {
            rc = pi.limiter != CFD_LIM_NONE ? cfd_gradient_limiter(m, &pi) : cfd_gradient(m, &pi);
            if (rc) break;
        }
        rc = cfd_residual(m, &pi, 1, nullptr);
        if (rc) break;
        double dummy;
        rc = cfd_residual_rms(m, &dummy);  // flag checks of the residual
        if (rc) break;
        hipLaunchKernelGGL(k_rms_seg, dim3(m->seg_rms.nblocks), dim3(kBlock), 0, m->stream, m->seg_rms.blk,
                           m->seg_rms.first, m->seg_rms.off, m->seg_rms.len, m->d.res, m->bpart);
        hipLaunchKernelGGL(k_reduce_seg<5>, dim3(K), dim3(kBlock), 0, m->stream, m->seg_rms.first, m->bpart, m->bred);
        if (hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * K, hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
            rc = cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: reading residual norms failed");
            break;
        }
        bool any = false;
        for (int c = 0; c < K; c++) {
            if (!active[c]) continue;
            const double *sv = m->hbred + (size_t)c * kBOut;
            const double t = sv[0] + sv[1] + sv[2] + sv[3] + sv[4];
            const double rms = sqrt(t / (5.0 * (double)(noff[c + 1] - noff[c])));
            cfd_solve_result &r = results[c];
            if (it == 0) r.first_rms = rms;
            r.final_rms = rms;
            r.iterations = it;
            if (!(rms == rms) || std::isinf(rms)) {
                r.status = CFD_E_NONFINITE;
                active[c] = 0;
            } else if (rms < o.tol) {
                r.converged = 1;
                active[c] = 0;
            }
            any = any || active[c];
        }
        if (!any) break;
        rc = cfd_delta_t(m, &pi, it, nullptr);
        if (!rc) rc = cfd_jacobian(m, &pi, 1, o.lin.mixed);
        if (rc) break;
        for (int c = 0; c < K; c++) solve[c] = active[c];
        rc = run_linsolve_batch(m, &o.lin, solve, iters, fres);
        if (rc) break;
        for (int c = 0; c < K; c++) {
            act_i[c] = active[c];
            if (active[c]) results[c].krylov_total += iters[c];
        }
        if (hipMemcpyAsync(d_active, act_i.data(), sizeof(int) * K, hipMemcpyHostToDevice, m->stream) != hipSuccess) {
            rc = cfd_set_error(ctx, CFD_E_HIP, "hipMemcpyAsync failed");
            break;
        }
        hipLaunchKernelGGL(k_update_seg, dim3((N + kBlock - 1) / kBlock), dim3(kBlock), 0, m->stream, N,
                           o.relaxation, m->order, m->ecase, d_active, m->x, m->d.qa);
        if (hipStreamSynchronize(m->stream) != hipSuccess) {  // act_i is a host buffer of the copy above
            rc = cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: update failed");
            break;
        }
        m->q_version++;
        for (int c = 0; c < K; c++)
            if (active[c]) results[c].iterations = it + 1;
    }
    // forces of every case
    if (!rc) {
        hipLaunchKernelGGL(k_forces_seg, dim3(m->seg_forces.nblocks), dim3(kBlock), 0, m->stream, m->seg_forces.blk,
                           m->seg_forces.first, m->seg_forces.off, m->seg_forces.len, m->bnode, m->btype, m->btag,
                           o.wall_tag, m->d.sa, m->d.qa, p.gamma, m->bpart);
        hipLaunchKernelGGL(k_reduce_seg<4>, dim3(K), dim3(kBlock), 0, m->stream, m->seg_forces.first, m->bpart,
                           m->bred);
        if (hipMemcpyAsync(m->hbred, m->bred, sizeof(double) * kBOut * K, hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess)
            rc = cfd_set_error(ctx, CFD_E_HIP, "cfd_batch_solve: reading forces failed");
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int c = 0; c < K && !rc; c++) {
        cfd_solve_result &r = results[c];
        const double *v = m->hbred + (size_t)c * kBOut;
        const double a = opts[c].alpha_deg * M_PI / 180.0;
        r.force[0] = v[0];
        r.force[1] = v[1];
        r.force[2] = v[2];
        r.drag = v[0] * cos(a) + v[1] * sin(a);
        r.lift = -v[0] * sin(a) + v[1] * cos(a);
        r.pressure_integral = v[3];
        r.seconds = secs;
    }
    // each case's final state, [5][n_c + b_c] with its ghosts after its nodes
    if (!rc && q_out) {
        const size_t NT = (size_t)N + B;
        std::vector<double> u(5 * NT);
        rc = cfd_field_get(m, CFD_FIELD_Q, u.data());
        for (int c = 0; c < K && !rc; c++) {
            if (!q_out[c]) continue;
            const size_t nc = (size_t)(noff[c + 1] - noff[c]), bc = (size_t)(boff[c + 1] - boff[c]), tc = nc + bc;
            for (int k = 0; k < 5; k++) {
                const double *src = u.data() + k * NT;
                memcpy(q_out[c] + k * tc, src + noff[c], sizeof(double) * nc);
                memcpy(q_out[c] + k * tc + nc, src + N + boff[c], sizeof(double) * bc);
            }
        }
    }
    if (device_bytes) *device_bytes = m->bytes;
    if (rc)
        for (int c = 0; c < K; c++)
            if (!results[c].status) results[c].status = rc;
    cfd_mesh_destroy(m);
    return rc;
}

extern "C" int cfd_batch_solve(cfd_ctx *ctx, int32_t K, const cfd_mesh_desc *meshes, const cfd_params *params,
                               const cfd_solve_opts *opts, cfd_solve_result *results) {
    return cfd_batch_solve_into(ctx, K, meshes, params, opts, results, nullptr, nullptr, true);
}
