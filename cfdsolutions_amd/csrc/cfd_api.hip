// C ABI of libcfd_hip.so: context and mesh handles, host-side structure
// builders (incidence lists, BSR pattern, level schedules), field transfer,
// and the steady implicit driver.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include <atomic>

#include "cfd_internal.h"
#include "setup_levels.h"

// ---------------------------------------------------------------------------
// errors and allocation
// ---------------------------------------------------------------------------
// The message of the last failure is kept per calling thread (meshes of one
// context may be driven from several threads) and, under the context's lock,
// on the context.
static thread_local std::string t_last_error;

int cfd_set_error(cfd_ctx *ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    t_last_error = buf;
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->lock);
        ctx->err = buf;
    }
    return code;
}

int cfd_hip_check(cfd_ctx *ctx, hipError_t e, const char *what) {
    return cfd_set_error(ctx, e == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

std::shared_mutex &cfd_capture_mutex() {
    static std::shared_mutex mu;
    return mu;
}

hipError_t cfd_memcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    CfdRuntimeGuard g;
    return hipMemcpy(dst, src, bytes, kind);
}

hipError_t cfd_host_alloc(void **p, size_t bytes) {
    CfdRuntimeGuard g;
    return hipHostMalloc(p, bytes, hipHostMallocDefault);
}

void *cfd_dalloc(cfd_mesh *m, size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 8;
    CfdRuntimeGuard g;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    m->allocs.emplace_back(p, bytes);
    m->bytes += (int64_t)bytes;
    return p;
}

void cfd_dfree(cfd_mesh *m, void *p) {
    if (!p) return;
    for (size_t k = 0; k < m->allocs.size(); k++) {
        if (m->allocs[k].first != p) continue;
        m->bytes -= (int64_t)m->allocs[k].second;
        m->allocs[k] = m->allocs.back();
        m->allocs.pop_back();
        CfdRuntimeGuard g;
        hipFree(p);
        return;
    }
}

extern "C" int cfd_ctx_create(int device, cfd_ctx **out) {
    if (!out) return CFD_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return CFD_E_HIP;
    cfd_ctx *c = new (std::nothrow) cfd_ctx();
    if (!c) return CFD_E_OOM;
    c->device = device;
    // this GPU's copy of the allocation-free solve's scratch array; a failed
    // lookup only disables cfd_linsolve_gs_noalloc
    if (hipSetDevice(device) == hipSuccess) {
        hipError_t e = hipSuccess;
        c->noalloc_scratch = noalloc_scratch_address(&e);
        if (!c->noalloc_scratch)
            cfd_set_error(c, CFD_E_HIP, "scratch array of cfd_linsolve_gs_noalloc: %s", hipGetErrorString(e));
    }
    *out = c;
    return CFD_OK;
}

extern "C" void cfd_ctx_destroy(cfd_ctx *ctx) { delete ctx; }

extern "C" const char *cfd_last_error(const cfd_ctx *ctx) {
    if (!ctx) return "null context";
    if (!t_last_error.empty()) return t_last_error.c_str();
    std::lock_guard<std::mutex> g(const_cast<cfd_ctx *>(ctx)->lock);
    t_last_error = ctx->err;
    return t_last_error.c_str();
}

// ---------------------------------------------------------------------------
// mesh creation
// ---------------------------------------------------------------------------

// Pattern of a block-SELL layout (BlockSell) over rows 0..n-1: row i holds
// the BSR slots rb(i) .. re(i)-1 in that order, with block columns ja[slot].
template <class RB, class RE>
static void build_block_sell(int n, RB rb, RE re, const int *ja, std::vector<int> &off, HostArray<int> &col,
                             HostArray<int> &src) {
    const int ns = (n + 63) / 64;
    off.assign(ns + 1, 0);
    for (int sl = 0; sl < ns; sl++) {
        int w = 0;
        for (int i = 64 * sl; i < std::min(n, 64 * sl + 64); i++) w = std::max(w, re(i) - rb(i));
        off[sl + 1] = off[sl] + 64 * w;
    }
    col = HostArray<int>((size_t)off[ns]);
    src = HostArray<int>((size_t)off[ns]);
    // every slot of a slice written by the thread that owns the slice; padding
    // slots get column 0 and source -1
    cfd_parallel_for(ns, [&](int64_t s0, int64_t s1) {
        for (int sl = (int)s0; sl < (int)s1; sl++) {
            const int w = (off[sl + 1] - off[sl]) / 64;
            for (int l = 0; l < 64; l++) {
                const int i = 64 * sl + l;
                const int len = i < n ? re(i) - rb(i) : 0;
                for (int k = 0; k < w; k++) {
                    const size_t q = (size_t)off[sl] + 64 * (size_t)k + l;
                    col[q] = k < len ? ja[rb(i) + k] : 0;
                    src[q] = k < len ? rb(i) + k : -1;
                }
            }
        }
    }, 1 << 8);
}

static int upload_block_sell(cfd_mesh *m, const std::vector<int> &off, const HostArray<int> &col,
                             const HostArray<int> &src, BlockSell &bs);

// Rows of each level followed by -1 up to a multiple of w.
static void pad_levels(const std::vector<int> &ptr, const std::vector<int> &rows, int w, std::vector<int> &out) {
    out.clear();
    for (size_t l = 0; l + 1 < ptr.size(); l++) {
        for (int k = ptr[l]; k < ptr[l + 1]; k++) out.push_back(rows[k]);
        while (out.size() % w) out.push_back(-1);
    }
}

static void build_levels(const std::vector<int> &ia, const std::vector<int> &ja, const std::vector<int> &iau, int N,
                         bool upper, std::vector<int> &ptr, std::vector<int> &rows, int &maxw) {
    std::vector<int> lev(N, 0);
    int nl = 0;
    if (!upper) {
        for (int i = 0; i < N; i++) {
            int l = 0;
            for (int k = ia[i]; k < iau[i]; k++) l = std::max(l, lev[ja[k]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    } else {
        for (int i = N - 1; i >= 0; i--) {
            int l = 0;
            for (int k = iau[i] + 1; k < ia[i + 1]; k++) l = std::max(l, lev[ja[k]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
    }
    ptr.assign(nl + 1, 0);
    for (int i = 0; i < N; i++) ptr[lev[i] + 1]++;
    maxw = 0;
    for (int l = 0; l < nl; l++) {
        maxw = std::max(maxw, ptr[l + 1]);
        ptr[l + 1] += ptr[l];
    }
    rows.assign(N, 0);
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < N; i++) rows[pos[lev[i]]++] = i;
}

template <class T>
static T *upload(cfd_mesh *m, const T *h, size_t n, int *rc) {
    T *d = (T *)cfd_dalloc(m, sizeof(T) * n);
    if (!d) {
        *rc = cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory");
        return nullptr;
    }
    if (n && cfd_memcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) {
        *rc = cfd_set_error(m->ctx, CFD_E_HIP, "hipMemcpy failed");
        return nullptr;
    }
    return d;
}

static int upload_block_sell(cfd_mesh *m, const std::vector<int> &off, const HostArray<int> &col,
                             const HostArray<int> &src, BlockSell &bs) {
    int rc = 0;
    bs.nblk = (int64_t)col.size();
    bs.off = upload(m, off.data(), off.size(), &rc);
    if (!rc && col.size()) bs.col = upload(m, col.data(), col.size(), &rc);
    if (!rc && src.size()) bs.src = upload(m, src.data(), src.size(), &rc);
    return rc;
}

extern "C" void cfd_mesh_destroy(cfd_mesh *m) {
    if (!m) return;
    hipSetDevice(m->ctx->device);
    if (m->stream) hipStreamSynchronize(m->stream);
    drop_krylov_graphs(m);
    CfdRuntimeGuard g;
    for (auto &a : m->allocs) hipFree(a.first);
    if (m->hred) hipHostFree(m->hred);
    if (m->hflag) hipHostFree(m->hflag);
    if (m->hbred) hipHostFree(m->hbred);
    if (m->hbcoef) hipHostFree(m->hbcoef);
    if (m->hbkdim) hipHostFree(m->hbkdim);
    if (m->stream) hipStreamDestroy(m->stream);
    delete m;
}

static void freestream_q(const cfd_params *p, double *q5) {
    // Solver_Set_Initial_Conditions + ComputeFarFieldCondition with
    // Get_TotalEnergy (Euler3D/src/Solver.cpp:340-362, Material.cpp:185-210)
    double et = p->inf_u * p->inf_u + p->inf_v * p->inf_v + p->inf_w * p->inf_w;
    et = p->inf_p / ((p->gamma - 1.0) * p->inf_rho) + 0.5 * et;
    q5[0] = p->inf_rho;
    q5[1] = p->inf_rho * p->inf_u;
    q5[2] = p->inf_rho * p->inf_v;
    q5[3] = p->inf_rho * p->inf_w;
    q5[4] = p->inf_rho * et;
}

extern "C" int cfd_mesh_create(cfd_ctx *ctx, const cfd_mesh_desc *d, const cfd_params *p, cfd_mesh **out) {
    return cfd_mesh_create_cases(ctx, d, p, nullptr, out);
}

// node_case (may be NULL): case index of every node of a union of
// independent meshes; the internal numbering then keeps each case's nodes
// contiguous (case-major, Morton order within a case).
int cfd_mesh_create_cases(cfd_ctx *ctx, const cfd_mesh_desc *d, const cfd_params *p, const int *node_case,
                          cfd_mesh **out) {
    if (!ctx || !d || !p || !out) return CFD_E_ARG;
    const int N = d->n_node, E = d->n_edge, B = d->n_bedge;
    if (N <= 0 || E < 0 || B < 0 || !d->edge || !d->edge_area || (B && (!d->bedge || !d->bedge_area || !d->bedge_type)) ||
        !d->xyz || !d->vol || !d->n2n_ia || !d->n2n_ja)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: invalid descriptor");
    if ((int64_t)N + B > INT32_MAX / 2 || 2LL * E + B > INT32_MAX / 2)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: mesh too large for int32 indices");
    // validate indices on the host before anything reaches a kernel
    for (int e = 0; e < E; e++) {
        const int l = d->edge[2 * e], r = d->edge[2 * e + 1];
        if (l < 0 || r < 0 || l >= N || r >= N || l >= r)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: edge %d has invalid nodes (%d, %d)", e, l, r);
    }
    for (int b = 0; b < B; b++) {
        const int l = d->bedge[2 * b], g = d->bedge[2 * b + 1];
        if (l < 0 || l >= N || g != N + b)
            return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: boundary edge %d must pair a node with ghost %d", b,
                                 N + b);
    }
    if (d->n2n_ia[0] != 0) return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: n2n_ia[0] != 0");
    for (int i = 0; i < N; i++) {
        if (d->n2n_ia[i + 1] < d->n2n_ia[i]) return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: bad n2n_ia");
        for (int k = d->n2n_ia[i]; k < d->n2n_ia[i + 1]; k++)
            if (d->n2n_ja[k] < 0 || d->n2n_ja[k] >= N || d->n2n_ja[k] == i)
                return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: bad neighbour in n2n_ja");
    }
    if (d->n2n_ia[N] != 2 * E)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: n2n lists (%d entries) do not match 2E = %d",
                             d->n2n_ia[N], 2 * E);
    if (hipSetDevice(ctx->device) != hipSuccess) return cfd_set_error(ctx, CFD_E_HIP, "hipSetDevice failed");

    cfd_mesh *m = new (std::nothrow) cfd_mesh();
    if (!m) return CFD_E_OOM;
    m->ctx = ctx;
    m->N = N;
    m->E = E;
    m->B = B;
    m->NT = N + B;
    int rc = CFD_OK;
    auto tph = std::chrono::steady_clock::now();
    auto phase = [&](const char *what) {
        if (!cfd_progress()) return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "cfd_mesh_create: %-28s %8.3f s\n", what, std::chrono::duration<double>(t - tph).count());
        tph = t;
    };
    try {
        hipError_t se;
        {
            CfdRuntimeGuard g;
            se = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
        }
        if (se != hipSuccess) {
            delete m;
            return cfd_set_error(ctx, CFD_E_HIP, "hipStreamCreate failed");
        }
        // Internal numbering of the node data: Morton order of the coordinates,
        // so the nodes a kernel has in flight on one XCD, and their neighbours,
        // occupy a compact range of memory that stays in that XCD's L2.
        // order[ii] = reference (API) index of internal node ii; ghosts keep
        // their index N + b.  The BSR system, its solvers and the [N][5]
        // vectors stay in the reference ordering.
        std::vector<int> order(N), iperm(N);
        {
            // bounding box per case, so a case of a union is numbered exactly
            // as the same mesh on its own
            int ncs = 1;
            if (node_case)
                for (int i = 0; i < N; i++) ncs = std::max(ncs, node_case[i] + 1);
            std::vector<double> blo(3 * (size_t)ncs, 1e300), bhi(3 * (size_t)ncs, -1e300);
            for (int i = 0; i < N; i++) {
                const size_t cb = 3 * (size_t)(node_case ? node_case[i] : 0);
                for (int k = 0; k < 3; k++) {
                    blo[cb + k] = std::min(blo[cb + k], d->xyz[3 * (size_t)i + k]);
                    bhi[cb + k] = std::max(bhi[cb + k], d->xyz[3 * (size_t)i + k]);
                }
            }
            struct Key {
                int cs;
                uint64_t morton;
                int node;
                bool operator<(const Key &o) const {
                    return cs != o.cs ? cs < o.cs : (morton != o.morton ? morton < o.morton : node < o.node);
                }
            };
            std::vector<Key> key(N);
            cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
                for (int i = (int)i0; i < (int)i1; i++) {
                    uint64_t c = 0;
                    uint32_t q3[3];
                    const double *lo = &blo[3 * (size_t)(node_case ? node_case[i] : 0)];
                    const double *hi = &bhi[3 * (size_t)(node_case ? node_case[i] : 0)];
                    for (int k = 0; k < 3; k++) {
                        const double span = hi[k] - lo[k];
                        const double f = span > 0 ? (d->xyz[3 * (size_t)i + k] - lo[k]) / span : 0.0;
                        q3[k] = (uint32_t)std::min(2097151.0, std::max(0.0, f * 2097151.0));
                    }
                    for (int bit = 20; bit >= 0; bit--)
                        for (int k = 0; k < 3; k++) c = (c << 1) | ((q3[k] >> bit) & 1u);
                    key[i] = {node_case ? node_case[i] : 0, c, i};
                }
            });
            // the keys order by case first: when the cases occupy ascending
            // node ranges (unions), each range sorts on its own
            std::vector<int> cut{0};
            bool ranges = node_case != nullptr;
            for (int i = 1; ranges && i < N; i++) {
                if (node_case[i] < node_case[i - 1]) ranges = false;
                else if (node_case[i] != node_case[i - 1]) cut.push_back(i);
            }
            cut.push_back(N);
            if (ranges && cut.size() > 2) {
                cfd_parallel_for(
                    (int64_t)cut.size() - 1,
                    [&](int64_t a, int64_t b) {
                        for (int64_t r = a; r < b; r++) std::sort(key.begin() + cut[r], key.begin() + cut[r + 1]);
                    },
                    1);
            } else {
                std::sort(key.begin(), key.end());
            }
            for (int t = 0; t < N; t++) {
                order[t] = key[t].node;
                iperm[key[t].node] = t;
            }
        }
        phase("morton order");
        m->h_order = order;
        // node incidence lists (internal numbering) in edge order: interior
        // edges by index, then boundary edges -- the order in which the
        // reference's sequential loops deliver each node's contributions
        // (a node's interior edges are its n2n neighbours, counted once each)
        std::vector<int> ip(N + 1, 0);
        cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
            for (int i = (int)i0; i < (int)i1; i++) ip[iperm[i] + 1] = d->n2n_ia[i + 1] - d->n2n_ia[i];
        });
        for (int b = 0; b < B; b++) ip[iperm[d->bedge[2 * b]] + 1]++;
        for (int i = 0; i < N; i++) ip[i + 1] += ip[i];
        HostArray<int2> inc(ip[N]);  // every entry written below (checked by the list ends)
        {
            // Edges and boundary edges in order, each appended to the lists of
            // its nodes.  A union's cases hold ascending ranges of nodes, edges
            // and boundary edges with no edge between cases, so each case's
            // ranges are filled on their own (in parallel) in the same order.
            std::vector<int> pos(ip.begin(), ip.end() - 1), ecut{0}, bcut{0};
            bool ranges = node_case != nullptr;
            auto ecase = [&](int e) { return node_case[d->edge[2 * e]]; };
            auto bcase = [&](int b) { return node_case[d->bedge[2 * b]]; };
            for (int e = 0; ranges && e < E; e++) {
                if (node_case[d->edge[2 * e + 1]] != ecase(e) || (e && ecase(e) < ecase(e - 1))) ranges = false;
                else if (e && ecase(e) != ecase(e - 1)) ecut.push_back(e);
            }
            for (int b = 1; ranges && b < B; b++) {
                if (bcase(b) < bcase(b - 1)) ranges = false;
                else if (bcase(b) != bcase(b - 1)) bcut.push_back(b);
            }
            // piece r: edges [ecut[r], ecut[r+1]) and boundary edges [bcut[r], bcut[r+1]) of one case
            const int ncs = node_case ? node_case[N - 1] + 1 : 1;
            if (ranges && E && B && (int)ecut.size() == ncs && (int)bcut.size() == ncs && ecase(0) == 0 &&
                bcase(0) == 0) {
                ecut.push_back(E);
                bcut.push_back(B);
            } else {
                ecut.assign({0, E});
                bcut.assign({0, B});
            }
            cfd_parallel_for(
                (int64_t)ecut.size() - 1,
                [&](int64_t r0, int64_t r1) {
                    for (int64_t r = r0; r < r1; r++) {
                        for (int e = ecut[r]; e < ecut[r + 1]; e++) {
                            // .y = e for the L end of the edge, ~e for the R end (roles follow
                            // the reference numbering, L < R, not the internal one)
                            const int l = iperm[d->edge[2 * e]], rr = iperm[d->edge[2 * e + 1]];
                            inc[pos[l]++] = make_int2(rr, e);
                            inc[pos[rr]++] = make_int2(l, ~e);
                        }
                        for (int b = bcut[r]; b < bcut[r + 1]; b++)
                            inc[pos[iperm[d->bedge[2 * b]]]++] = make_int2(N + b, E + b);
                    }
                },
                1);
            for (int i = 0; i < N; i++)
                if (pos[i] != ip[i + 1]) {
                    cfd_mesh_destroy(m);
                    return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: n2n lists do not match the edges at node %d",
                                         order[i]);
                }
        }
        phase("incidence lists");
        // SELL-64 copy of the incidence lists with the area vector of each
        // incidence (see cfd_mesh::nbr); padding entries are never read.
        const int nsl = (N + kSlice - 1) / kSlice;
        std::vector<int> sptr(nsl + 1, 0);
        {
            int64_t tot = 0;
            for (int sl = 0; sl < nsl; sl++) {
                int w = 0;
                for (int i = sl * kSlice; i < std::min(N, (sl + 1) * kSlice); i++) w = std::max(w, ip[i + 1] - ip[i]);
                tot += (int64_t)kSlice * w;
                if (tot > INT32_MAX) {
                    cfd_mesh_destroy(m);
                    return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: mesh too large for 32-bit incidence offsets");
                }
                sptr[sl + 1] = (int)tot;
            }
        }
        m->nsell = (size_t)sptr[nsl];
        // uninitialised storage: the per-slice pass below writes every entry
        // once (padding included), in parallel, instead of a serial zero fill
        HostArray<int> nbr(m->nsell);
        HostArray<double> isx(m->nsell), isy(m->nsell), isz(m->nsell);
        m->h_sell_edge = HostArray<int>(m->nsell);  // every entry written below
        cfd_parallel_for(
            nsl,
            [&](int64_t s0, int64_t s1) {
                for (int sl = (int)s0; sl < (int)s1; sl++) {
                    const int w = (sptr[sl + 1] - sptr[sl]) / kSlice;
                    for (int l = 0; l < kSlice; l++) {
                        const int i = sl * kSlice + l;
                        const int cnt = i < N ? ip[i + 1] - ip[i] : 0;
                        for (int k = 0; k < w; k++) {
                            const size_t idx = (size_t)sptr[sl] + (size_t)k * kSlice + l;
                            if (k >= cnt) {  // padding, never read
                                nbr[idx] = 0;
                                isx[idx] = isy[idx] = isz[idx] = 0.0;
                                m->h_sell_edge[idx] = -1;
                                continue;
                            }
                            const int2 en = inc[ip[i] + k];
                            m->h_sell_edge[idx] = en.y >= 0 ? en.y : ~en.y;  // boundary: E + b
                            const double *S;
                            if (en.x >= N) {
                                nbr[idx] = en.x;
                                S = d->bedge_area + 3 * (size_t)(en.x - N);
                            } else if (en.y >= 0) {
                                nbr[idx] = en.x;
                                S = d->edge_area + 3 * (size_t)en.y;
                            } else {
                                nbr[idx] = ~en.x;
                                S = d->edge_area + 3 * (size_t)(~en.y);
                            }
                            isx[idx] = S[0];
                            isy[idx] = S[1];
                            isz[idx] = S[2];
                        }
                    }
                }
            },
            64);
        phase("sell incidences");
        // BSR pattern, Create_CRS_SolverBlockMatrix (Euler3D/src/Solver.cpp:612-675):
        // row i holds i and its neighbours in ascending column order.
        m->h_ia.assign(N + 1, 0);
        for (int i = 0; i < N; i++) m->h_ia[i + 1] = m->h_ia[i] + (d->n2n_ia[i + 1] - d->n2n_ia[i]) + 1;
        m->nnz = m->h_ia[N];
        m->h_ja.resize(m->nnz);
        m->h_iau.resize(N);
        cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
            for (int i = (int)i0; i < (int)i1; i++) {
                int k = m->h_ia[i];
                m->h_ja[k++] = i;
                for (int q = d->n2n_ia[i]; q < d->n2n_ia[i + 1]; q++) m->h_ja[k++] = d->n2n_ja[q];
                std::sort(m->h_ja.begin() + m->h_ia[i], m->h_ja.begin() + m->h_ia[i + 1]);
                for (int q = m->h_ia[i]; q < m->h_ia[i + 1]; q++)
                    if (m->h_ja[q] == i) m->h_iau[i] = q;
            }
        });
        phase("bsr pattern");
        // BSR block of the off-diagonal (row of the node, column of the neighbour)
        HostArray<int> off(inc.size());  // every entry written below (unless a neighbour is missing)
        std::vector<int> missing(N, -1);  // per node: a neighbour missing from its n2n list
        cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
            for (int ii = (int)i0; ii < (int)i1; ii++) {
                const int i = order[ii];
                for (int k = ip[ii]; k < ip[ii + 1]; k++) {
                    if (inc[k].x >= N) {
                        off[k] = -1;
                        continue;
                    }
                    const int j = order[inc[k].x];
                    auto b0 = m->h_ja.begin() + m->h_ia[i], b1 = m->h_ja.begin() + m->h_ia[i + 1];
                    auto it = std::lower_bound(b0, b1, j);
                    if (it == b1 || *it != j) {
                        missing[ii] = j;
                        break;
                    }
                    off[k] = (int)(it - m->h_ja.begin());
                }
            }
        });
        for (int ii = 0; ii < N; ii++)
            if (missing[ii] >= 0) {
                const int i = order[ii], j = missing[ii];
                cfd_mesh_destroy(m);
                return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: edge (%d,%d) missing from n2n lists", i, j);
            }
        HostArray<int> soff(m->nsell);
        cfd_parallel_for(
            nsl,
            [&](int64_t s0, int64_t s1) {
                for (int sl = (int)s0; sl < (int)s1; sl++) {
                    const int w = (sptr[sl + 1] - sptr[sl]) / kSlice;
                    for (int l = 0; l < kSlice; l++) {
                        const int i = sl * kSlice + l;
                        const int cnt = i < N ? ip[i + 1] - ip[i] : 0;
                        for (int k = 0; k < w; k++)
                            soff[(size_t)sptr[sl] + (size_t)k * kSlice + l] = k < cnt ? off[ip[i] + k] : -1;
                    }
                }
            },
            64);
        phase("off-diagonal slots");
        // (the natural-order ILU(0) merge map is built on first use,
        // cfd_ensure_ilu_map)
        std::vector<int> lrows, urows;
        build_levels(m->h_ia, m->h_ja, m->h_iau, N, false, m->lower.ptr, lrows, m->lower.maxwidth);
        build_levels(m->h_ia, m->h_ja, m->h_iau, N, true, m->upper.ptr, urows, m->upper.maxwidth);
        m->lower.nlevels = (int)m->lower.ptr.size() - 1;
        m->upper.nlevels = (int)m->upper.ptr.size() - 1;
        phase("levels");
        // multicolour ordering for ILU(0): greedy colouring of the node graph in
        // reference order; rows sorted by (colour, index)
        std::vector<int> mperm(N), mia(N + 1, 0), miau(N), mlrows, murows;
        HostArray<int> mja(m->nnz), msrc(m->nnz);  // filled in full below
        std::unique_ptr<int[]> mmoff, mmqq;
        int64_t mc_nmap = 0;
        {
            // Greedy colouring in reference order.  The cases of a union hold
            // ascending node ranges with no edge between them, so each range
            // is coloured on its own (in parallel), exactly as in one pass.
            std::vector<int> colour(N, -1), cut{0};
            bool ranges = node_case != nullptr;
            for (int i = 1; ranges && i < N; i++) {
                if (node_case[i] < node_case[i - 1]) ranges = false;
                else if (node_case[i] != node_case[i - 1]) cut.push_back(i);
            }
            if (!ranges) cut.assign(1, 0);
            cut.push_back(N);
            if (cut.size() > 2) {  // no edge may leave its range (checked range by range, in parallel)
                std::atomic<bool> leaves{false};
                cfd_parallel_for(
                    (int64_t)cut.size() - 1,
                    [&](int64_t r0, int64_t r1) {
                        for (int64_t r = r0; r < r1 && !leaves.load(std::memory_order_relaxed); r++)
                            for (int q = d->n2n_ia[cut[r]]; q < d->n2n_ia[cut[r + 1]]; q++)
                                if (d->n2n_ja[q] < cut[r] || d->n2n_ja[q] >= cut[r + 1]) {
                                    leaves = true;
                                    break;
                                }
                    },
                    1);
                if (leaves) cut.assign({0, N});
            }
            phase("multicolour: ranges");
            std::vector<int> ncs(cut.size() - 1, 0);
            cfd_parallel_for(
                (int64_t)cut.size() - 1,
                [&](int64_t r0, int64_t r1) {
                    std::vector<int> mark;
                    for (int64_t r = r0; r < r1; r++) {
                        int nc = 0;
                        for (int i = cut[r]; i < cut[r + 1]; i++) {
                            mark.assign(nc + 1, 0);
                            for (int q = d->n2n_ia[i]; q < d->n2n_ia[i + 1]; q++) {
                                const int cj = colour[d->n2n_ja[q]];
                                if (cj >= 0) mark[cj] = 1;
                            }
                            int c = 0;
                            while (mark[c]) c++;
                            colour[i] = c;
                            nc = std::max(nc, c + 1);
                        }
                        ncs[r] = nc;
                    }
                },
                1);
            phase("multicolour: colouring");
            int nc = 1;
            for (int v : ncs) nc = std::max(nc, v);
            m->ncolor = nc;
            std::vector<int> cnt(nc + 1, 0), miperm(N);
            for (int i = 0; i < N; i++) cnt[colour[i] + 1]++;
            for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
            for (int i = 0; i < N; i++) mperm[cnt[colour[i]]++] = i;  // ascending index within a colour
            m->mc_cstart.assign(nc + 1, 0);
            for (int c = 0; c < nc; c++) m->mc_cstart[c + 1] = cnt[c];  // cnt[c] now ends colour c
            for (int t = 0; t < N; t++) miperm[mperm[t]] = t;
            for (int t = 0; t < N; t++) mia[t + 1] = mia[t] + (m->h_ia[mperm[t] + 1] - m->h_ia[mperm[t]]);
            cfd_parallel_for(N, [&](int64_t t0, int64_t t1) {
                std::vector<std::pair<int, int>> row;
                for (int t = (int)t0; t < (int)t1; t++) {
                    const int oi = mperm[t];
                    row.clear();
                    for (int k = m->h_ia[oi]; k < m->h_ia[oi + 1]; k++) row.push_back({miperm[m->h_ja[k]], k});
                    std::sort(row.begin(), row.end());
                    int k2 = mia[t];
                    for (auto &e : row) {
                        if (e.first == t) miau[t] = k2;
                        mja[k2] = e.first;
                        msrc[k2] = e.second;
                        k2++;
                    }
                }
            });
            phase("multicolour: pattern");
            if (!build_merge_map_par(mia.data(), mja.data(), miau.data(), N, mmoff, mmqq, mc_nmap)) {
                cfd_mesh_destroy(m);
                return cfd_set_error(ctx, CFD_E_ARG, "cfd_mesh_create: ILU(0) merge map too large");
            }
            phase("multicolour: merge map");
            build_levels_by_colour(mia.data(), mja.data(), miau.data(), N, false, m->mc_cstart, m->mc_lower.ptr, mlrows,
                                   m->mc_lower.maxwidth);
            build_levels_by_colour(mia.data(), mja.data(), miau.data(), N, true, m->mc_cstart, m->mc_upper.ptr, murows,
                                   m->mc_upper.maxwidth);
            m->mc_lower.nlevels = (int)m->mc_lower.ptr.size() - 1;
            m->mc_upper.nlevels = (int)m->mc_upper.ptr.size() - 1;
        }
        phase("multicolour: levels");
        // neighbour lists (reference order within each list), internal numbering
        std::vector<int> nia(N + 1, 0), nja((size_t)d->n2n_ia[N]);
        for (int ii = 0; ii < N; ii++) nia[ii + 1] = nia[ii] + (d->n2n_ia[order[ii] + 1] - d->n2n_ia[order[ii]]);
        cfd_parallel_for(N, [&](int64_t i0, int64_t i1) {
            for (int ii = (int)i0; ii < (int)i1; ii++) {
                const int i = order[ii];
                int k = nia[ii];
                for (int q = d->n2n_ia[i]; q < d->n2n_ia[i + 1]; q++) nja[k++] = iperm[d->n2n_ja[q]];
            }
        });

        // packed node records: q (freestream), coordinates
        HostArray<double> qa((size_t)kQA * m->NT);
        std::vector<double> sa((size_t)4 * B, 0.0), vol(N);
        double q5[5];
        freestream_q(p, q5);
        cfd_parallel_for(m->NT, [&](int64_t i0, int64_t i1) {
            for (int i = (int)i0; i < (int)i1; i++) {
                for (int k = 0; k < 5; k++) qa[(size_t)kQA * i + k] = q5[k];
                for (int k = 0; k < 3; k++)  // ghost records hold no coordinates
                    qa[(size_t)kQA * i + 5 + k] = i < N ? d->xyz[3 * (size_t)order[i] + k] : 0.0;
                if (i < N) vol[i] = d->vol[order[i]];
            }
        });
        for (int b = 0; b < B; b++)
            for (int k = 0; k < 3; k++) sa[4 * (size_t)b + k] = d->bedge_area[3 * (size_t)b + k];
        std::vector<int> bnode(B), btag(B, 0);
        for (int b = 0; b < B; b++) {
            bnode[b] = iperm[d->bedge[2 * b]];
            if (d->bedge_tag) btag[b] = d->bedge_tag[b];
        }

        phase("node records");
        m->order = upload(m, order.data(), order.size(), &rc);
        if (!rc) m->inc_ptr = upload(m, ip.data(), ip.size(), &rc);
        if (!rc) m->sptr = upload(m, sptr.data(), sptr.size(), &rc);
        if (!rc) m->nbr = upload(m, nbr.data(), nbr.size(), &rc);
        if (!rc) m->soff = upload(m, soff.data(), soff.size(), &rc);
        if (!rc) m->d.isx = upload(m, isx.data(), isx.size(), &rc);
        if (!rc) m->d.isy = upload(m, isy.data(), isy.size(), &rc);
        if (!rc) m->d.isz = upload(m, isz.data(), isz.size(), &rc);
        if (!rc) m->iperm = upload(m, iperm.data(), iperm.size(), &rc);
        if (!rc) m->n2n_ia = upload(m, nia.data(), nia.size(), &rc);
        if (!rc) m->n2n_ja = upload(m, nja.data(), nja.size(), &rc);
        if (!rc) m->bnode = upload(m, bnode.data(), bnode.size(), &rc);
        if (!rc) m->btype = upload(m, d->bedge_type, (size_t)B, &rc);
        if (!rc) m->btag = upload(m, btag.data(), btag.size(), &rc);
        if (!rc) m->d.qa = upload(m, qa.data(), qa.size(), &rc);
        if (!rc) m->d.sa = upload(m, sa.data(), sa.size(), &rc);
        if (!rc) m->d.vol = upload(m, vol.data(), vol.size(), &rc);
        if (!rc) m->ia = upload(m, m->h_ia.data(), m->h_ia.size(), &rc);
        if (!rc) m->ja = upload(m, m->h_ja.data(), m->h_ja.size(), &rc);
        if (!rc) m->iau = upload(m, m->h_iau.data(), m->h_iau.size(), &rc);
        if (!rc) m->mc_perm = upload(m, mperm.data(), mperm.size(), &rc);
        if (!rc) m->mc_ia = upload(m, mia.data(), mia.size(), &rc);
        if (!rc) m->mc_ja = upload(m, mja.data(), mja.size(), &rc);
        if (!rc) m->mc_iau = upload(m, miau.data(), miau.size(), &rc);
        if (!rc) m->mc_src = upload(m, msrc.data(), msrc.size(), &rc);
        if (!rc) m->mc_moff = upload(m, mmoff.get(), (size_t)m->nnz, &rc);
        if (!rc) m->mc_mqq = upload(m, mmqq.get(), (size_t)std::max<int64_t>(mc_nmap, 1), &rc);
        phase("uploads");
        if (!rc) {  // block-SELL patterns of A and of the multicolour factors
            std::vector<int> off;
            HostArray<int> col(0), src(0);
            const std::vector<int> &hia = m->h_ia;
            build_block_sell(N, [&](int i) { return hia[i]; }, [&](int i) { return hia[i + 1]; }, m->h_ja.data(), off, col,
                             src);
            rc = upload_block_sell(m, off, col, src, m->a_sell);
            if (!rc) {
                build_block_sell(N, [&](int t) { return mia[t]; }, [&](int t) { return miau[t]; }, mja.data(), off, col,
                                 src);
                rc = upload_block_sell(m, off, col, src, m->mcl_sell);
            }
            if (!rc) {
                build_block_sell(N, [&](int t) { return miau[t] + 1; }, [&](int t) { return mia[t + 1]; }, mja.data(), off,
                                 col, src);
                rc = upload_block_sell(m, off, col, src, m->mcu_sell);
            }
        }
        phase("block-sell");
        if (!rc) m->mc_lower.rows = upload(m, mlrows.data(), mlrows.size(), &rc);
        if (!rc) m->mc_upper.rows = upload(m, murows.data(), murows.size(), &rc);
        if (!rc) m->lower.rows = upload(m, lrows.data(), lrows.size(), &rc);
        if (!rc) m->upper.rows = upload(m, urows.data(), urows.size(), &rc);
        {
            std::vector<int> s12;
            pad_levels(m->lower.ptr, lrows, 12, s12);
            m->lower.n12 = (int)s12.size();
            if (!rc) m->lower.slots12 = upload(m, s12.data(), s12.size(), &rc);
            pad_levels(m->upper.ptr, urows, 12, s12);
            m->upper.n12 = (int)s12.size();
            if (!rc) m->upper.slots12 = upload(m, s12.data(), s12.size(), &rc);
            std::vector<int> zeros((size_t)N + 1, 0);
            if (!rc) m->ready = upload(m, zeros.data(), (size_t)N, &rc);
            if (!rc) m->epoch = upload(m, zeros.data(), 1, &rc);
            const char *sf = getenv("CFD_SYNC_FREE");
            m->sync_free = !(sf && sf[0] == '0');
            const char *gr = getenv("CFD_GRAPHS");
            m->use_graphs = !(gr && gr[0] == '0');
            // one row per lane fills the GPU only with enough rows; smaller
            // systems keep five lanes per row (CFD_SELL=0/1 forces either)
            const char *se = getenv("CFD_SELL");
            m->use_sell = se ? se[0] == '1' : N >= kSellMinRows;
        }
        phase("levels upload");
        if (rc) {
            cfd_mesh_destroy(m);
            return rc;
        }
        m->d.qb = (double *)cfd_dalloc(m, sizeof(double) * kQB * (size_t)N);
        m->d.lsq = (double *)cfd_dalloc(m, sizeof(double) * 8 * (size_t)N);
        m->d.res = (double *)cfd_dalloc(m, sizeof(double) * kRES * (size_t)N);
        m->dt = (double *)cfd_dalloc(m, sizeof(double) * (size_t)N);
        m->b = (double *)cfd_dalloc(m, sizeof(double) * 5 * (size_t)N);
        m->x = (double *)cfd_dalloc(m, sizeof(double) * 5 * (size_t)N);
        m->roe_cand = (uint64_t *)cfd_dalloc(m, sizeof(uint64_t) * (size_t)N);
        m->roe_blk = (int *)cfd_dalloc(m, sizeof(int) * (size_t)((N + 4095) / 4096));
        m->risk = (int *)cfd_dalloc(m, sizeof(int));
        m->red = (double *)cfd_dalloc(m, sizeof(double) * 1024 * 65);
        m->redout = (double *)cfd_dalloc(m, sizeof(double) * 4 * 65);
        m->flag = (int *)cfd_dalloc(m, sizeof(int) * 4);
        m->eigbits = (unsigned long long *)cfd_dalloc(m, sizeof(unsigned long long) * 6);
        if (!m->d.qb || !m->d.lsq || !m->d.res || !m->dt || !m->b || !m->x || !m->roe_cand || !m->roe_blk ||
            !m->risk || !m->red || !m->redout || !m->flag || !m->eigbits) {
            cfd_mesh_destroy(m);
            return cfd_set_error(ctx, CFD_E_OOM, "cfd_mesh_create: out of device memory");
        }
        if (cfd_host_alloc((void **)&m->hred, sizeof(double) * 4 * 65) != hipSuccess ||
            cfd_host_alloc((void **)&m->hflag, sizeof(int) * 4) != hipSuccess) {
            cfd_mesh_destroy(m);
            return cfd_set_error(ctx, CFD_E_OOM, "cfd_mesh_create: pinned allocation failed");
        }
        hipMemsetAsync(m->d.qb, 0, sizeof(double) * kQB * (size_t)N, m->stream);
        hipMemsetAsync(m->roe_blk, 0, sizeof(int) * (size_t)((N + 4095) / 4096), m->stream);
        hipMemsetAsync(m->flag, 0, sizeof(int) * 4, m->stream);
        hipMemsetAsync(m->d.res, 0, sizeof(double) * kRES * (size_t)N, m->stream);
        phase("workspaces");
        rc = launch_lsq_coef(m);
        if (!rc && hipStreamSynchronize(m->stream) != hipSuccess) rc = cfd_set_error(ctx, CFD_E_HIP, "LSQ setup failed");
        if (rc) {
            cfd_mesh_destroy(m);
            return rc;
        }
    } catch (const std::bad_alloc &) {
        cfd_mesh_destroy(m);
        return cfd_set_error(ctx, CFD_E_OOM, "cfd_mesh_create: host out of memory");
    }
    phase("lsq");
    *out = m;
    return CFD_OK;
}

// The natural-order ILU(0) merge map (cfd_mesh::ilu_moff, ilu_mqq), built
// and uploaded on the first natural-order factorisation: the multicolour
// factorisation, the sweeps and the Gauss-Seidel solvers do not use it.
int cfd_ensure_ilu_map(cfd_mesh *m) {
    if (m->ilu_moff) return CFD_OK;
    std::unique_ptr<int[]> moff, mqq;
    int64_t total = 0;
    if (!build_merge_map_par(m->h_ia.data(), m->h_ja.data(), m->h_iau.data(), m->N, moff, mqq, total))
        return cfd_set_error(m->ctx, CFD_E_ARG, "ILU(0) merge map too large");
    int rc = CFD_OK;
    int *mo = upload(m, moff.get(), (size_t)m->nnz, &rc);
    int *mq = rc ? nullptr : upload(m, mqq.get(), (size_t)std::max<int64_t>(total, 1), &rc);
    if (rc) return rc;
    m->ilu_moff = mo;
    m->ilu_mqq = mq;
    m->ilu_nmap = total;
    return CFD_OK;
}

extern "C" int64_t cfd_mesh_device_bytes(const cfd_mesh *m) { return m ? m->bytes : 0; }

// Device bytes of a mesh handle after a steady solve with FGMRES(km): every
// allocation of cfd_mesh_create plus the Jacobian, ILU(0) and Krylov
// workspace.  Layout padding is bounded: SELL slices pad each slice to its
// widest node (taken as <= 25% of the incidences plus one slice), levels pad
// to whole wavefronts (counted for up to 4096 levels).
extern "C" int64_t cfd_footprint_bytes(int32_t N, int32_t E, int32_t B, int32_t km, int32_t mixed) {
    const int64_t n = N, e = E, b = B, nnz = n + 2 * e, inc = 2 * e + b;
    const int64_t s = mixed ? 4 : 8;
    const int64_t nsell = inc + inc / 4 + (int64_t)kSlice * 32;
    const int64_t lvl = std::min<int64_t>(n, 4096);
    int64_t bytes = 0;
    // qa, qb, lsq, res, vol, dt, boundary areas, right-hand side, update
    bytes += 8 * (kQA * (n + b) + kQB * n + 8 * n + kRES * n + n + n + 4 * b + 5 * n + 5 * n);
    // order, inverse order, inc_ptr, sptr, n2n lists, boundary node/type/tag, Roe-pass candidates and block
    // flags, ready flags, epoch, risk
    bytes += 4 * (n + n + (n + 1) + (n / kSlice + 2) + (n + 1) + 2 * e + 3 * b + 2 * n + (n / 4096 + 1) + n + 2);
    bytes += nsell * (4 + 4 + 3 * 8);                       // nbr, soff, isx, isy, isz
    bytes += 4 * ((n + 1) + nnz + n) + 4 * nnz;             // BSR ia, ja, iau; ILU(0) merge offsets
    bytes += 4 * 8 * nnz;                                   // ILU(0) merge map (<= 8 upper blocks per lower entry)
    bytes += 4 * (n + (n + 1) + 3 * nnz + n + 2 * n) + 4 * 8 * nnz;  // multicolour pattern, slot map, levels, merge map
    bytes += s * 5 * n;                                     // multicolour work vector
    bytes += 4 * (2 * n + 2 * n + (11 + 11) * lvl);        // level rows and padded slot lists
    bytes += 8 * (1024 * 65 + 4 * 65 + 6) + 4 * 4;         // reduction scratch, eigenvalue bits, flags
    bytes += 25 * s * nnz * 2 + 25 * s * n;                 // Jacobian, ILU(0) factors, inverse pivots
    bytes += s * 5 * n * (2 * (int64_t)km + 2) + 8 * 5 * n; // Krylov V, Z, w; FP64 true residual
    // block-SELL copies: A, and the multicolour factors (strict lower + upper
    // blocks, inverted diagonals); slices padded to their longest row, taken
    // as <= 25% of the blocks plus 64 slots per slice
    const int64_t sell = nnz + nnz / 4 + 64 * (n / 64 + 1);
    bytes += 2 * sell * (4 + 4 + 25 * s) + 25 * s * (n + 64);
    return bytes;
}

extern "C" void *cfd_mesh_stream(cfd_mesh *m) { return m ? (void *)m->stream : nullptr; }

extern "C" int cfd_mesh_sync(cfd_mesh *m) {
    if (!m) return CFD_E_ARG;
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    return CFD_OK;
}

int cfd_ensure_fp32(cfd_mesh *m) {
    if (!m->f_ready) {
        const size_t N = m->N, NT = m->NT, EB = (size_t)m->E + m->B;
        m->f.qa = (float *)cfd_dalloc(m, sizeof(float) * kQA * NT);
        m->f.qb = (float *)cfd_dalloc(m, sizeof(float) * kQB * N);
        m->f.sa = (float *)cfd_dalloc(m, sizeof(float) * 4 * (size_t)m->B);
        m->f.vol = (float *)cfd_dalloc(m, sizeof(float) * N);
        m->f.lsq = (float *)cfd_dalloc(m, sizeof(float) * 8 * N);
        m->f.res = (float *)cfd_dalloc(m, sizeof(float) * kRES * N);
        m->f.isx = (float *)cfd_dalloc(m, sizeof(float) * m->nsell);
        m->f.isy = (float *)cfd_dalloc(m, sizeof(float) * m->nsell);
        m->f.isz = (float *)cfd_dalloc(m, sizeof(float) * m->nsell);
        if (!m->f.qa || !m->f.qb || !m->f.sa || !m->f.vol || !m->f.lsq || !m->f.res || !m->f.isx || !m->f.isy ||
            !m->f.isz)
            return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (FP32 buffers)");
        int rc = launch_convert_geometry_fp32(m);
        if (rc) return rc;
        m->f_ready = true;
        m->f_version = -1;
    }
    if (m->f_version != m->q_version) {
        int rc = launch_convert_state_to_fp32(m);
        if (rc) return rc;
        m->f_version = m->q_version;
    }
    return CFD_OK;
}

// ---------------------------------------------------------------------------
// Work-edge patches of the flux-once residual.  Built from the device
// incidence lists: a node's incidence with neighbour j is an interior edge
// (node is L when nbr >= 0, R when nbr = ~j) or boundary edge b (nbr = N + b),
// with the edge's area vector.  An edge inside a patch becomes one work edge
// (taken from its L end); an edge between patches becomes a work edge of
// each; a boundary edge one of its node's patch.  Work edges of a patch are
// ordered by reference edge index (boundary edges after the interior ones),
// and each node lists the slots of its incidences in ascending order (sign
// bit: the node is R), which is the order of its incidence list: each node
// adds its fluxes in the reference's order.  Patches hold kPatch consecutive
// nodes, restarting at every case of a union.
// ---------------------------------------------------------------------------
constexpr int kPatch = 256;  // nodes per patch: one per thread of the patch kernel's workgroup

// The fast build runs the flux-once patch kernel at second order; the exact
// build runs the node-centred kernel, since its patches (whose work edges go
// in the reference's edge order, giving the same bits) cost a host build of
// about 0.15 s per million nodes, more than a steady solve's residual
// evaluations save.  CFD_FLUX_ONCE=0/1 (read at each call) forces either
// kernel in both builds.
bool cfd_residual_uses_patches(const cfd_params *p) {
    const char *const once = getenv("CFD_FLUX_ONCE");
    if (once && (once[0] == '0' || once[0] == '1')) return once[0] == '1';
    return p->arith == CFD_ARITH_FAST && p->order == 2;
}

int64_t cfd_patch_footprint_bytes(int64_t n, int64_t e, int64_t b, int64_t cases) {
    const int64_t inc = 2 * e + b, nsell = inc + inc / 4 + (int64_t)kSlice * 32;
    // per work edge (at most one per incidence): its two nodes and its area
    // vector in FP64 and FP32; a slot per incidence list entry; patch offsets
    return inc * (4 + 4 + 24 + 12) + 4 * nsell + 8 * (n / kPatch + cases + 2);
}

int cfd_ensure_patches(cfd_mesh *m, bool fp32, int order) {
    PatchSet &ps = m->patch[order];
    if (!ps.pw_ptr) {
        const auto t0 = std::chrono::steady_clock::now();
        const int N = m->N, nsl = (N + kSlice - 1) / kSlice;
        // patches of kPatch nodes from the start of the mesh, or of each case
        // of a union (so a case's patches, and its sums, are those it has alone)
        std::vector<int> pstart;
        const std::vector<int> whole{0, N};
        const std::vector<int> &cut = m->ncase ? m->case_node : whole;
        for (size_t c = 0; c + 1 < cut.size(); c++)
            for (int n = cut[c]; n < cut[c + 1]; n += kPatch) pstart.push_back(n);
        pstart.push_back(N);
        const int P = (int)pstart.size() - 1;
        std::vector<int> ip(N + 1), sptr(nsl + 1);
        HostArray<int> nbr(m->nsell);  // (downloaded in full)
        HostArray<double> sx(m->nsell), sy(m->nsell), sz(m->nsell);
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        CFD_HIP(m, cfd_memcpy(ip.data(), m->inc_ptr, sizeof(int) * ip.size(), hipMemcpyDeviceToHost));
        CFD_HIP(m, cfd_memcpy(sptr.data(), m->sptr, sizeof(int) * sptr.size(), hipMemcpyDeviceToHost));
        CFD_HIP(m, cfd_memcpy(nbr.data(), m->nbr, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
        CFD_HIP(m, cfd_memcpy(sx.data(), m->d.isx, sizeof(double) * sx.size(), hipMemcpyDeviceToHost));
        CFD_HIP(m, cfd_memcpy(sy.data(), m->d.isy, sizeof(double) * sy.size(), hipMemcpyDeviceToHost));
        CFD_HIP(m, cfd_memcpy(sz.data(), m->d.isz, sizeof(double) * sz.size(), hipMemcpyDeviceToHost));
        auto sell = [&](int i, int k) { return (size_t)sptr[i / kSlice] + (size_t)kSlice * k + i % kSlice; };
        // work edges of each patch: (key, l, r, incidence of the area vector);
        // key: the first end in the patch (node order) or the reference edge
        // index (edge order: interior edges by index, then boundary edges by
        // index, the order of each node's incidence list, along which its
        // slots then ascend)
        const bool by_edge = order == kPatchEdgeOrder;
        struct Work {
            int key, l, r;
            size_t inc;
        };
        std::vector<std::vector<Work>> pw(P);
        cfd_parallel_for(
            P,
            [&](int64_t p0, int64_t p1) {
                for (int p = (int)p0; p < (int)p1; p++) {
                    const int n0 = pstart[p], n1 = pstart[p + 1];
                    std::vector<Work> &w = pw[p];
                    for (int i = n0; i < n1; i++) {
                        for (int k = 0; k < ip[i + 1] - ip[i]; k++) {
                            const size_t x = sell(i, k);
                            const int jr = nbr[x];
                            if (jr >= N) {
                                w.push_back({by_edge ? m->h_sell_edge[x] : i - n0, i, jr, x});
                                continue;
                            }
                            const bool isL = jr >= 0;
                            const int j = isL ? jr : ~jr;
                            const bool inside = j >= n0 && j < n1;
                            if (inside && !isL) continue;  // taken from its L end
                            const int key = by_edge ? m->h_sell_edge[x] : (inside ? std::min(i, j) : i) - n0;
                            w.push_back({key, isL ? i : j, isL ? j : i, x});
                        }
                    }
                    // (edge order: the keys are distinct; node order: ties by L, then R)
                    std::sort(w.begin(), w.end(), [](const Work &a, const Work &b) {
                        return a.key != b.key ? a.key < b.key : (a.l != b.l ? a.l < b.l : a.r < b.r);
                    });
                }
            },
            1);
        std::vector<int> wptr(P + 1, 0);
        for (int p = 0; p < P; p++) wptr[p + 1] = wptr[p] + (int)pw[p].size();
        const int64_t W = wptr[P];
        HostArray<int> wl(W), wr(W);  // every entry written below
        HostArray<double> ws(3 * (size_t)W);
        std::vector<int> slot(m->nsell, 0);
        bool ok = true;
        std::vector<char> bad(P, 0);
        cfd_parallel_for(
            P,
            [&](int64_t p0, int64_t p1) {
                std::vector<int> cnt;
                for (int p = (int)p0; p < (int)p1; p++) {
                    const int n0 = pstart[p], n1 = pstart[p + 1];
                    // each node's slots, ascending as the work edges are visited in order
                    cnt.assign(n1 - n0, 0);
                    auto add = [&](int i, int v) {
                        int &c = cnt[i - n0];
                        if (c < ip[i + 1] - ip[i]) slot[sell(i, c)] = v;
                        c++;
                    };
                    for (int q = 0; q < (int)pw[p].size(); q++) {
                        const Work &u = pw[p][q];
                        const size_t o = (size_t)wptr[p] + q;
                        wl[o] = u.l;
                        wr[o] = u.r;
                        ws[3 * o] = sx[u.inc];
                        ws[3 * o + 1] = sy[u.inc];
                        ws[3 * o + 2] = sz[u.inc];
                        if (u.l >= n0 && u.l < n1) add(u.l, q);
                        if (u.r >= n0 && u.r < n1) add(u.r, q | (int)0x80000000u);
                    }
                    for (int i = n0; i < n1; i++)
                        if (cnt[i - n0] != ip[i + 1] - ip[i]) bad[p] = 1;
                }
            },
            1);
        for (char b : bad) ok = ok && !b;
        if (!ok) return cfd_set_error(m->ctx, CFD_E_STATE, "work-edge patches do not cover the incidence lists");
        int rc = 0;
        ps.pw_ptr = upload(m, wptr.data(), wptr.size(), &rc);
        if (!rc) ps.pn_start = upload(m, pstart.data(), pstart.size(), &rc);
        if (!rc) ps.pw_l = upload(m, wl.data(), wl.size(), &rc);
        if (!rc) ps.pw_r = upload(m, wr.data(), wr.size(), &rc);
        if (!rc) ps.pn_slot = upload(m, slot.data(), slot.size(), &rc);
        if (!rc) ps.ws = upload(m, ws.data(), ws.size(), &rc);
        if (rc) {
            ps.pw_ptr = nullptr;
            return rc;
        }
        ps.npatch = P;
        ps.npw = W;
        if (cfd_progress())
            fprintf(stderr, "cfd_ensure_patches: %d patches, %lld work edges (%s order) %8.3f s\n", P, (long long)W,
                    order == kPatchEdgeOrder ? "edge" : "node",
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    if (fp32 && !ps.ws_f) {
        std::vector<double> ws(3 * (size_t)ps.npw);
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        CFD_HIP(m, cfd_memcpy(ws.data(), ps.ws, sizeof(double) * ws.size(), hipMemcpyDeviceToHost));
        std::vector<float> wf(ws.begin(), ws.end());
        int rc = 0;
        ps.ws_f = upload(m, wf.data(), wf.size(), &rc);
        if (rc) return rc;
    }
    return CFD_OK;
}

// ---------------------------------------------------------------------------
// state reset (for reuse of a resident mesh by another solve)
// ---------------------------------------------------------------------------
static int check_params(cfd_mesh *m, const cfd_params *p);

__global__ void k_set_state(int NT, double q0, double q1, double q2, double q3, double q4, double *qa) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NT) return;
    double *r = qa + kQA * (size_t)i;
    r[0] = q0;
    r[1] = q1;
    r[2] = q2;
    r[3] = q3;
    r[4] = q4;
}

extern "C" int cfd_mesh_reset(cfd_mesh *m, const cfd_params *p) {
    int rc = check_params(m, p);
    if (rc) return rc;
    hipSetDevice(m->ctx->device);
    double q5[5];
    freestream_q(p, q5);
    hipLaunchKernelGGL(k_set_state, dim3((m->NT + 255) / 256), dim3(256), 0, m->stream, m->NT, q5[0], q5[1], q5[2],
                       q5[3], q5[4], m->d.qa);
    CFD_HIP(m, hipGetLastError());
    m->q_version++;
    m->jac_valid = false;
    m->prec_valid = false;
    m->asell_valid = false;
    m->residual_valid = false;
    return CFD_OK;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
// reference (API) index of internal node record i (ghosts are not renumbered)
static inline size_t api_index(const cfd_mesh *m, size_t i) { return i < (size_t)m->N ? (size_t)m->h_order[i] : i; }

extern "C" int cfd_field_set(cfd_mesh *m, int field, const double *h) {
    if (!m || !h) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    const size_t N = m->N, NT = m->NT;
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    if (field == CFD_FIELD_Q || field == CFD_FIELD_GRAD || field == CFD_FIELD_PHI) {
        const bool isq = field == CFD_FIELD_Q;
        const size_t rec = isq ? kQA : kQB, n = isq ? NT : N;
        double *dev = isq ? m->d.qa : m->d.qb;
        std::vector<double> buf(rec * n);
        CFD_HIP(m, cfd_memcpy(buf.data(), dev, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        const int nf = field == CFD_FIELD_GRAD ? 15 : 5, o = field == CFD_FIELD_PHI ? kPHI : 0;
        for (int f = 0; f < nf; f++)
            for (size_t i = 0; i < n; i++) buf[rec * i + o + f] = h[f * n + api_index(m, i)];
        CFD_HIP(m, cfd_memcpy(dev, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
        m->q_version++;
        return CFD_OK;
    }
    if (field == CFD_FIELD_RES_CONV || field == CFD_FIELD_RES_DISS || field == CFD_FIELD_DT_ACC) {
        std::vector<double> buf((size_t)kRES * N);
        CFD_HIP(m, cfd_memcpy(buf.data(), m->d.res, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        const int nf = field == CFD_FIELD_DT_ACC ? 1 : 5;
        const int o = field == CFD_FIELD_RES_CONV ? 0 : (field == CFD_FIELD_RES_DISS ? 5 : 10);
        for (int f = 0; f < nf; f++)
            for (size_t i = 0; i < N; i++) buf[kRES * i + o + f] = h[f * N + api_index(m, i)];
        CFD_HIP(m, cfd_memcpy(m->d.res, buf.data(), sizeof(double) * buf.size(), hipMemcpyHostToDevice));
        m->res_prec = CFD_FP64;
        return CFD_OK;
    }
    if (field == CFD_FIELD_DT) {
        std::vector<double> buf(N);
        for (size_t i = 0; i < N; i++) buf[i] = h[api_index(m, i)];
        CFD_HIP(m, cfd_memcpy(m->dt, buf.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        return CFD_OK;
    }
    double *dev = nullptr;
    size_t n = 0;
    switch (field) {
    case CFD_FIELD_JAC_B: dev = m->b; n = 5 * N; break;
    case CFD_FIELD_X: dev = m->x; n = 5 * N; break;
    default: return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_field_set: unknown field %d", field);
    }
    CFD_HIP(m, cfd_memcpy(dev, h, sizeof(double) * n, hipMemcpyHostToDevice));
    return CFD_OK;
}

static int check_residual_flags(cfd_mesh *m);

extern "C" int cfd_field_get(cfd_mesh *m, int field, double *h) {
    if (!m || !h) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    const size_t N = m->N, NT = m->NT;
    if (field == CFD_FIELD_RES_CONV || field == CFD_FIELD_RES_DISS || field == CFD_FIELD_DT_ACC) {
        int rc = check_residual_flags(m);
        if (rc) return rc;
    }
    if ((field == CFD_FIELD_RES_CONV || field == CFD_FIELD_RES_DISS || field == CFD_FIELD_DT_ACC) &&
        m->res_prec == CFD_FP32) {
        int rc = launch_residual_to_fp64(m);
        if (rc) return rc;
        m->res_prec = CFD_FP64;
    }
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    if (field == CFD_FIELD_Q || field == CFD_FIELD_GRAD || field == CFD_FIELD_PHI) {
        const bool isq = field == CFD_FIELD_Q;
        const size_t rec = isq ? kQA : kQB, n = isq ? NT : N;
        std::vector<double> buf(rec * n);
        if (!isq && m->f_ready && m->res_prec == CFD_FP32) {
            // not reached: FP32 fields are exposed through the FP64 mirror only
        }
        CFD_HIP(m, cfd_memcpy(buf.data(), isq ? m->d.qa : m->d.qb, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        const int nf = field == CFD_FIELD_GRAD ? 15 : 5, o = field == CFD_FIELD_PHI ? kPHI : 0;
        for (int f = 0; f < nf; f++)
            for (size_t i = 0; i < n; i++) h[f * n + api_index(m, i)] = buf[rec * i + o + f];
        return CFD_OK;
    }
    if (field == CFD_FIELD_RES_CONV || field == CFD_FIELD_RES_DISS || field == CFD_FIELD_DT_ACC) {
        std::vector<double> buf((size_t)kRES * N);
        CFD_HIP(m, cfd_memcpy(buf.data(), m->d.res, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        const int nf = field == CFD_FIELD_DT_ACC ? 1 : 5;
        const int o = field == CFD_FIELD_RES_CONV ? 0 : (field == CFD_FIELD_RES_DISS ? 5 : 10);
        for (int f = 0; f < nf; f++)
            for (size_t i = 0; i < N; i++) h[f * N + api_index(m, i)] = buf[kRES * i + o + f];
        return CFD_OK;
    }
    if (field == CFD_FIELD_DT) {
        std::vector<double> buf(N);
        CFD_HIP(m, cfd_memcpy(buf.data(), m->dt, sizeof(double) * N, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) h[api_index(m, i)] = buf[i];
        return CFD_OK;
    }
    const double *dev = nullptr;
    size_t n = 0;
    switch (field) {
    case CFD_FIELD_JAC_B: dev = m->b; n = 5 * N; break;
    case CFD_FIELD_X: dev = m->x; n = 5 * N; break;
    default: return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_field_get: unknown field %d", field);
    }
    CFD_HIP(m, cfd_memcpy(h, dev, sizeof(double) * n, hipMemcpyDeviceToHost));
    return CFD_OK;
}

// ---------------------------------------------------------------------------
// hot-path entry points
// ---------------------------------------------------------------------------
static int check_residual_flags(cfd_mesh *m);

static int check_params(cfd_mesh *m, const cfd_params *p) {
    if (!m || !p) return CFD_E_ARG;
    if (p->flux < 0 || p->flux > 2 || (p->order != 1 && p->order != 2) || p->limiter < 0 || p->limiter > 2 ||
        p->gradient < 0 || p->gradient > 1 || (p->precision != CFD_FP64 && p->precision != CFD_FP32) ||
        (p->arith != CFD_ARITH_EXACT && p->arith != CFD_ARITH_FAST) || !(p->gamma > 1.0))
        return cfd_set_error(m->ctx, CFD_E_ARG, "invalid cfd_params");
    hipSetDevice(m->ctx->device);
    return CFD_OK;
}

static int read_flags(cfd_mesh *m) {
    CFD_HIP(m, hipMemcpyAsync(m->hflag, m->flag, 4 * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    return CFD_OK;
}

extern "C" int cfd_bc(cfd_mesh *m, const cfd_params *p, int iter) {
    int rc = check_params(m, p);
    if (rc) return rc;
    CFD_HIP(m, hipMemsetAsync(m->flag, 0, 4 * sizeof(int), m->stream));
    if (p->precision == CFD_FP32) {
        rc = cfd_ensure_fp32(m);
        if (!rc) rc = launch_bc<float>(m, m->f, p, iter);
    } else {
        rc = launch_bc<double>(m, m->d, p, iter);
        m->q_version++;
    }
    if (rc) return rc;
    rc = read_flags(m);
    if (rc) return rc;
    if (m->hflag[3])
        return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_bc: %d boundary edges with an unusable wave configuration",
                             m->hflag[3]);
    if (m->hflag[0])
        return cfd_set_error(m->ctx, CFD_E_NEGATIVE_PRESSURE, "cfd_bc: negative pressure on %d boundary edges",
                             m->hflag[0]);
    if (p->precision == CFD_FP32) m->f_version = m->q_version;
    return CFD_OK;
}

// Runs the launcher of the build p->arith selects (res_exact or res_fast)
// on the buffers of p->precision.
#define CFD_RES_DISPATCH(m, p, launcher, ...)                                                             \
    ((p)->precision == CFD_FP32                                                                           \
         ? ((p)->arith == CFD_ARITH_FAST ? res_fast::launcher<float>((m), (m)->f, (p), ##__VA_ARGS__)      \
                                         : res_exact::launcher<float>((m), (m)->f, (p), ##__VA_ARGS__))    \
         : ((p)->arith == CFD_ARITH_FAST ? res_fast::launcher<double>((m), (m)->d, (p), ##__VA_ARGS__)     \
                                         : res_exact::launcher<double>((m), (m)->d, (p), ##__VA_ARGS__)))

extern "C" int cfd_gradient(cfd_mesh *m, const cfd_params *p) {
    int rc = check_params(m, p);
    if (!rc && p->precision == CFD_FP32) rc = cfd_ensure_fp32(m);
    return rc ? rc : CFD_RES_DISPATCH(m, p, launch_gradient);
}

extern "C" int cfd_limiter(cfd_mesh *m, const cfd_params *p) {
    int rc = check_params(m, p);
    if (!rc && p->precision == CFD_FP32) rc = cfd_ensure_fp32(m);
    return rc ? rc : CFD_RES_DISPATCH(m, p, launch_limiter);
}

extern "C" int cfd_gradient_limiter(cfd_mesh *m, const cfd_params *p) {
    int rc = check_params(m, p);
    if (!rc && p->precision == CFD_FP32) rc = cfd_ensure_fp32(m);
    return rc ? rc : CFD_RES_DISPATCH(m, p, launch_gradient_limiter);
}

extern "C" int cfd_residual(cfd_mesh *m, const cfd_params *p, int add_time, double *eig6) {
    (void)add_time;  // the accumulator is always produced; it is only consumed with AddTime
    int rc = check_params(m, p);
    if (rc) return rc;
    const bool stats = eig6 != nullptr;
    if (stats) {
        unsigned long long init[6];
        const double mn = DBL_MAX, mx = DBL_MIN;
        for (int t = 0; t < 3; t++) {
            memcpy(&init[2 * t], &mn, 8);
            memcpy(&init[2 * t + 1], &mx, 8);
        }
        CFD_HIP(m, hipMemcpyAsync(m->eigbits, init, sizeof(init), hipMemcpyHostToDevice, m->stream));
    }
    CFD_HIP(m, hipMemsetAsync(m->flag, 0, 4 * sizeof(int), m->stream));
    if (p->precision == CFD_FP32) rc = cfd_ensure_fp32(m);
    if (!rc) rc = CFD_RES_DISPATCH(m, p, launch_residual, stats);
    m->res_prec = p->precision == CFD_FP32 ? CFD_FP32 : CFD_FP64;
    if (rc) return rc;
    m->residual_valid = true;
    m->hllc_check = p->flux == CFD_FLUX_HLLC;
    if (stats) {
        unsigned long long bits[6];
        CFD_HIP(m, hipMemcpyAsync(bits, m->eigbits, sizeof(bits), hipMemcpyDeviceToHost, m->stream));
        CFD_HIP(m, hipStreamSynchronize(m->stream));
        for (int t = 0; t < 6; t++) memcpy(&eig6[t], &bits[t], 8);
    }
    return CFD_OK;
}

// Reports an HLLC wave-configuration failure of the last residual (read back
// lazily so the residual launch itself never waits on the host).
static int check_residual_flags(cfd_mesh *m) {
    if (!m->hllc_check) return CFD_OK;
    m->hllc_check = false;
    int rc = read_flags(m);
    if (rc) return rc;
    if (m->hflag[1]) return cfd_set_error(m->ctx, CFD_E_HLLC, "HLLC: no wave configuration on %d nodes", m->hflag[1]);
    return CFD_OK;
}

static double cfl_of(const cfd_params *p, int iter) {
    // Compute_DeltaT, Euler3D/src/Time_Step.cpp:27-35
    if ((p->cfl_ramp > 1) && (p->cfl_max > p->cfl_min)) {
        if (iter < p->cfl_ramp)
            return p->cfl_min + (p->cfl_max - p->cfl_min) * (((double)iter) / ((double)(p->cfl_ramp - 1)));
        return p->cfl_max;
    }
    return p->cfl_max;
}

extern "C" int cfd_delta_t(cfd_mesh *m, const cfd_params *p, int iter, double *cfl) {
    int rc = check_params(m, p);
    if (rc) return rc;
    if (m->res_prec != CFD_FP64) {
        rc = launch_residual_to_fp64(m);
        if (rc) return rc;
        m->res_prec = CFD_FP64;
    }
    const double c = cfl_of(p, iter);
    if (cfl) *cfl = c;
    return launch_delta_t(m, c);
}

extern "C" int cfd_residual_rms(cfd_mesh *m, double *rms) {
    if (!m || !rms) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    if (m->res_prec != CFD_FP64) {
        int rc = launch_residual_to_fp64(m);
        if (rc) return rc;
        m->res_prec = CFD_FP64;
    }
    double s[5];
    int rc = check_residual_flags(m);
    if (rc) return rc;
    rc = launch_rms(m, s);
    if (rc) return rc;
    double t = s[0] + s[1] + s[2] + s[3] + s[4];
    *rms = sqrt(t / (5.0 * (double)m->N));
    return CFD_OK;
}

static int ensure_jacobian_storage(cfd_mesh *m, int mixed) {
    if (mixed) {
        if (!m->af) m->af = (float *)cfd_dalloc(m, sizeof(float) * 25 * (size_t)m->nnz);
        if (!m->af) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (FP32 Jacobian)");
    } else {
        if (!m->a) m->a = (double *)cfd_dalloc(m, sizeof(double) * 25 * (size_t)m->nnz);
        if (!m->a) return cfd_set_error(m->ctx, CFD_E_OOM, "out of device memory (Jacobian)");
    }
    return CFD_OK;
}

extern "C" int cfd_jacobian(cfd_mesh *m, const cfd_params *p, int rhs_mode, int mixed) {
    int rc = check_params(m, p);
    if (rc) return rc;
    if (!m->residual_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "cfd_jacobian: call cfd_residual first");
    rc = check_residual_flags(m);
    if (rc) return rc;
    if (m->res_prec != CFD_FP64) {
        rc = launch_residual_to_fp64(m);
        if (rc) return rc;
        m->res_prec = CFD_FP64;
    }
    rc = ensure_jacobian_storage(m, mixed);
    if (rc) return rc;
    return launch_jacobian(m, p, rhs_mode, mixed);
}

extern "C" int cfd_jacobian_fd(cfd_mesh *m, const cfd_params *p, int method, int iter, int add_time, int rhs_mode,
                               int mixed) {
    int rc = check_params(m, p);
    if (rc) return rc;
    if (method < CFD_JAC_FD_CENTRAL || method > CFD_JAC_FD_ALTERNATE || iter < 0)
        return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_jacobian_fd: invalid method %d or iteration %d", method, iter);
    if (p->precision != CFD_FP64)
        return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_jacobian_fd: needs the FP64 state path (precision CFD_FP64)");
    if (!m->residual_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "cfd_jacobian_fd: call cfd_residual first");
    rc = check_residual_flags(m);
    if (rc) return rc;
    if (m->res_prec != CFD_FP64) {
        rc = launch_residual_to_fp64(m);
        if (rc) return rc;
        m->res_prec = CFD_FP64;
    }
    rc = ensure_jacobian_storage(m, mixed);
    if (rc) return rc;
    CFD_HIP(m, hipMemsetAsync(m->flag, 0, 4 * sizeof(int), m->stream));
    rc = launch_jacobian_fd(m, p, method, iter, add_time, rhs_mode, mixed);
    if (rc) return rc;
    // a boundary instance that does not apply, or a negative ghost pressure,
    // in a perturbed state (the reference stops with error())
    rc = read_flags(m);
    if (rc) return rc;
    if (m->hflag[3])
        return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_jacobian_fd: %d perturbed boundary states without a wave "
                             "configuration", m->hflag[3]);
    if (m->hflag[0])
        return cfd_set_error(m->ctx, CFD_E_NEGATIVE_PRESSURE, "cfd_jacobian_fd: %d perturbed ghost states with "
                             "negative pressure", m->hflag[0]);
    if (m->hflag[1])
        return cfd_set_error(m->ctx, CFD_E_HLLC, "cfd_jacobian_fd: HLLC without a wave configuration in %d perturbed "
                             "fluxes", m->hflag[1]);
    return CFD_OK;
}

extern "C" int32_t cfd_bsr_structure(cfd_mesh *m, int32_t *ia, int32_t *ja, int32_t *iau) {
    if (!m) return CFD_E_ARG;
    if (ia) memcpy(ia, m->h_ia.data(), sizeof(int) * m->h_ia.size());
    if (ja) memcpy(ja, m->h_ja.data(), sizeof(int) * m->h_ja.size());
    if (iau) memcpy(iau, m->h_iau.data(), sizeof(int) * m->h_iau.size());
    return m->nnz;
}

extern "C" int cfd_jacobian_get(cfd_mesh *m, double *a, double *b) {
    if (!m) return CFD_E_ARG;
    if (!m->jac_valid) return cfd_set_error(m->ctx, CFD_E_STATE, "cfd_jacobian_get: no Jacobian");
    hipSetDevice(m->ctx->device);
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    if (a) {
        if (m->jac_mixed) {
            std::vector<float> tmp(25 * (size_t)m->nnz);
            CFD_HIP(m, cfd_memcpy(tmp.data(), m->af, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < tmp.size(); i++) a[i] = tmp[i];
        } else {
            CFD_HIP(m, cfd_memcpy(a, m->a, sizeof(double) * 25 * (size_t)m->nnz, hipMemcpyDeviceToHost));
        }
    }
    if (b) CFD_HIP(m, cfd_memcpy(b, m->b, sizeof(double) * 5 * (size_t)m->N, hipMemcpyDeviceToHost));
    return CFD_OK;
}

extern "C" int cfd_jacobian_set(cfd_mesh *m, const double *a, const double *b) {
    if (!m || !a || !b) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    int rc = ensure_jacobian_storage(m, 0);
    if (rc) return rc;
    CFD_HIP(m, hipStreamSynchronize(m->stream));
    CFD_HIP(m, cfd_memcpy(m->a, a, sizeof(double) * 25 * (size_t)m->nnz, hipMemcpyHostToDevice));
    CFD_HIP(m, cfd_memcpy(m->b, b, sizeof(double) * 5 * (size_t)m->N, hipMemcpyHostToDevice));
    m->jac_valid = true;
    m->jac_mixed = false;
    m->prec_valid = false;
    m->asell_valid = false;
    return CFD_OK;
}

extern "C" int cfd_linsolve(cfd_mesh *m, const cfd_linsolve_opts *o, cfd_linsolve_stats *st) {
    if (!m || !o) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    cfd_linsolve_stats tmp;
    return run_linsolve(m, o, st ? st : &tmp);
}

extern "C" int cfd_linsolve_gs_noalloc(cfd_mesh *m, int sweeps, cfd_linsolve_stats *st) {
    if (!m) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    cfd_linsolve_stats tmp;
    return run_gs_noalloc(m, sweeps, st ? st : &tmp);
}

extern "C" int cfd_update(cfd_mesh *m, double relax) {
    if (!m) return CFD_E_ARG;
    hipSetDevice(m->ctx->device);
    m->q_version++;
    return launch_update(m, relax);
}

extern "C" int cfd_forces(cfd_mesh *m, const cfd_params *p, int wall_tag, double alpha_deg, double *force3,
                          double *lift, double *drag, double *pint) {
    int rc = check_params(m, p);
    if (rc) return rc;
    double v[4];
    rc = launch_forces(m, p, wall_tag, v);
    if (rc) return rc;
    const double a = alpha_deg * M_PI / 180.0;
    if (force3) {
        force3[0] = v[0];
        force3[1] = v[1];
        force3[2] = v[2];
    }
    if (drag) *drag = v[0] * cos(a) + v[1] * sin(a);
    if (lift) *lift = -v[0] * sin(a) + v[1] * cos(a);
    if (pint) *pint = v[3];
    return CFD_OK;
}

// The Jacobian of one Newton step: first-order flux, V/dt on the diagonal,
// B = -(Conv + Diss); approximate Roe or finite differences of the scheme's
// own flux (o->jacobian).
int solve_jacobian(cfd_mesh *m, const cfd_params &p, const cfd_solve_opts *o, int it) {
    if (o->jacobian == CFD_SOLVE_JAC_FD) {
        cfd_params p1 = p;
        p1.order = 1;
        return cfd_jacobian_fd(m, &p1, CFD_JAC_FD_CENTRAL, it, 1, 1, o->lin.mixed);
    }
    return cfd_jacobian(m, &p, 1, o->lin.mixed);
}

extern "C" int cfd_solve_steady(cfd_mesh *m, const cfd_params *p0, const cfd_solve_opts *o, cfd_solve_result *res,
                                double *rms_hist) {
    int rc = check_params(m, p0);
    if (rc) return rc;
    if (!o || !res) return CFD_E_ARG;
    if (o->jacobian != CFD_SOLVE_JAC_APPROX && o->jacobian != CFD_SOLVE_JAC_FD)
        return cfd_set_error(m->ctx, CFD_E_ARG, "cfd_solve_steady: unknown Jacobian %d", o->jacobian);
    const auto t0 = std::chrono::steady_clock::now();
    *res = cfd_solve_result{};
    cfd_params p = *p0;
    p.precision = CFD_FP64;  // the nonlinear residual is always FP64
    for (int it = 0; it < o->max_iter; it++) {
        cfd_params pi = p;
        if (pi.order == 2 && it < o->first_order_iters) pi.order = 1;
        rc = cfd_bc(m, &pi, it);
        if (rc) break;
        if (pi.order == 2) {
            rc = pi.limiter != CFD_LIM_NONE ? cfd_gradient_limiter(m, &pi) : cfd_gradient(m, &pi);
            if (rc) break;
        }
        rc = cfd_residual(m, &pi, 1, nullptr);
        if (rc) break;
        double rms = 0.0;
        rc = cfd_residual_rms(m, &rms);
        if (rc) break;
        if (rms_hist) rms_hist[it] = rms;
        if (it == 0) res->first_rms = rms;
        res->final_rms = rms;
        res->iterations = it;
        if (!(rms == rms) || std::isinf(rms)) {
            rc = cfd_set_error(m->ctx, CFD_E_NONFINITE, "non-finite residual at iteration %d", it);
            break;
        }
        if (rms < o->tol) {
            res->converged = 1;
            break;
        }
        rc = cfd_delta_t(m, &pi, it, nullptr);
        if (!rc) rc = solve_jacobian(m, pi, o, it);
        if (rc) break;
        cfd_linsolve_stats st{};
        if (o->autotune && !m->tuned) {
            cfd_tune_report tr;
            rc = autotune_mesh(m, &pi, &o->lin, &tr, [&] { return run_linsolve(m, &o->lin, &st); });
            if (rc) break;
            m->tuned = true;
        }
        rc = cfd_linsolve(m, &o->lin, &st);
        if (rc) break;
        res->krylov_total += st.iterations;
        rc = cfd_update(m, o->relaxation);
        if (rc) break;
        res->iterations = it + 1;
        if (cfd_progress())
            fprintf(stderr, "cfd_solve_steady: iteration %d rms %.3e krylov %d (%d in all) %.2f s\n", it, rms,
                    st.iterations, res->krylov_total,
                    std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    }
    res->status = rc;
    if (!rc) {
        int rc2 = cfd_forces(m, &p, o->wall_tag, o->alpha_deg, res->force, &res->lift, &res->drag,
                             &res->pressure_integral);
        if (rc2) rc = rc2;
    }
    hipStreamSynchronize(m->stream);
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
