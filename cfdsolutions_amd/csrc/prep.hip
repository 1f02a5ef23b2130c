// Median-dual preprocessing on the GPU: from nodes, boundary triangles and
// tetrahedra to the arrays the residual path consumes (node->node lists,
// interior and boundary edges, edge area vectors, control volumes), bit for
// bit those of the host library's cfdm_preprocess and of the reference:
//   - node->node lists in first-appearance order over each node's cells,
//     Euler3D/src/Connectivity_Maps.cpp:44-340;
//   - interior edges (L < R, node-major), Connectivity_Maps.cpp:733-762;
//   - surface orientation towards the domain, Connectivity_Maps.cpp:782-910;
//   - boundary edges and area vectors, Euler3D/src/Area_Volume.cpp:462-572;
//   - interior edge areas and dual volumes, Area_Volume.cpp:660-821;
//   - outward boundary normals, Area_Volume.cpp:913-916.
//
// The reference accumulates the edge areas and volumes in one sequential
// sweep over the tetrahedra.  Here every edge and every node owns its sum:
// it gathers its contributions in the sweep's order (ascending tetrahedron,
// then local edge, then the two sub-faces), so the sums are the same bits
// with no atomics on floating-point values.  Integer lists built with atomic
// slots are sorted per owner before use, so every result is deterministic.
#include "cfd_internal.h"

#include <cstring>
#include <memory>

#include <hipcub/hipcub.hpp>

namespace {

constexpr int kPrepBlock = 256;
constexpr int kMaxNbr = 64;  // neighbours per node (as the host library)

// Local edges of a tetrahedron and the corner nodes of the two sub-faces each
// edge's dual face is made of (the constant table of Area_Volume.cpp:42-103).
__constant__ int cTetEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int cTetFaceL[6][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}, {1, 3, 2}, {0, 2, 3}};
__constant__ int cTetFaceR[6][3] = {{0, 3, 1}, {0, 1, 2}, {0, 2, 3}, {1, 3, 2}, {0, 3, 1}, {1, 3, 2}};

enum PrepError : int { kErrNbr = 1, kErrTri = 2, kErrEdge = 4 };

struct P3 {
    double v[3];
};

// Vector3D::cross, MATH/src/Vector3D.cpp:200-206 (same operation order).
__device__ __forceinline__ P3 cross3(const P3 &a, const P3 &b) {
    P3 t;
    t.v[0] = +a.v[1] * b.v[2] - a.v[2] * b.v[1];
    t.v[1] = -a.v[0] * b.v[2] + a.v[2] * b.v[0];
    t.v[2] = +a.v[0] * b.v[1] - a.v[1] * b.v[0];
    return t;
}
__device__ __forceinline__ double dot3(const P3 &a, const P3 &b) {
    return (a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]);
}
__device__ __forceinline__ void normalize3(P3 &a) {
    const double mag = sqrt(a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]);
    if (mag > 1.0e-20) {
        a.v[0] = a.v[0] / mag;
        a.v[1] = a.v[1] / mag;
        a.v[2] = a.v[2] / mag;
    }
}

__device__ __forceinline__ int gtid() { return blockIdx.x * kPrepBlock + threadIdx.x; }

// UGRID winding translation {2,1,0} (MeshIO.cpp:22-44).
__global__ void k_translate(int NT, const int *tri, int *ftri) {
    const int c = gtid();
    if (c >= NT) return;
    ftri[3 * (size_t)c + 2] = tri[3 * (size_t)c + 0];
    ftri[3 * (size_t)c + 1] = tri[3 * (size_t)c + 1];
    ftri[3 * (size_t)c + 0] = tri[3 * (size_t)c + 2];
}

// node -> cell lists; cell ids: triangles 0..NT-1, then tetrahedra NT + c
__global__ void k_count_cells(int NT, int NC, const int *ftri, const int *tet, int *cnt) {
    const int c = gtid();
    if (c < NT) {
        for (int a = 0; a < 3; a++) atomicAdd(cnt + ftri[3 * (size_t)c + a], 1);
    } else if (c < NT + NC) {
        for (int a = 0; a < 4; a++) atomicAdd(cnt + tet[4 * (size_t)(c - NT) + a], 1);
    }
}

__global__ void k_fill_cells(int NT, int NC, const int *ftri, const int *tet, int *pos, int *list) {
    const int c = gtid();
    if (c < NT) {
        for (int a = 0; a < 3; a++) list[atomicAdd(pos + ftri[3 * (size_t)c + a], 1)] = c;
    } else if (c < NT + NC) {
        for (int a = 0; a < 4; a++) list[atomicAdd(pos + tet[4 * (size_t)(c - NT) + a], 1)] = c;
    }
}

// ascending order within each segment (insertion sort: segments are short)
__global__ void k_sort_segments(int n, const int *ptr, int *list) {
    const int i = gtid();
    if (i >= n) return;
    const int b = ptr[i], e = ptr[i + 1];
    for (int k = b + 1; k < e; k++) {
        const int v = list[k];
        int q = k - 1;
        while (q >= b && list[q] > v) {
            list[q + 1] = list[q];
            q--;
        }
        list[q + 1] = v;
    }
}

// Node->node list of one node in first-appearance order over its cells
// (ascending cell id), with the reference's vertex order per cell.  COUNT:
// writes the list length; otherwise writes the list at out.
template <bool COUNT>
__global__ __launch_bounds__(64) void k_n2n(int N, int NT, const int *c_ptr, const int *c_list, const int *ftri,
                                            const int *tet, int *cnt, const int *ia, int *ja, int *err) {
    __shared__ int nb[kMaxNbr][64];
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= N) return;
    const int t = threadIdx.x;
    int m = 0;
    bool ok = true;
    auto add = [&](int v) {
        for (int q = 0; q < m; q++)
            if (nb[q][t] == v) return;
        if (m >= kMaxNbr) {
            ok = false;
            return;
        }
        nb[m++][t] = v;
    };
    for (int q = c_ptr[n]; q < c_ptr[n + 1] && ok; q++) {
        const int g = c_list[q];
        if (g < NT) {
            const int *s = ftri + 3 * (size_t)g;
            if (n == s[0]) {
                add(s[1]);
                add(s[2]);
            } else if (n == s[1]) {
                add(s[0]);
                add(s[2]);
            } else {
                add(s[0]);
                add(s[1]);
            }
        } else {
            const int *s = tet + 4 * (size_t)(g - NT);
            if (n == s[0]) {
                add(s[1]);
                add(s[2]);
                add(s[3]);
            } else if (n == s[1]) {
                add(s[0]);
                add(s[2]);
                add(s[3]);
            } else if (n == s[2]) {
                add(s[1]);
                add(s[0]);
                add(s[3]);
            } else {
                add(s[0]);
                add(s[1]);
                add(s[2]);
            }
        }
    }
    if (!ok) {
        atomicOr(err, kErrNbr);
        if (COUNT) cnt[n] = 0;
        return;
    }
    if (COUNT) {
        cnt[n] = m;
    } else {
        int *o = ja + ia[n];
        for (int q = 0; q < m; q++) o[q] = nb[q][t];
    }
}

// interior edges: the neighbours v > n of each node n, in list order
__global__ void k_fwd_count(int N, const int *ia, const int *ja, int *cnt) {
    const int n = gtid();
    if (n >= N) return;
    int k = 0;
    for (int q = ia[n]; q < ia[n + 1]; q++) k += ja[q] > n;
    cnt[n] = k;
}

__global__ void k_edges(int N, const int *ia, const int *ja, const int *fwd, int *edge) {
    const int n = gtid();
    if (n >= N) return;
    int e = fwd[n];
    for (int q = ia[n]; q < ia[n + 1]; q++) {
        if (ja[q] > n) {
            edge[2 * (size_t)e] = n;
            edge[2 * (size_t)e + 1] = ja[q];
            e++;
        }
    }
}

// Surface orientation (Connectivity_Maps.cpp:782-910): the normal of each
// triangle must point into the domain, judged against the centroid of the
// one tetrahedron holding all three of its nodes.
__global__ void k_orient(int NT, const int *c_ptr, const int *c_list, const int *tet, const double *X, int *ftri,
                         int *err) {
    const int c = gtid();
    if (c >= NT) return;
    int *s = ftri + 3 * (size_t)c;
    const int s0 = s[0], s1 = s[1], s2 = s[2];
    int tcell = -1;
    for (int q = c_ptr[s0]; q < c_ptr[s0 + 1] && tcell < 0; q++) {
        const int g = c_list[q];
        if (g < NT) continue;
        const int *t = tet + 4 * (size_t)(g - NT);
        int hit = 0;
        for (int a = 0; a < 4; a++) hit += (t[a] == s1) + (t[a] == s2);
        if (hit == 2) tcell = g - NT;
    }
    if (tcell < 0) {
        atomicOr(err, kErrTri);
        return;
    }
    const int sn[3] = {s0, s1, s2};
    const int *tn = tet + 4 * (size_t)tcell;
    P3 sv1, sv2, scen, vcen, scvc;
    for (int d = 0; d < 3; d++) {
        sv1.v[d] = X[3 * (size_t)s1 + d] - X[3 * (size_t)s0 + d];
        sv2.v[d] = X[3 * (size_t)s2 + d] - X[3 * (size_t)s0 + d];
    }
    P3 snv = cross3(sv1, sv2);
    normalize3(snv);
    for (int d = 0; d < 3; d++) {
        scen.v[d] = 0.0;
        for (int a = 0; a < 3; a++) scen.v[d] += X[3 * (size_t)sn[a] + d];
        scen.v[d] /= 3.0;
        vcen.v[d] = 0.0;
        for (int a = 0; a < 4; a++) vcen.v[d] += X[3 * (size_t)tn[a] + d];
        vcen.v[d] /= 4.0;
        scvc.v[d] = vcen.v[d] - scen.v[d];
    }
    normalize3(scvc);
    if (dot3(scvc, snv) < 0.0) {
        s[0] = s2;
        s[2] = s0;
    }
}

// Boundary edges of each triangle corner and their area vectors, outward
// (Area_Volume.cpp:462-572 and :913-916).
__global__ void k_boundary(int N, int NT, const int *ftri, const int *tag, const double *X, int *bedge, double *bs,
                           int *btag) {
    const int c = gtid();
    if (c >= NT) return;
    const int *s = ftri + 3 * (size_t)c;
    double surcen[3];
    for (int d = 0; d < 3; d++)
        surcen[d] = (X[3 * (size_t)s[0] + d] + X[3 * (size_t)s[1] + d] + X[3 * (size_t)s[2] + d]) / 3.0;
    for (int k = 0; k < 3; k++) {
        const int n0 = s[k], na = s[(k + 1) % 3], nb = s[(k + 2) % 3];
        P3 V1, V2, V3;
        for (int d = 0; d < 3; d++) {
            const double e1 = (X[3 * (size_t)n0 + d] + X[3 * (size_t)na + d]) / 2.0;
            const double e2 = (X[3 * (size_t)n0 + d] + X[3 * (size_t)nb + d]) / 2.0;
            V1.v[d] = e1 - X[3 * (size_t)n0 + d];
            V2.v[d] = surcen[d] - X[3 * (size_t)n0 + d];
            V3.v[d] = e2 - X[3 * (size_t)n0 + d];
        }
        const P3 c1 = cross3(V1, V2), c2 = cross3(V2, V3);
        const size_t b = 3 * (size_t)c + k;
        for (int d = 0; d < 3; d++) {
            double a = 0.0;
            a += 0.5 * (c1.v[d] + c2.v[d]);
            bs[3 * b + d] = -1.0 * a;
        }
        bedge[2 * b] = n0;
        bedge[2 * b + 1] = N + (int)b;
        btag[b] = tag[c];
    }
}

// interior edge of each local edge of each tetrahedron (through the forward
// list of its lower node), counted per edge
__global__ void k_tet_edges(int NC, const int *tet, const int *fwd, const int *edge, int *te2e, int *cnt, int *err) {
    const int c = gtid();
    if (c >= NC) return;
    const int *node = tet + 4 * (size_t)c;
    for (int j = 0; j < 6; j++) {
        const int a = node[cTetEdge[j][0]], b = node[cTetEdge[j][1]];
        const int L = a < b ? a : b, R = a < b ? b : a;
        int e = -1;
        for (int q = fwd[L]; q < fwd[L + 1]; q++)
            if (edge[2 * (size_t)q + 1] == R) {
                e = q;
                break;
            }
        te2e[6 * (size_t)c + j] = e;
        if (e < 0) atomicOr(err, kErrEdge);
        else atomicAdd(cnt + e, 1);
    }
}

__global__ void k_fill_edges(int NC, const int *te2e, int *pos, int *list) {
    const int c = gtid();
    if (c >= NC) return;
    for (int j = 0; j < 6; j++) {
        const int e = te2e[6 * (size_t)c + j];
        if (e >= 0) list[atomicAdd(pos + e, 1)] = 6 * c + j;
    }
}

// The two sub-face cross products of local edge j of tetrahedron `node` (the
// operation order of Area_Volume.cpp:660-821).
struct EdgeFaces {
    P3 cr1, cr2, V3v, V4;
};

__device__ __forceinline__ EdgeFaces edge_faces(const int *node, int j, const double *X) {
    double volcen[3];
    for (int d = 0; d < 3; d++) {
        volcen[d] = 0.0;
        for (int a = 0; a < 4; a++) volcen[d] += X[3 * (size_t)node[a] + d];
        volcen[d] /= 4.0;
    }
    const int e0 = node[cTetEdge[j][0]], e1 = node[cTetEdge[j][1]];
    double fl[3], fr[3], ec[3];
    for (int d = 0; d < 3; d++) {
        fl[d] = 0.0;
        fr[d] = 0.0;
        for (int a = 0; a < 3; a++) {
            fl[d] += X[3 * (size_t)node[cTetFaceL[j][a]] + d];
            fr[d] += X[3 * (size_t)node[cTetFaceR[j][a]] + d];
        }
        fl[d] /= 3.0;
        fr[d] /= 3.0;
        ec[d] = 0.0;
        ec[d] += X[3 * (size_t)e0 + d];
        ec[d] += X[3 * (size_t)e1 + d];
        ec[d] /= 2.0;
    }
    EdgeFaces f;
    P3 V1, V2;
    for (int d = 0; d < 3; d++) {
        V1.v[d] = volcen[d] - ec[d];
        V2.v[d] = fl[d] - ec[d];
        f.V3v.v[d] = ec[d] - X[3 * (size_t)e1 + d];
        f.V4.v[d] = X[3 * (size_t)e0 + d] - ec[d];
    }
    f.cr1 = cross3(V1, V2);
    for (int d = 0; d < 3; d++) {
        V1.v[d] = fr[d] - ec[d];
        V2.v[d] = volcen[d] - ec[d];
    }
    f.cr2 = cross3(V1, V2);
    return f;
}

// Area vector of each interior edge: its tetrahedra's contributions in
// ascending tetrahedron order (each tetrahedron holds the edge once).
__global__ void k_edge_area(int E, const int *e_ptr, const int *e_list, const int *tet, const double *X,
                            double *S) {
    const int e = gtid();
    if (e >= E) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int q = e_ptr[e]; q < e_ptr[e + 1]; q++) {
        const int te = e_list[q], c = te / 6, j = te % 6;
        const int *node = tet + 4 * (size_t)c;
        const EdgeFaces f = edge_faces(node, j, X);
        double av[3];
        for (int d = 0; d < 3; d++) av[d] = 0.5 * f.cr1.v[d];
        for (int d = 0; d < 3; d++) av[d] += 0.5 * f.cr2.v[d];
        if (node[cTetEdge[j][1]] > node[cTetEdge[j][0]])
            for (int d = 0; d < 3; d++) av[d] = -1.0 * av[d];
        for (int d = 0; d < 3; d++) s[d] += av[d];
    }
    for (int d = 0; d < 3; d++) S[3 * (size_t)e + d] = s[d];
}

// Dual volume of each node: for each of its tetrahedra (ascending) and each
// local edge holding it (ascending), the first then the second sub-face term.
__global__ void k_volume(int N, int NT, const int *c_ptr, const int *c_list, const int *tet, const double *X,
                         double *vol) {
    const int n = gtid();
    if (n >= N) return;
    double v = 0.0;
    for (int q = c_ptr[n]; q < c_ptr[n + 1]; q++) {
        const int g = c_list[q];
        if (g < NT) continue;
        const int *node = tet + 4 * (size_t)(g - NT);
        for (int j = 0; j < 6; j++) {
            const bool is1 = node[cTetEdge[j][1]] == n, is0 = node[cTetEdge[j][0]] == n;
            if (!is0 && !is1) continue;
            const EdgeFaces f = edge_faces(node, j, X);
            const P3 &w = is1 ? f.V3v : f.V4;
            v += (1.0 / 6.0) * dot3(w, f.cr1);
            v += (1.0 / 6.0) * dot3(w, f.cr2);
        }
    }
    vol[n] = v;
}

// Device buffers of one preprocessing run, freed together.
struct PrepBufs {
    std::vector<void *> ptrs;
    ~PrepBufs() {
        CfdRuntimeGuard g;
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    T *get(size_t n, hipError_t *e) {
        void *p = nullptr;
        {
            CfdRuntimeGuard g;
            *e = hipMalloc(&p, sizeof(T) * (n ? n : 1));
        }
        if (*e != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T *)p;
    }
};

inline int grid(int64_t n) { return (int)((n + kPrepBlock - 1) / kPrepBlock); }

}  // namespace

struct cfd_prep {
    int32_t n_node = 0, n_tri = 0, n_tet = 0, n_edge = 0, n_bedge = 0, n_n2n = 0;
    std::vector<int32_t> tri, edge, bedge, btag, n2n_ia, n2n_ja;
    std::vector<double> edge_area, bedge_area, vol;
};

extern "C" int cfd_preprocess(cfd_ctx *ctx, int32_t n_node, const double *xyz, int32_t n_tri, const int32_t *tri,
                              const int32_t *tri_tag, int32_t n_tet, const int32_t *tet, cfd_prep **out) {
    if (!ctx || !out || n_node <= 0 || n_tri < 0 || n_tet <= 0 || !xyz || !tet || (n_tri && (!tri || !tri_tag)))
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: invalid arguments");
    *out = nullptr;
    const int N = n_node, NT = n_tri, NC = n_tet;
    if ((int64_t)NT + NC > INT32_MAX / 8 || 3LL * NT > INT32_MAX / 4 || 6LL * NC > INT32_MAX / 2 ||
        4LL * NC + 3LL * NT > INT32_MAX / 2)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: mesh too large for int32 indices");
    for (int64_t k = 0; k < 3LL * NT; k++)
        if (tri[k] < 0 || tri[k] >= N) return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: bad triangle node");
    for (int64_t k = 0; k < 4LL * NC; k++)
        if (tet[k] < 0 || tet[k] >= N) return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: bad tetrahedron node");
    if (hipSetDevice(ctx->device) != hipSuccess) return cfd_set_error(ctx, CFD_E_HIP, "hipSetDevice failed");

    hipStream_t st = nullptr;
    {
        CfdRuntimeGuard g;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
            return cfd_set_error(ctx, CFD_E_HIP, "cfd_preprocess: hipStreamCreate failed");
    }
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() {
            CfdRuntimeGuard g;
            (void)hipStreamDestroy(s);
        }
    } sg{st};
    PrepBufs bufs;
    hipError_t he = hipSuccess;
#define PREP_CHECK(call)                                                                                   \
    do {                                                                                                   \
        he = (call);                                                                                       \
        if (he != hipSuccess)                                                                              \
            return cfd_set_error(ctx, he == hipErrorOutOfMemory ? CFD_E_OOM : CFD_E_HIP,                   \
                                 "cfd_preprocess: %s: %s", #call, hipGetErrorString(he));                 \
    } while (0)
#define PREP_ALLOC(ptr, T, n)                                                                              \
    do {                                                                                                   \
        ptr = bufs.get<T>((size_t)(n), &he);                                                               \
        if (!ptr) return cfd_set_error(ctx, CFD_E_OOM, "cfd_preprocess: out of device memory");           \
    } while (0)

    // exclusive prefix sum of cnt[0..n) into ptr[0..n] (ptr[0] = 0)
    size_t scan_bytes = 0;
    void *scan_tmp = nullptr;
    auto scan = [&](const int *cnt, int *ptr, int n) -> hipError_t {
        size_t need = 0;
        hipError_t e = hipcub::DeviceScan::InclusiveSum(nullptr, need, cnt, ptr + 1, n, st);
        if (e != hipSuccess) return e;
        if (need > scan_bytes) {
            scan_tmp = bufs.get<char>(need, &e);
            if (!scan_tmp) return e;
            scan_bytes = need;
        }
        e = hipMemsetAsync(ptr, 0, sizeof(int), st);
        if (e != hipSuccess) return e;
        return hipcub::DeviceScan::InclusiveSum(scan_tmp, scan_bytes, cnt, ptr + 1, n, st);
    };
    auto to_host = [&](void *dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) return e;
        CfdRuntimeGuard g;
        return hipStreamSynchronize(st);
    };

    double *dX;
    int *dtri, *dftri, *dtag, *dtet, *derr;
    PREP_ALLOC(dX, double, 3 * (size_t)N);
    PREP_ALLOC(dtri, int, 3 * (size_t)NT);
    PREP_ALLOC(dftri, int, 3 * (size_t)NT);
    PREP_ALLOC(dtag, int, NT);
    PREP_ALLOC(dtet, int, 4 * (size_t)NC);
    PREP_ALLOC(derr, int, 1);
    PREP_CHECK(hipMemcpyAsync(dX, xyz, sizeof(double) * 3 * (size_t)N, hipMemcpyHostToDevice, st));
    if (NT) {
        PREP_CHECK(hipMemcpyAsync(dtri, tri, sizeof(int) * 3 * (size_t)NT, hipMemcpyHostToDevice, st));
        PREP_CHECK(hipMemcpyAsync(dtag, tri_tag, sizeof(int) * (size_t)NT, hipMemcpyHostToDevice, st));
    }
    PREP_CHECK(hipMemcpyAsync(dtet, tet, sizeof(int) * 4 * (size_t)NC, hipMemcpyHostToDevice, st));
    PREP_CHECK(hipMemsetAsync(derr, 0, sizeof(int), st));
    if (NT) hipLaunchKernelGGL(k_translate, dim3(grid(NT)), dim3(kPrepBlock), 0, st, NT, dtri, dftri);

    // node -> cell lists, ascending cell id
    int *ccnt, *cptr, *cpos, *clist;
    const int64_t ncell = 3LL * NT + 4LL * NC;
    PREP_ALLOC(ccnt, int, N);
    PREP_ALLOC(cptr, int, N + 1);
    PREP_ALLOC(cpos, int, N + 1);
    PREP_ALLOC(clist, int, ncell);
    PREP_CHECK(hipMemsetAsync(ccnt, 0, sizeof(int) * (size_t)N, st));
    hipLaunchKernelGGL(k_count_cells, dim3(grid(NT + NC)), dim3(kPrepBlock), 0, st, NT, NC, dftri, dtet, ccnt);
    PREP_CHECK(scan(ccnt, cptr, N));
    PREP_CHECK(hipMemcpyAsync(cpos, cptr, sizeof(int) * (size_t)(N + 1), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_fill_cells, dim3(grid(NT + NC)), dim3(kPrepBlock), 0, st, NT, NC, dftri, dtet, cpos, clist);
    hipLaunchKernelGGL(k_sort_segments, dim3(grid(N)), dim3(kPrepBlock), 0, st, N, cptr, clist);

    // node -> node lists
    int *ncnt, *nia, *nja;
    PREP_ALLOC(ncnt, int, N);
    PREP_ALLOC(nia, int, N + 1);
    hipLaunchKernelGGL(k_n2n<true>, dim3((N + 63) / 64), dim3(64), 0, st, N, NT, cptr, clist, dftri, dtet, ncnt,
                       nullptr, nullptr, derr);
    PREP_CHECK(scan(ncnt, nia, N));
    int nnz = 0, err = 0;
    PREP_CHECK(to_host(&nnz, nia + N, sizeof(int)));
    PREP_CHECK(to_host(&err, derr, sizeof(int)));
    if (err & kErrNbr) return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: a node has more than 64 neighbours");
    PREP_ALLOC(nja, int, nnz);
    hipLaunchKernelGGL(k_n2n<false>, dim3((N + 63) / 64), dim3(64), 0, st, N, NT, cptr, clist, dftri, dtet, nullptr,
                       nia, nja, derr);

    // interior edges
    int *fcnt, *fwd, *dedge;
    PREP_ALLOC(fcnt, int, N);
    PREP_ALLOC(fwd, int, N + 1);
    hipLaunchKernelGGL(k_fwd_count, dim3(grid(N)), dim3(kPrepBlock), 0, st, N, nia, nja, fcnt);
    PREP_CHECK(scan(fcnt, fwd, N));
    int E = 0;
    PREP_CHECK(to_host(&E, fwd + N, sizeof(int)));
    PREP_ALLOC(dedge, int, 2 * (size_t)E);
    hipLaunchKernelGGL(k_edges, dim3(grid(N)), dim3(kPrepBlock), 0, st, N, nia, nja, fwd, dedge);

    // surface orientation, boundary edges and areas
    const int B = 3 * NT;
    int *dbedge, *dbtag;
    double *dbs;
    PREP_ALLOC(dbedge, int, 2 * (size_t)B);
    PREP_ALLOC(dbtag, int, B);
    PREP_ALLOC(dbs, double, 3 * (size_t)B);
    if (NT) {
        hipLaunchKernelGGL(k_orient, dim3(grid(NT)), dim3(kPrepBlock), 0, st, NT, cptr, clist, dtet, dX, dftri, derr);
        hipLaunchKernelGGL(k_boundary, dim3(grid(NT)), dim3(kPrepBlock), 0, st, N, NT, dftri, dtag, dX, dbedge, dbs,
                           dbtag);
    }

    // edge -> (tetrahedron, local edge) lists, then areas and volumes
    int *te2e, *ecnt, *eptr, *epos, *elist;
    double *dS, *dvol;
    PREP_ALLOC(te2e, int, 6 * (size_t)NC);
    PREP_ALLOC(ecnt, int, E);
    PREP_ALLOC(eptr, int, E + 1);
    PREP_ALLOC(epos, int, E + 1);
    PREP_ALLOC(elist, int, 6 * (size_t)NC);
    PREP_ALLOC(dS, double, 3 * (size_t)E);
    PREP_ALLOC(dvol, double, N);
    PREP_CHECK(hipMemsetAsync(ecnt, 0, sizeof(int) * (size_t)(E ? E : 1), st));
    hipLaunchKernelGGL(k_tet_edges, dim3(grid(NC)), dim3(kPrepBlock), 0, st, NC, dtet, fwd, dedge, te2e, ecnt, derr);
    PREP_CHECK(scan(ecnt, eptr, E));
    PREP_CHECK(hipMemcpyAsync(epos, eptr, sizeof(int) * (size_t)(E + 1), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_fill_edges, dim3(grid(NC)), dim3(kPrepBlock), 0, st, NC, te2e, epos, elist);
    hipLaunchKernelGGL(k_sort_segments, dim3(grid(E)), dim3(kPrepBlock), 0, st, E, eptr, elist);
    hipLaunchKernelGGL(k_edge_area, dim3(grid(E)), dim3(kPrepBlock), 0, st, E, eptr, elist, dtet, dX, dS);
    hipLaunchKernelGGL(k_volume, dim3(grid(N)), dim3(kPrepBlock), 0, st, N, NT, cptr, clist, dtet, dX, dvol);
    PREP_CHECK(hipGetLastError());
    PREP_CHECK(to_host(&err, derr, sizeof(int)));
    if (err & kErrTri) return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: a triangle has no tetrahedron");
    if (err & kErrEdge) return cfd_set_error(ctx, CFD_E_ARG, "cfd_preprocess: a tetrahedron edge is not in the graph");

    cfd_prep *p = new (std::nothrow) cfd_prep();
    if (!p) return CFD_E_OOM;
    std::unique_ptr<cfd_prep> own(p);
    try {
        p->n_node = N;
        p->n_tri = NT;
        p->n_tet = NC;
        p->n_edge = E;
        p->n_bedge = B;
        p->n_n2n = nnz;
        p->tri.resize(3 * (size_t)NT);
        p->edge.resize(2 * (size_t)E);
        p->bedge.resize(2 * (size_t)B);
        p->btag.resize(B);
        p->n2n_ia.resize(N + 1);
        p->n2n_ja.resize(nnz);
        p->edge_area.resize(3 * (size_t)E);
        p->bedge_area.resize(3 * (size_t)B);
        p->vol.resize(N);
    } catch (const std::bad_alloc &) {
        return cfd_set_error(ctx, CFD_E_OOM, "cfd_preprocess: out of host memory");
    }
    auto get = [&](auto &v, const void *src) -> hipError_t {
        if (v.empty()) return hipSuccess;
        return hipMemcpyAsync(v.data(), src, sizeof(v[0]) * v.size(), hipMemcpyDeviceToHost, st);
    };
    PREP_CHECK(get(p->tri, dftri));
    PREP_CHECK(get(p->edge, dedge));
    PREP_CHECK(get(p->bedge, dbedge));
    PREP_CHECK(get(p->btag, dbtag));
    PREP_CHECK(get(p->n2n_ia, nia));
    PREP_CHECK(get(p->n2n_ja, nja));
    PREP_CHECK(get(p->edge_area, dS));
    PREP_CHECK(get(p->bedge_area, dbs));
    PREP_CHECK(get(p->vol, dvol));
    {
        CfdRuntimeGuard g;
        PREP_CHECK(hipStreamSynchronize(st));
    }
#undef PREP_CHECK
#undef PREP_ALLOC
    *out = own.release();
    return CFD_OK;
}

extern "C" int cfd_prep_info(const cfd_prep *p, cfd_prep_counts *c) {
    if (!p || !c) return CFD_E_ARG;
    c->n_node = p->n_node;
    c->n_tri = p->n_tri;
    c->n_tet = p->n_tet;
    c->n_edge = p->n_edge;
    c->n_bedge = p->n_bedge;
    c->n_n2n = p->n_n2n;
    return CFD_OK;
}

extern "C" int cfd_prep_get(const cfd_prep *p, int32_t *tri, int32_t *edge, double *edge_area, int32_t *bedge,
                            double *bedge_area, int32_t *bedge_tag, double *vol, int32_t *n2n_ia, int32_t *n2n_ja) {
    if (!p) return CFD_E_ARG;
    auto cp = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(tri, p->tri);
    cp(edge, p->edge);
    cp(edge_area, p->edge_area);
    cp(bedge, p->bedge);
    cp(bedge_area, p->bedge_area);
    cp(bedge_tag, p->btag);
    cp(vol, p->vol);
    cp(n2n_ia, p->n2n_ia);
    cp(n2n_ja, p->n2n_ja);
    return CFD_OK;
}

extern "C" void cfd_prep_destroy(cfd_prep *p) { delete p; }
