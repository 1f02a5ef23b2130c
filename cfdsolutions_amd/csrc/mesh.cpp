// Host-side mesh generation and median-dual preprocessing.
//
// The connectivity and metric routines below reproduce the reference's
// preprocessing (Euler3D/src/MeshIO.cpp, Euler3D/src/Connectivity_Maps.cpp,
// Euler3D/src/Area_Volume.cpp) in O(N+E) with the same floating-point
// operation order, so edge areas and dual volumes are bit-identical to the
// reference's for the same UGRID input.  Storage is flat arrays instead of the
// reference's per-node List objects.
#include "cfd_mesh.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

struct cfdm_mesh {
    int nNode = 0, nTri = 0, nTet = 0;
    std::vector<double> xyz;   // 3N
    std::vector<int> tri;      // 3*nTri, node order as in the UGRID file
    std::vector<int> triTag;   // nTri
    std::vector<int> tet;      // 4*nTet
    // preprocessed
    bool ready = false;
    std::vector<int> ftri;      // triangles after winding translation + orientation fix
    std::vector<int> n2nIA, n2nJA;
    std::vector<int> edge;      // 2E
    std::vector<double> edgeS;  // 3E
    std::vector<int> bedge;     // 2B
    std::vector<double> bedgeS; // 3B
    std::vector<int> btag;      // B
    std::vector<double> vol;    // N
};

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *cfdm_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Counter-based random numbers (SplitMix64), shared with the state generators
// ---------------------------------------------------------------------------
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

// Uniform in [0,1) from (seed, index, stream).
static inline double rng_u01(uint64_t seed, uint64_t idx, uint64_t stream) {
    uint64_t r = splitmix64(seed ^ splitmix64(idx * 16ULL + stream));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee on the node graph of the tetrahedra
// ---------------------------------------------------------------------------
static void build_node_graph(int n, const std::vector<int> &tet, std::vector<int> &ia,
                             std::vector<int> &ja) {
    // Edge pairs of all tets, deduplicated per node by sort.
    const int nt = (int)(tet.size() / 4);
    std::vector<int> cnt(n + 1, 0);
    for (int t = 0; t < nt; t++)
        for (int a = 0; a < 4; a++) cnt[tet[4 * t + a] + 1] += 3;
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    std::vector<int> tmp(cnt[n]);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int t = 0; t < nt; t++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                if (a != b) tmp[pos[tet[4 * t + a]]++] = tet[4 * t + b];
    ia.assign(n + 1, 0);
    ja.clear();
    ja.reserve(cnt[n] / 2);
    for (int i = 0; i < n; i++) {
        auto b = tmp.begin() + cnt[i], e = tmp.begin() + cnt[i + 1];
        std::sort(b, e);
        auto u = std::unique(b, e);
        ja.insert(ja.end(), b, u);
        ia[i + 1] = (int)ja.size();
    }
}

static std::vector<int> rcm_order(int n, const std::vector<int> &ia, const std::vector<int> &ja) {
    std::vector<int> deg(n);
    for (int i = 0; i < n; i++) deg[i] = ia[i + 1] - ia[i];
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> level(n, -1);
    std::vector<char> placed(n, 0);
    std::vector<int> nb;

    auto bfs_levels = [&](int root, int &last_level_min_deg_node) {
        std::vector<int> q;
        q.push_back(root);
        std::fill(level.begin(), level.end(), -1);
        level[root] = 0;
        size_t h = 0;
        int maxl = 0;
        while (h < q.size()) {
            int u = q[h++];
            for (int k = ia[u]; k < ia[u + 1]; k++) {
                int v = ja[k];
                if (level[v] < 0) {
                    level[v] = level[u] + 1;
                    maxl = std::max(maxl, level[v]);
                    q.push_back(v);
                }
            }
        }
        int best = -1;
        for (int u : q)
            if (level[u] == maxl && (best < 0 || deg[u] < deg[best])) best = u;
        last_level_min_deg_node = best;
        return maxl;
    };

    for (int seed = 0; seed < n; seed++) {
        if (placed[seed]) continue;
        // pseudo-peripheral start node (George-Liu)
        int start = seed;
        int ecc = bfs_levels(start, start);
        for (int it = 0; it < 8; it++) {
            int cand;
            int e2 = bfs_levels(start, cand);
            if (e2 <= ecc && it > 0) break;
            ecc = e2;
            start = cand;
        }
        size_t h = order.size();
        order.push_back(start);
        placed[start] = 1;
        while (h < order.size()) {
            int u = order[h++];
            nb.clear();
            for (int k = ia[u]; k < ia[u + 1]; k++)
                if (!placed[ja[k]]) nb.push_back(ja[k]);
            std::sort(nb.begin(), nb.end(), [&](int a, int b) {
                return deg[a] != deg[b] ? deg[a] < deg[b] : a < b;
            });
            for (int v : nb) {
                placed[v] = 1;
                order.push_back(v);
            }
        }
    }
    std::reverse(order.begin(), order.end());
    return order;  // order[new] = old
}

static void renumber(cfdm_mesh &m, const std::vector<int> &order) {
    const int n = m.nNode;
    std::vector<int> newid(n);
    for (int k = 0; k < n; k++) newid[order[k]] = k;
    std::vector<double> x(3 * (size_t)n);
    for (int k = 0; k < n; k++)
        for (int d = 0; d < 3; d++) x[3 * (size_t)k + d] = m.xyz[3 * (size_t)order[k] + d];
    m.xyz.swap(x);
    for (auto &v : m.tet) v = newid[v];
    for (auto &v : m.tri) v = newid[v];
}

// ---------------------------------------------------------------------------
// Box / bump-channel generator
// ---------------------------------------------------------------------------
extern "C" int cfdm_box_generate(int nx, int ny, int nz, double lx, double ly, double lz,
                                 double bump_h, double jitter, uint64_t seed, int reorder,
                                 cfdm_mesh **out) {
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0) || jitter < 0 ||
        jitter >= 0.5)
        return fail(CFDM_E_ARG, "cfdm_box_generate: invalid arguments");
    const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t ntet = 6LL * nx * ny * nz;
    if (nn > INT32_MAX / 4 || ntet > INT32_MAX / 4)
        return fail(CFDM_E_ARG, "cfdm_box_generate: mesh too large for int32 indices");
    cfdm_mesh *m = new (std::nothrow) cfdm_mesh();
    if (!m) return fail(CFDM_E_OOM, "cfdm_box_generate: out of memory");
    try {
        m->nNode = (int)nn;
        m->xyz.resize(3 * (size_t)nn);
        auto nid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
        const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
        double R = 0.0;
        if (bump_h > 0) R = (0.25 + bump_h * bump_h) / (2.0 * bump_h);
        for (int k = 0; k <= nz; k++)
            for (int j = 0; j <= ny; j++)
                for (int i = 0; i <= nx; i++) {
                    const int id = nid(i, j, k);
                    double x = lx * (double)i / (double)nx;
                    double eta = (double)j / (double)ny;
                    double yb = 0.0;
                    if (bump_h > 0 && x >= 1.0 && x <= 2.0)
                        yb = std::sqrt(R * R - (x - 1.5) * (x - 1.5)) - (R - bump_h);
                    double y = yb + eta * (ly - yb);
                    double z = lz * (double)k / (double)nz;
                    if (jitter > 0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
                        x += hx * jitter * (2.0 * rng_u01(seed, (uint64_t)id, 0) - 1.0);
                        y += hy * jitter * (2.0 * rng_u01(seed, (uint64_t)id, 1) - 1.0);
                        z += hz * jitter * (2.0 * rng_u01(seed, (uint64_t)id, 2) - 1.0);
                    }
                    m->xyz[3 * (size_t)id + 0] = x;
                    m->xyz[3 * (size_t)id + 1] = y;
                    m->xyz[3 * (size_t)id + 2] = z;
                }
        // Kuhn split: the 6 axis permutations walk from corner (0,0,0) to (1,1,1).
        static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
        static const int odd[6] = {0, 1, 1, 0, 0, 1};
        m->nTet = (int)ntet;
        m->tet.resize(4 * (size_t)ntet);
        size_t t = 0;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++)
                    for (int p = 0; p < 6; p++) {
                        int c[3] = {0, 0, 0};
                        int v[4];
                        v[0] = nid(i, j, k);
                        c[perm[p][0]] = 1;
                        v[1] = nid(i + c[0], j + c[1], k + c[2]);
                        c[perm[p][1]] = 1;
                        v[2] = nid(i + c[0], j + c[1], k + c[2]);
                        v[3] = nid(i + 1, j + 1, k + 1);
                        if (odd[p]) std::swap(v[2], v[3]);
                        for (int a = 0; a < 4; a++) m->tet[4 * t + a] = v[a];
                        t++;
                    }
        // Boundary triangles: tet faces whose three nodes share a boundary plane.
        static const int face[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
        auto ijk = [&](int id, int &i, int &j, int &k) {
            i = id % (nx + 1);
            j = (id / (nx + 1)) % (ny + 1);
            k = id / ((nx + 1) * (ny + 1));
        };
        for (size_t tt = 0; tt < (size_t)ntet; tt++) {
            for (int f = 0; f < 4; f++) {
                int a = m->tet[4 * tt + face[f][0]], b = m->tet[4 * tt + face[f][1]],
                    c = m->tet[4 * tt + face[f][2]];
                int ia_, ja_, ka_, ib, jb, kb, ic, jc, kc;
                ijk(a, ia_, ja_, ka_);
                ijk(b, ib, jb, kb);
                ijk(c, ic, jc, kc);
                int tag = 0;
                if (ia_ == 0 && ib == 0 && ic == 0) tag = CFDM_TAG_XMIN;
                else if (ia_ == nx && ib == nx && ic == nx) tag = CFDM_TAG_XMAX;
                else if (ja_ == 0 && jb == 0 && jc == 0) tag = CFDM_TAG_YMIN;
                else if (ja_ == ny && jb == ny && jc == ny) tag = CFDM_TAG_YMAX;
                else if (ka_ == 0 && kb == 0 && kc == 0) tag = CFDM_TAG_ZMIN;
                else if (ka_ == nz && kb == nz && kc == nz) tag = CFDM_TAG_ZMAX;
                if (tag) {
                    m->tri.push_back(a);
                    m->tri.push_back(b);
                    m->tri.push_back(c);
                    m->triTag.push_back(tag);
                }
            }
        }
        m->nTri = (int)m->triTag.size();
        if (reorder) {
            std::vector<int> ia, ja;
            build_node_graph(m->nNode, m->tet, ia, ja);
            renumber(*m, rcm_order(m->nNode, ia, ja));
        }
    } catch (const std::bad_alloc &) {
        delete m;
        return fail(CFDM_E_OOM, "cfdm_box_generate: out of memory");
    }
    *out = m;
    return CFDM_OK;
}

// ---------------------------------------------------------------------------
// UGRID I/O (format of Euler3D/src/MeshIO.cpp:50-218; triangles, tets only)
// ---------------------------------------------------------------------------
extern "C" int cfdm_write_ugrid(const cfdm_mesh *m, const char *path) {
    if (!m || !path) return fail(CFDM_E_ARG, "cfdm_write_ugrid: null argument");
    FILE *fp = fopen(path, "w");
    if (!fp) return fail(CFDM_E_IO, "cfdm_write_ugrid: cannot open %s", path);
    fprintf(fp, "%d %d %d %d %d %d %d\n", m->nNode, m->nTri, 0, m->nTet, 0, 0, 0);
    for (int n = 0; n < m->nNode; n++)
        fprintf(fp, "%.17g %.17g %.17g\n", m->xyz[3 * (size_t)n], m->xyz[3 * (size_t)n + 1],
                m->xyz[3 * (size_t)n + 2]);
    for (int c = 0; c < m->nTri; c++)
        fprintf(fp, "%d %d %d\n", m->tri[3 * (size_t)c] + 1, m->tri[3 * (size_t)c + 1] + 1,
                m->tri[3 * (size_t)c + 2] + 1);
    for (int c = 0; c < m->nTri; c++) fprintf(fp, "%d\n", m->triTag[c]);
    for (int c = 0; c < m->nTet; c++)
        fprintf(fp, "%d %d %d %d\n", m->tet[4 * (size_t)c] + 1, m->tet[4 * (size_t)c + 1] + 1,
                m->tet[4 * (size_t)c + 2] + 1, m->tet[4 * (size_t)c + 3] + 1);
    int rc = ferror(fp) ? CFDM_E_IO : CFDM_OK;
    fclose(fp);
    if (rc) return fail(rc, "cfdm_write_ugrid: write error on %s", path);
    return CFDM_OK;
}

extern "C" int cfdm_read_ugrid(const char *path, cfdm_mesh **out) {
    if (!path || !out) return fail(CFDM_E_ARG, "cfdm_read_ugrid: null argument");
    FILE *fp = fopen(path, "r");
    if (!fp) return fail(CFDM_E_IO, "cfdm_read_ugrid: cannot open %s", path);
    int cnt[7];
    if (fscanf(fp, "%d %d %d %d %d %d %d", &cnt[0], &cnt[1], &cnt[2], &cnt[3], &cnt[4], &cnt[5],
               &cnt[6]) != 7) {
        fclose(fp);
        return fail(CFDM_E_IO, "cfdm_read_ugrid: bad header in %s", path);
    }
    if (cnt[2] || cnt[4] || cnt[5] || cnt[6] || cnt[0] <= 0 || cnt[1] < 0 || cnt[3] <= 0) {
        fclose(fp);
        return fail(CFDM_E_MESH, "cfdm_read_ugrid: only triangle/tetrahedron meshes are supported");
    }
    cfdm_mesh *m = new (std::nothrow) cfdm_mesh();
    if (!m) {
        fclose(fp);
        return fail(CFDM_E_OOM, "cfdm_read_ugrid: out of memory");
    }
    bool ok = true;
    try {
        m->nNode = cnt[0];
        m->nTri = cnt[1];
        m->nTet = cnt[3];
        m->xyz.resize(3 * (size_t)m->nNode);
        m->tri.resize(3 * (size_t)m->nTri);
        m->triTag.resize(m->nTri);
        m->tet.resize(4 * (size_t)m->nTet);
        for (size_t i = 0; ok && i < m->xyz.size(); i++) ok = fscanf(fp, "%lg", &m->xyz[i]) == 1;
        for (size_t i = 0; ok && i < m->tri.size(); i++) {
            ok = fscanf(fp, "%d", &m->tri[i]) == 1;
            m->tri[i]--;
        }
        for (size_t i = 0; ok && i < m->triTag.size(); i++) ok = fscanf(fp, "%d", &m->triTag[i]) == 1;
        for (size_t i = 0; ok && i < m->tet.size(); i++) {
            ok = fscanf(fp, "%d", &m->tet[i]) == 1;
            m->tet[i]--;
        }
    } catch (const std::bad_alloc &) {
        fclose(fp);
        delete m;
        return fail(CFDM_E_OOM, "cfdm_read_ugrid: out of memory");
    }
    fclose(fp);
    if (!ok) {
        delete m;
        return fail(CFDM_E_IO, "cfdm_read_ugrid: truncated file %s", path);
    }
    for (int v : m->tri)
        if (v < 0 || v >= m->nNode) ok = false;
    for (int v : m->tet)
        if (v < 0 || v >= m->nNode) ok = false;
    if (!ok) {
        delete m;
        return fail(CFDM_E_MESH, "cfdm_read_ugrid: node index out of range in %s", path);
    }
    *out = m;
    return CFDM_OK;
}

extern "C" void cfdm_free(cfdm_mesh *m) { delete m; }

// ---------------------------------------------------------------------------
// Preprocessing
// ---------------------------------------------------------------------------
namespace {

struct V3 {
    double v[3];
};

// Vector3D::cross, MATH/src/Vector3D.cpp:200-206 (same operation order).
static inline V3 cross(const V3 &a, const V3 &b) {
    V3 t;
    t.v[0] = +a.v[1] * b.v[2] - a.v[2] * b.v[1];
    t.v[1] = -a.v[0] * b.v[2] + a.v[2] * b.v[0];
    t.v[2] = +a.v[0] * b.v[1] - a.v[1] * b.v[0];
    return t;
}
static inline double dot(const V3 &a, const V3 &b) {
    return (a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]);
}
static inline void normalize(V3 &a) {
    double mag = std::sqrt(a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]);
    if (mag > 1.0e-20) {
        a.v[0] = a.v[0] / mag;
        a.v[1] = a.v[1] / mag;
        a.v[2] = a.v[2] / mag;
    }
}

// Local edge numbering of a tetrahedron with its left/right faces,
// Euler3D/src/Area_Volume.cpp:42-103.
static const int kTetEdge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int kTetFaceL[6][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}, {1, 3, 2}, {0, 2, 3}};
static const int kTetFaceR[6][3] = {{0, 3, 1}, {0, 1, 2}, {0, 2, 3}, {1, 3, 2}, {0, 3, 1}, {1, 3, 2}};

}  // namespace

extern "C" int cfdm_preprocess(cfdm_mesh *m) {
    if (!m) return fail(CFDM_E_ARG, "cfdm_preprocess: null mesh");
    try {
        const int N = m->nNode, NT = m->nTri, NE4 = m->nTet;
        const double *X = m->xyz.data();
        // UGRID winding translation for triangles {2,1,0}, MeshIO.cpp:22-44.
        m->ftri.resize(3 * (size_t)NT);
        for (int c = 0; c < NT; c++) {
            m->ftri[3 * (size_t)c + 2] = m->tri[3 * (size_t)c + 0];
            m->ftri[3 * (size_t)c + 1] = m->tri[3 * (size_t)c + 1];
            m->ftri[3 * (size_t)c + 0] = m->tri[3 * (size_t)c + 2];
        }
        // node -> cell (global ids: triangles first, then tets), ascending.
        std::vector<int> n2cIA(N + 1, 0);
        for (int v : m->ftri) n2cIA[v + 1]++;
        for (int v : m->tet) n2cIA[v + 1]++;
        for (int i = 0; i < N; i++) n2cIA[i + 1] += n2cIA[i];
        std::vector<int> n2cJA(n2cIA[N]);
        {
            std::vector<int> pos(n2cIA.begin(), n2cIA.end() - 1);
            for (int c = 0; c < NT; c++)
                for (int a = 0; a < 3; a++) n2cJA[pos[m->ftri[3 * (size_t)c + a]]++] = c;
            for (int c = 0; c < NE4; c++)
                for (int a = 0; a < 4; a++) n2cJA[pos[m->tet[4 * (size_t)c + a]]++] = NT + c;
        }
        // node -> node in first-appearance order, Connectivity_Maps.cpp:44-340.
        std::vector<int> n2nIA(N + 1, 0);
        std::vector<int> n2nJA;
        n2nJA.reserve((size_t)N * 15);
        int nb[64];
        for (int n = 0; n < N; n++) {
            int cnt = 0;
            auto add = [&](int v) {
                for (int q = 0; q < cnt; q++)
                    if (nb[q] == v) return true;
                if (cnt >= 64) return false;
                nb[cnt++] = v;
                return true;
            };
            bool ok = true;
            for (int q = n2cIA[n]; q < n2cIA[n + 1]; q++) {
                int g = n2cJA[q];
                if (g < NT) {
                    const int *t = &m->ftri[3 * (size_t)g];
                    int l0, l1;
                    if (n == t[0]) { l0 = t[1]; l1 = t[2]; }
                    else if (n == t[1]) { l0 = t[0]; l1 = t[2]; }
                    else { l0 = t[0]; l1 = t[1]; }
                    ok = ok && add(l0) && add(l1);
                } else {
                    const int *t = &m->tet[4 * (size_t)(g - NT)];
                    int l[3];
                    if (n == t[0]) { l[0] = t[1]; l[1] = t[2]; l[2] = t[3]; }
                    else if (n == t[1]) { l[0] = t[0]; l[1] = t[2]; l[2] = t[3]; }
                    else if (n == t[2]) { l[0] = t[1]; l[1] = t[0]; l[2] = t[3]; }
                    else { l[0] = t[0]; l[1] = t[1]; l[2] = t[2]; }
                    ok = ok && add(l[0]) && add(l[1]) && add(l[2]);
                }
            }
            if (!ok) return fail(CFDM_E_MESH, "cfdm_preprocess: node %d has more than 64 neighbours", n);
            n2nJA.insert(n2nJA.end(), nb, nb + cnt);
            n2nIA[n + 1] = (int)n2nJA.size();
        }
        // Interior edges, node-major, Connectivity_Maps.cpp:733-762.
        std::vector<int> edge;
        edge.reserve(n2nJA.size());
        for (int n = 0; n < N; n++)
            for (int q = n2nIA[n]; q < n2nIA[n + 1]; q++)
                if (n2nJA[q] > n) {
                    edge.push_back(n);
                    edge.push_back(n2nJA[q]);
                }
        const int E = (int)(edge.size() / 2);
        // Position of each (L, neighbour) pair: edge id, searched through L's list.
        std::vector<int> fwdIA(N + 1, 0);
        for (int e = 0; e < E; e++) fwdIA[edge[2 * (size_t)e] + 1]++;
        for (int i = 0; i < N; i++) fwdIA[i + 1] += fwdIA[i];
        auto find_edge = [&](int a, int b) {
            int L = a < b ? a : b, R = a < b ? b : a;
            for (int e = fwdIA[L]; e < fwdIA[L + 1]; e++)
                if (edge[2 * (size_t)e + 1] == R) return e;
            return -1;
        };
        // Surface orientation: normal must point into the domain,
        // Connectivity_Maps.cpp:782-910.
        for (int c = 0; c < NT; c++) {
            int *s = &m->ftri[3 * (size_t)c];
            // the one tetrahedron that holds all three nodes
            int tcell = -1;
            for (int q = n2cIA[s[0]]; q < n2cIA[s[0] + 1] && tcell < 0; q++) {
                int g = n2cJA[q];
                if (g < NT) continue;
                const int *t = &m->tet[4 * (size_t)(g - NT)];
                int hit = 0;
                for (int a = 0; a < 4; a++) hit += (t[a] == s[1]) + (t[a] == s[2]);
                if (hit == 2) tcell = g - NT;
            }
            if (tcell < 0) return fail(CFDM_E_MESH, "cfdm_preprocess: triangle %d has no tetrahedron", c);
            V3 sv1, sv2, scen, vcen, scvc;
            for (int d = 0; d < 3; d++) {
                sv1.v[d] = X[3 * (size_t)s[1] + d] - X[3 * (size_t)s[0] + d];
                sv2.v[d] = X[3 * (size_t)s[2] + d] - X[3 * (size_t)s[0] + d];
            }
            V3 snv = cross(sv1, sv2);
            normalize(snv);
            for (int d = 0; d < 3; d++) {
                scen.v[d] = 0.0;
                for (int a = 0; a < 3; a++) scen.v[d] += X[3 * (size_t)s[a] + d];
                scen.v[d] /= 3.0;
                vcen.v[d] = 0.0;
                for (int a = 0; a < 4; a++) vcen.v[d] += X[3 * (size_t)m->tet[4 * (size_t)tcell + a] + d];
                vcen.v[d] /= 4.0;
                scvc.v[d] = vcen.v[d] - scen.v[d];
            }
            normalize(scvc);
            if (dot(scvc, snv) < 0.0) std::swap(s[0], s[2]);
        }
        // Boundary edges and their area vectors, Area_Volume.cpp:462-572.
        const int B = 3 * NT;
        m->bedge.assign(2 * (size_t)B, 0);
        m->bedgeS.assign(3 * (size_t)B, 0.0);
        m->btag.assign(B, 0);
        for (int c = 0; c < NT; c++) {
            const int *s = &m->ftri[3 * (size_t)c];
            double surcen[3];
            for (int d = 0; d < 3; d++)
                surcen[d] = (X[3 * (size_t)s[0] + d] + X[3 * (size_t)s[1] + d] + X[3 * (size_t)s[2] + d]) / 3.0;
            for (int k = 0; k < 3; k++) {
                const int n0 = s[k];
                // node 0: edges to nodes 1 and 2; node 1: to 2 and 0; node 2: to 0 and 1
                const int na = s[(k + 1) % 3], nbn = s[(k + 2) % 3];
                V3 V1, V2, V3v;
                for (int d = 0; d < 3; d++) {
                    double e1 = (X[3 * (size_t)n0 + d] + X[3 * (size_t)na + d]) / 2.0;
                    double e2 = (X[3 * (size_t)n0 + d] + X[3 * (size_t)nbn + d]) / 2.0;
                    V1.v[d] = e1 - X[3 * (size_t)n0 + d];
                    V2.v[d] = surcen[d] - X[3 * (size_t)n0 + d];
                    V3v.v[d] = e2 - X[3 * (size_t)n0 + d];
                }
                V3 c1 = cross(V1, V2), c2 = cross(V2, V3v);
                const size_t b = 3 * (size_t)c + k;
                for (int d = 0; d < 3; d++) m->bedgeS[3 * b + d] += 0.5 * (c1.v[d] + c2.v[d]);
                m->bedge[2 * b + 0] = n0;
                m->bedge[2 * b + 1] = N + (int)b;
                m->btag[b] = m->triTag[c];
            }
        }
        // Interior edge areas and dual volumes, Area_Volume.cpp:660-821.
        std::vector<double> S(3 * (size_t)E, 0.0);
        std::vector<double> vol(N, 0.0);
        for (int c = 0; c < NE4; c++) {
            const int *node = &m->tet[4 * (size_t)c];
            double volcen[3];
            for (int d = 0; d < 3; d++) {
                volcen[d] = 0.0;
                for (int a = 0; a < 4; a++) volcen[d] += X[3 * (size_t)node[a] + d];
                volcen[d] /= 4.0;
            }
            for (int j = 0; j < 6; j++) {
                const int e0 = node[kTetEdge[j][0]], e1 = node[kTetEdge[j][1]];
                double fl[3], fr[3], ec[3];
                for (int d = 0; d < 3; d++) {
                    fl[d] = 0.0;
                    fr[d] = 0.0;
                    for (int a = 0; a < 3; a++) {
                        fl[d] += X[3 * (size_t)node[kTetFaceL[j][a]] + d];
                        fr[d] += X[3 * (size_t)node[kTetFaceR[j][a]] + d];
                    }
                    fl[d] /= 3.0;
                    fr[d] /= 3.0;
                    ec[d] = 0.0;
                    ec[d] += X[3 * (size_t)e0 + d];
                    ec[d] += X[3 * (size_t)e1 + d];
                    ec[d] /= 2.0;
                }
                V3 V1, V2, V3v, V4;
                for (int d = 0; d < 3; d++) {
                    V1.v[d] = volcen[d] - ec[d];
                    V2.v[d] = fl[d] - ec[d];
                    V3v.v[d] = ec[d] - X[3 * (size_t)e1 + d];
                    V4.v[d] = X[3 * (size_t)e0 + d] - ec[d];
                }
                V3 cr = cross(V1, V2);
                vol[e1] += (1.0 / 6.0) * dot(V3v, cr);
                vol[e0] += (1.0 / 6.0) * dot(V4, cr);
                double av[3];
                for (int d = 0; d < 3; d++) av[d] = 0.5 * cr.v[d];
                for (int d = 0; d < 3; d++) {
                    V1.v[d] = fr[d] - ec[d];
                    V2.v[d] = volcen[d] - ec[d];
                }
                cr = cross(V1, V2);
                vol[e1] += (1.0 / 6.0) * dot(V3v, cr);
                vol[e0] += (1.0 / 6.0) * dot(V4, cr);
                for (int d = 0; d < 3; d++) av[d] += 0.5 * cr.v[d];
                if (e1 > e0)
                    for (int d = 0; d < 3; d++) av[d] = -1.0 * av[d];
                const int e = find_edge(e0, e1);
                if (e < 0) return fail(CFDM_E_MESH, "cfdm_preprocess: tetrahedron edge not in node graph");
                for (int d = 0; d < 3; d++) S[3 * (size_t)e + d] += av[d];
            }
        }
        // Outward boundary normals, Area_Volume.cpp:913-916.
        for (auto &v : m->bedgeS) v = -1.0 * v;
        m->n2nIA.swap(n2nIA);
        m->n2nJA.swap(n2nJA);
        m->edge.swap(edge);
        m->edgeS.swap(S);
        m->vol.swap(vol);
        m->ready = true;
    } catch (const std::bad_alloc &) {
        return fail(CFDM_E_OOM, "cfdm_preprocess: out of memory");
    }
    return CFDM_OK;
}

extern "C" int cfdm_get_info(const cfdm_mesh *m, cfdm_info *info) {
    if (!m || !info) return fail(CFDM_E_ARG, "cfdm_get_info: null argument");
    info->n_node = m->nNode;
    info->n_tri = m->nTri;
    info->n_tet = m->nTet;
    info->n_edge = m->ready ? (int32_t)(m->edge.size() / 2) : 0;
    info->n_bedge = m->ready ? (int32_t)(m->bedge.size() / 2) : 0;
    info->n_n2n = m->ready ? (int32_t)m->n2nJA.size() : 0;
    return CFDM_OK;
}

extern "C" int cfdm_get_arrays(const cfdm_mesh *m, double *xyz, int32_t *tet, int32_t *tri,
                               int32_t *tri_tag, int32_t *edge, double *edge_area, int32_t *bedge,
                               double *bedge_area, int32_t *bedge_tag, double *vol, int32_t *n2n_ia,
                               int32_t *n2n_ja) {
    if (!m) return fail(CFDM_E_ARG, "cfdm_get_arrays: null mesh");
    const bool need_pre = edge || edge_area || bedge || bedge_area || bedge_tag || vol || n2n_ia || n2n_ja;
    if (need_pre && !m->ready) return fail(CFDM_E_ARG, "cfdm_get_arrays: mesh not preprocessed");
    auto cp = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    cp(xyz, m->xyz);
    cp(tet, m->tet);
    cp(tri, m->ready ? m->ftri : m->tri);
    cp(tri_tag, m->triTag);
    cp(edge, m->edge);
    cp(edge_area, m->edgeS);
    cp(bedge, m->bedge);
    cp(bedge_area, m->bedgeS);
    cp(bedge_tag, m->btag);
    cp(vol, m->vol);
    cp(n2n_ia, m->n2nIA);
    cp(n2n_ja, m->n2nJA);
    return CFDM_OK;
}

// ---------------------------------------------------------------------------
// Restart and solution files (conservative variables)
// ---------------------------------------------------------------------------
// Restart_Writer, Euler3D/src/RestartIO.cpp:45-109: iteration, node count,
// ghost count and variable count (int32), then the five variables of every
// node and ghost, interleaved per node (native byte order, as fwrite).
extern "C" int cfdm_restart_write(const char *path, int32_t iteration, int32_t n_node, int32_t n_ghost,
                                  const double *q) {
    if (!path || !q || n_node < 0 || n_ghost < 0) return fail(CFDM_E_ARG, "cfdm_restart_write: invalid arguments");
    const size_t nt = (size_t)n_node + (size_t)n_ghost;
    std::vector<double> buf(5 * nt);
    for (size_t i = 0; i < nt; i++)
        for (int k = 0; k < 5; k++) buf[5 * i + k] = q[k * nt + i];
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail(CFDM_E_IO, "cfdm_restart_write: cannot write %s", path);
    const int32_t head[4] = {iteration, n_node, n_ghost, 5};
    const bool ok = fwrite(head, sizeof(int32_t), 4, fp) == 4 && fwrite(buf.data(), sizeof(double), buf.size(), fp) ==
                                                                      buf.size();
    if (fclose(fp) != 0 || !ok) return fail(CFDM_E_IO, "cfdm_restart_write: write to %s failed", path);
    return CFDM_OK;
}

// Restart_Reader, Euler3D/src/RestartIO.cpp:112-179: the counts must match
// the mesh (the reference stops otherwise); q receives [5][n_node + n_ghost].
extern "C" int cfdm_restart_read(const char *path, int32_t n_node, int32_t n_ghost, int32_t *iteration, double *q) {
    if (!path || !q || n_node < 0 || n_ghost < 0) return fail(CFDM_E_ARG, "cfdm_restart_read: invalid arguments");
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail(CFDM_E_IO, "cfdm_restart_read: cannot open %s", path);
    int32_t head[4] = {0, 0, 0, 0};
    const size_t nt = (size_t)n_node + (size_t)n_ghost;
    std::vector<double> buf(5 * nt);
    int rc = CFDM_OK;
    if (fread(head, sizeof(int32_t), 4, fp) != 4)
        rc = fail(CFDM_E_IO, "cfdm_restart_read: %s is truncated", path);
    else if (head[1] != n_node)
        rc = fail(CFDM_E_ARG, "cfdm_restart_read: %s holds %d nodes, the mesh %d", path, head[1], n_node);
    else if (head[2] != n_ghost)
        rc = fail(CFDM_E_ARG, "cfdm_restart_read: %s holds %d ghosts, the mesh %d", path, head[2], n_ghost);
    else if (head[3] != 5)
        rc = fail(CFDM_E_ARG, "cfdm_restart_read: %s holds %d variables, expected 5", path, head[3]);
    else if (fread(buf.data(), sizeof(double), buf.size(), fp) != buf.size())
        rc = fail(CFDM_E_IO, "cfdm_restart_read: %s is truncated", path);
    fclose(fp);
    if (rc) return rc;
    for (size_t i = 0; i < nt; i++)
        for (int k = 0; k < 5; k++) q[k * nt + i] = buf[5 * i + k];
    if (iteration) *iteration = head[0];
    return CFDM_OK;
}

// VTK_Writer, Euler3D/src/MeshIO.cpp:219-399 (conservative variables, GENERIC
// non-dimensionalisation): legacy ASCII unstructured grid with the boundary
// triangles (their surface tag as BoundaryIds) and the tetrahedra as cells,
// and density, velocity components, pressure (Get_Pressure), temperature
// (Get_Temperature), Mach (Get_Mach, Euler3D/src/Material.cpp:701-1100) and
// the velocity vector as point data.  q: [5][ldq], nodes first.
extern "C" int cfdm_vtk_write(const cfdm_mesh *m, const char *path, const double *q, int32_t ldq, double gamma,
                              double r_gas) {
    if (!m || !path || !q || ldq < m->nNode) return fail(CFDM_E_ARG, "cfdm_vtk_write: invalid arguments");
    FILE *fp = fopen(path, "w");
    if (!fp) return fail(CFDM_E_IO, "cfdm_vtk_write: cannot write %s", path);
    const int N = m->nNode;
    const std::vector<int> &tri = m->ready ? m->ftri : m->tri;  // triangles as the solver holds them
    fprintf(fp, "# vtk DataFile Version 2.0\n");
    fprintf(fp, "Solver solution data\n");
    fprintf(fp, "ASCII \n");
    fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
    fprintf(fp, "POINTS %d double\n", N);
    for (int i = 0; i < N; i++)
        fprintf(fp, "%f %f %f\n", m->xyz[3 * (size_t)i], m->xyz[3 * (size_t)i + 1], m->xyz[3 * (size_t)i + 2]);
    const int ncell = m->nTri + m->nTet;
    fprintf(fp, "CELLS %d %d\n", ncell, 4 * m->nTri + 5 * m->nTet);
    for (int c = 0; c < m->nTri; c++)
        fprintf(fp, "3 %d %d %d\n", tri[3 * (size_t)c], tri[3 * (size_t)c + 1], tri[3 * (size_t)c + 2]);
    for (int c = 0; c < m->nTet; c++)
        fprintf(fp, "4 %d %d %d %d\n", m->tet[4 * (size_t)c], m->tet[4 * (size_t)c + 1], m->tet[4 * (size_t)c + 2],
                m->tet[4 * (size_t)c + 3]);
    fprintf(fp, "CELL_TYPES %d\n", ncell);
    for (int c = 0; c < m->nTri; c++) fprintf(fp, "5\n");
    for (int c = 0; c < m->nTet; c++) fprintf(fp, "10\n");
    fprintf(fp, "CELL_DATA %d\n", ncell);
    fprintf(fp, "SCALARS BoundaryIds Int 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int c = 0; c < m->nTri; c++) fprintf(fp, "%d\n", m->triTag[c]);
    for (int c = 0; c < m->nTet; c++) fprintf(fp, "0\n");
    fprintf(fp, "POINT_DATA %d\n", N);
    const double *Q1 = q, *Q2 = q + ldq, *Q3 = q + 2 * (size_t)ldq, *Q4 = q + 3 * (size_t)ldq,
                 *Q5 = q + 4 * (size_t)ldq;
    const char *names[4] = {"Density", "X_Velocity", "Y_Velocity", "Z_Velocity"};
    for (int f = 0; f < 4; f++) {
        fprintf(fp, "SCALARS %s double 1\n", names[f]);
        fprintf(fp, "LOOKUP_TABLE default\n");
        for (int i = 0; i < N; i++) {
            const double v = f == 0 ? Q1[i] : (f == 1 ? Q2[i] : (f == 2 ? Q3[i] : Q4[i])) / Q1[i];
            fprintf(fp, "%22.15e\n", v);
        }
    }
    for (int f = 0; f < 3; f++) {
        fprintf(fp, "SCALARS %s double 1\n", f == 0 ? "Pressure" : (f == 1 ? "Temperature" : "Mach"));
        fprintf(fp, "LOOKUP_TABLE default\n");
        for (int i = 0; i < N; i++) {
            double v;
            if (f < 2) {
                const double p = (gamma - 1.0) * (Q5[i] - 0.5 * (Q2[i] * Q2[i] + Q3[i] * Q3[i] + Q4[i] * Q4[i]) / Q1[i]);
                v = f == 0 ? p : p / (r_gas * Q1[i]);
            } else {
                const double q2 = (Q2[i] * Q2[i] + Q3[i] * Q3[i] + Q4[i] * Q4[i]) / (Q1[i] * Q1[i]);
                const double p = (gamma - 1.0) * (Q5[i] - 0.5 * Q1[i] * q2);
                v = sqrt(q2 * Q1[i] / (gamma * p));
            }
            fprintf(fp, "%22.15e\n", v);
        }
    }
    fprintf(fp, "VECTORS Velocity double \n");
    for (int i = 0; i < N; i++) fprintf(fp, "%22.15e \t %22.15e \t %22.15e\n", Q2[i] / Q1[i], Q3[i] / Q1[i], Q4[i] / Q1[i]);
    if (fclose(fp) != 0) return fail(CFDM_E_IO, "cfdm_vtk_write: write to %s failed", path);
    return CFDM_OK;
}
