// Internal definitions shared by the translation units of libcfd_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "cfd_hip.h"

struct cfd_ctx {
    int device = 0;
    std::mutex lock;  // guards err
    std::string err;
    // device address of this GPU's copy of the scratch array of
    // cfd_linsolve_gs_noalloc (looked up when the context is created), or
    // NULL if the lookup failed
    double *noalloc_scratch = nullptr;
};

// Device buffers of one precision for the residual path.  Node data used by
// neighbour gathers is packed per node so one gather touches whole sectors:
//   qa [NT][8]  = q1..q5, x, y, z        (ghost records hold q only)
//   qb [N][24]  = grad (Q1x,Q1y,Q1z,...,Q5z), phi1..phi5, 4 pad
//   sa [B][4]   = outward area vector x, y, z, pad of each boundary edge
//   lsq [N][8]  = r11 r12 r13 r22 r23 ro22 ro33 pad
//   res [N][12] = convective residual (5), dissipative residual (5), time-step
//                 accumulator, pad; one contiguous record per node because
//                 nodes are processed in a locality order, not storage order.
//   isx, isy, isz = area vector of every node incidence (edge orientation,
//                 L -> R, boundary outward) in the sliced layout of
//                 cfd_mesh::nbr, so each step of a node walk is one coalesced
//                 load per wavefront.
template <class T>
struct ResidualBufs {
    T *qa = nullptr;
    T *qb = nullptr;
    T *sa = nullptr;
    T *isx = nullptr, *isy = nullptr, *isz = nullptr;
    T *vol = nullptr;
    T *lsq = nullptr;
    T *res = nullptr;
};
constexpr int kQA = 8;
constexpr int kQB = 20;  // gradient (15) and limiter (5) of a node, unpadded
constexpr int kPHI = 15;  // offset of phi in a qb record
constexpr int kRES = 12;  // residual record: conv 0-4, diss 5-9, dt accumulator 10

struct LevelSchedule {
    std::vector<int> ptr;   // host: level l holds rows[ptr[l]..ptr[l+1])
    int *rows = nullptr;    // device: rows sorted by level
    int nlevels = 0;
    int maxwidth = 0;
    // rows in level order with every level padded (-1) to a multiple of 12
    // (sync-free sweeps, 12 rows of 5 lanes per wavefront)
    int *slots12 = nullptr;
    int n12 = 0;
};

constexpr int kMaxKrylov = 64;   // largest FGMRES restart length
constexpr int kDotParts = 512;   // partial sums of one Krylov dot product, at most
// Blocks of 256 threads for a Krylov dot product over n entries: one entry per
// thread up to kDotParts blocks, grid strides beyond.  Batched solves give each
// case the blocks its solve alone would use, so both take them from here.
inline int dot_parts(size_t n) {
    const size_t g = (n + 256 - 1) / 256;
    return (int)(g < 1 ? 1 : (g < (size_t)kDotParts ? g : kDotParts));
}
constexpr int kBOut = 2 * (kMaxKrylov + 1) + 1;  // per-case results of one batched Arnoldi step

// Block map of a segmented reduction over the cases of a union mesh: case c
// gets nb[c] blocks that stride over its own range exactly as a launch of
// nb[c] blocks over that case alone does, and its partial sums are combined
// in the same order, so per-case results equal the single-mesh ones bit for
// bit.
struct SegMap {
    int nblocks = 0;
    int2 *blk = nullptr;                 // device [nblocks]: (case, block within the case)
    int *first = nullptr;                // device [ncase+1]: first block of each case
    int *off = nullptr, *len = nullptr;  // device [ncase]: range of each case
    std::vector<int> h_first;
};

// Captured graphs of the FGMRES Arnoldi steps (one per step index), valid
// for one Krylov buffer, restart length and preconditioner.
struct KrylovGraphs {
    const void *key = nullptr, *lu = nullptr;
    int prec = -1, m = 0;
    std::vector<void *> exec;  // hipGraphExec_t
};

// Block rows in a sliced ELL layout with the 25 entries of a block spread
// across lanes: the k-th block of row i (slice s = i/64, lane l = i%64) is
// slot p = off[s] + 64k + l, and entry e of slot p is value 25(p - l) + 64e + l.
// Each group of 64 slots is a [25][64] tile, so for every (k, e) the 64 rows
// of a slice read one contiguous 64-value run.  Row lengths come from the
// BSR pattern the layout mirrors; padding slots have src -1.
// Rows from which the block-SELL kernels (one row per lane) are used: 2048
// wavefronts, two per SIMD of the 256 CUs.
constexpr int kSellMinRows = 131072;

// Runs fn(lo, hi) over contiguous chunks of [0, n), each of at least `grain`
// items, on up to 16 host threads; used by the host-side mesh set-up, whose
// loops write disjoint outputs, so results do not depend on the thread count.
template <class F>
void cfd_parallel_for(int64_t n, F fn, int64_t grain = 1 << 14) {
    const int64_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const int nt = (int)std::max<int64_t>(1, std::min(hw, n / std::max<int64_t>(1, grain)));
    if (nt <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int t = 1; t < nt; t++) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    fn((int64_t)0, n / nt);
    for (auto &x : th) x.join();
}

// Host array without value initialisation, for large set-up arrays that a
// parallel pass writes in full.
template <class T>
struct HostArray {
    std::unique_ptr<T[]> p;
    size_t n;
    explicit HostArray(size_t count) : p(new T[count ? count : 1]), n(count) {}
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
    size_t size() const { return n; }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
};

// CFD_PROGRESS=1: the Newton loops of cfd_solve_steady and cfd_batch_solve
// print one line per iteration to stderr (long solves stay observable).
inline bool cfd_progress() {
    static const bool on = [] {
        const char *e = getenv("CFD_PROGRESS");
        return e && e[0] == '1';
    }();
    return on;
}

struct BlockSell {
    int64_t nblk = 0;     // slots, a multiple of 64
    int *off = nullptr;   // [nslice + 1]
    int *col = nullptr;   // [nblk] block column (0 for padding)
    int *src = nullptr;   // [nblk] BSR slot it is copied from, -1 for padding
};

// Work-edge patches of the flux-once residual: patch p holds the internal
// nodes [pn_start[p], pn_start[p+1]) (at most 256, never across cases of a
// union) and its work edges pw_ptr[p] .. pw_ptr[p+1]: each interior edge with
// both ends in the patch once, each edge to another patch, each boundary edge
// of a patch node.  The work edges of a patch go in one of two orders:
//   kPatchNodeOrder: by their first end in the patch (then L, R): the chunks
//     of a patch touch compact groups of nodes (the fast build);
//   kPatchEdgeOrder: by reference edge index, boundary edges last, so each
//     node adds its fluxes in the reference's order (the exact build).
constexpr int kPatchNodeOrder = 0, kPatchEdgeOrder = 1;
struct PatchSet {
    int npatch = 0;
    int64_t npw = 0;
    int *pn_start = nullptr;  // [npatch + 1]
    int *pw_ptr = nullptr;    // [npatch + 1]
    int *pw_l = nullptr;      // [npw] internal index of the L end
    int *pw_r = nullptr;      // [npw] internal index of the R end, or N + b for boundary edge b
    int *pn_slot = nullptr;   // SELL as nbr: patch-local slot of each incidence, ascending per node; bit 31: R end
    double *ws = nullptr;     // [npw][3] area vector, oriented L -> R (outward for boundary edges)
    float *ws_f = nullptr;    // the same in FP32 (made on first FP32 use)
};

struct cfd_mesh {
    cfd_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    int N = 0, E = 0, B = 0, NT = 0;
    int64_t bytes = 0;
    std::vector<std::pair<void *, size_t>> allocs;  // device allocations and their sizes

    // topology
    int *order = nullptr;     // [N] reference index of each internal node (internal = Morton order)
    std::vector<int> h_order;
    int *inc_ptr = nullptr;   // [N+1]
    // node incidence lists, in the order the reference's sequential loops
    // deliver each node's contributions: interior edges by edge index, then
    // boundary edges (inc_ptr gives each node's count)
    // The same incidence lists in a sliced ELL layout (SELL-64): slice s holds
    // internal nodes 64s..64s+63; entry k of the node in lane l sits at
    // sptr[s] + 64 k + l, so the 64 lanes of a wavefront read consecutive
    // words at every step.  nbr = j if the node is the L end of the edge,
    // ~j if it is the R end, N + b for boundary edge b (ghost record).
    int *sptr = nullptr;      // [ceil(N/64)+1]
    int *nbr = nullptr;       // [sptr[last]]
    int *soff = nullptr;      // [sptr[last]] BSR block (row of the node, column of the neighbour), -1 for ghosts
    size_t nsell = 0;
    int *n2n_ia = nullptr, *n2n_ja = nullptr;
    int *bnode = nullptr, *btype = nullptr, *btag = nullptr;  // [B]
    // Roe pressure correction: bit n of roe_cand[a] set where node a's test
    // with its n-th neighbour (reference numbering and order) can zero a
    // limiter (bit 63: every neighbour from the 63rd on); roe_blk[a >> 12] ==
    // mark_epoch where any node of that block has one.
    uint64_t *roe_cand = nullptr;  // [N]
    int *roe_blk = nullptr;        // [ceil(N / 4096)]
    int *iperm = nullptr;          // [N] internal index of each reference node
    int *risk = nullptr;      // [1] nodes with a near-vacuum face state
    int mark_epoch = 0;

    // Work-edge patches of the flux-once residual, built on first use, one
    // set per work-edge order (kPatchNodeOrder for the fast build,
    // kPatchEdgeOrder for the exact build)
    PatchSet patch[2];
    // host, SELL as nbr: the reference edge index of each incidence (E + b for
    // boundary edge b; -1 for padding), which orders kPatchEdgeOrder's work edges
    HostArray<int> h_sell_edge{0};

    ResidualBufs<double> d;   // FP64
    ResidualBufs<float> f;    // FP32 (allocated on first use)
    bool f_ready = false;

    // time step and linear system (FP64 unless noted)
    double *dt = nullptr;     // [N]
    int nnz = 0;
    std::vector<int> h_ia, h_ja, h_iau;
    int *ia = nullptr, *ja = nullptr, *iau = nullptr;
    // ILU(0) merge map: for lower entry kk = (i, k) of the BSR pattern and the
    // t-th upper block of row k, ilu_mqq[ilu_moff[kk] + t] is the slot of the
    // block of row i in the same column, or -1
    int *ilu_moff = nullptr, *ilu_mqq = nullptr;  // built on first use (cfd_ensure_ilu_map)
    int64_t ilu_nmap = 0;
    double *a = nullptr, *b = nullptr, *x = nullptr;
    float *af = nullptr;      // FP32 Jacobian (mixed)
    double *lu = nullptr, *dinv = nullptr;
    float *luf = nullptr, *dinvf = nullptr;
    double *gslu = nullptr;   // [N][25] no-pivot LU of the diagonal blocks (block GS)
    double *rowbuf = nullptr; // [N] per-row sums of the GS sweep
    LevelSchedule lower, upper;
    // multicolour ILU(0): rows ordered by the colour of a greedy colouring of
    // the node graph (in reference order), then by index, so the triangular
    // factors have one level per colour
    int ncolor = 0;
    int *mc_perm = nullptr;                               // [N] original row of each permuted row
    int *mc_ia = nullptr, *mc_ja = nullptr, *mc_iau = nullptr;  // permuted BSR pattern
    int *mc_src = nullptr;                                // [nnz] original slot of each permuted slot
    int *mc_moff = nullptr, *mc_mqq = nullptr;            // ILU(0) merge map of the permuted pattern
    LevelSchedule mc_lower, mc_upper;
    double *zp = nullptr;                                 // [N][5] permuted work vector
    float *zpf = nullptr;
    bool sync_free = true;    // single-launch triangular sweeps (else one launch per level)
    int *ready = nullptr;     // [N] sweep completion epochs
    int *epoch = nullptr;     // [1] current sweep epoch
    // block-SELL copies for the Krylov loop: A (SpMV), and the multicolour
    // ILU(0) factors split into strict lower and upper parts plus the
    // inverted diagonal blocks (one block per row, slot = row)
    BlockSell a_sell, mcl_sell, mcu_sell;
    double *asv = nullptr, *lsv = nullptr, *usv = nullptr, *dsv = nullptr;
    float *asvf = nullptr, *lsvf = nullptr, *usvf = nullptr, *dsvf = nullptr;
    // FP16 block-SELL copies of the multicolour factors (linear solves with
    // mixed >= 2) and of A (mixed = 3): each block's entries over its largest
    // magnitude, with that scale kept per block in FP32
    _Float16 *lsvh = nullptr, *usvh = nullptr, *dsvh = nullptr, *asvh = nullptr;
    float *lsvs = nullptr, *usvs = nullptr, *dsvs = nullptr, *asvs = nullptr;
    bool asell_half = false;  // the SELL copy of A is the FP16 one
    bool asell_valid = false;    // A's SELL copy matches the current Jacobian
    bool use_sell = false;       // Krylov loop on the block-SELL copies (else the BSR arrays)
    std::vector<int> mc_cstart;  // [ncolor + 1] first permuted row of each colour
    bool jac_valid = false, jac_mixed = false;
    bool prec_valid = false;
    int prec_kind = -1;
    bool prec_mixed = false;
    bool prec_half = false;  // the sweeps read the FP16 factor copies

    // autotuned variants (cfd_autotune): gather mode of the fused
    // gradient+limiter per precision (FP64, FP32; -1: the fixed rule), and
    // whether a steady solve has tuned this mesh
    int8_t tune_gl_stage[2] = {-1, -1};
    bool tuned = false;

    // Krylov workspace
    bool use_graphs = true;   // replay captured Arnoldi steps
    KrylovGraphs kgraph;
    int kry_m64 = 0, kry_m32 = 0;  // restart length the FP64 / FP32 Krylov basis holds
    double *V = nullptr, *Z = nullptr, *w = nullptr, *r = nullptr;
    float *Vf = nullptr, *Zf = nullptr, *wf = nullptr;
    // reductions
    double *red = nullptr;     // device scratch for partial sums
    double *redout = nullptr;  // device final results (up to 128)
    double *hred = nullptr;    // pinned host mirror
    int *flag = nullptr;       // device error flags [4]
    int *hflag = nullptr;      // pinned host mirror
    unsigned long long *eigbits = nullptr;  // [6]
    bool residual_valid = false;
    int res_prec = CFD_FP64;   // precision of the last residual
    bool hllc_check = false;   // HLLC anomaly flag not yet read back
    int64_t q_version = 0, f_version = -1;  // FP32 mirror of the state is current when equal

    // union of independent cases (batched solves); ncase = 0 for one mesh.
    // Nodes, boundary edges and [N][5] vector entries of case c occupy one
    // contiguous range in both numberings.
    int ncase = 0;
    std::vector<int> case_node, case_bedge;  // host offsets [ncase+1]
    double *case_inf = nullptr;              // device [ncase][8] freestream rho, u, v, w, p, outflow p
    int *bcase = nullptr;                    // device [B] case of each boundary edge
    int *ecase = nullptr;                    // device [N] case of each node
    int *case_active = nullptr;              // device [ncase] 1 while a case still iterates
    // device [ncase] failure bits of each case, raised by the kernels that
    // flag a failure for the whole mesh: 1 negative ghost pressure, 2 no
    // boundary-condition instance applies, 4 no HLLC wave configuration
    int *case_flag = nullptr;
    SegMap seg_rms, seg_dot, seg_forces;
    double *bpart = nullptr;                 // device partial sums of segmented reductions
    double *bred = nullptr, *hbred = nullptr;  // per-case results [ncase][kBOut] (device, pinned host)
    double *bcoef = nullptr, *hbcoef = nullptr;  // [ncase][kMaxKrylov+1] per-case coefficients (device, pinned)
    int *bkdim = nullptr, *hbkdim = nullptr;     // [ncase] per-case Krylov dimension (device, pinned)
    KrylovGraphs kgraph_b;                   // batched Arnoldi steps
};

// ---------------------------------------------------------------------------
// helpers implemented in cfd_api.hip
// ---------------------------------------------------------------------------
int cfd_set_error(cfd_ctx *ctx, int code, const char *fmt, ...);
int cfd_mesh_create_cases(cfd_ctx *ctx, const cfd_mesh_desc *d, const cfd_params *p, const int *node_case,
                          cfd_mesh **out);
int cfd_hip_check(cfd_ctx *ctx, hipError_t e, const char *what);
// Stream capture beside other threads.  HIP (as CUDA) counts device and
// pinned-host allocation and free, stream creation and destruction and
// synchronous copies and fills as unsafe while any stream of the process is
// being captured; the relaxed capture mode only stops it from rejecting
// them.  Worker threads of the solve server drive meshes concurrently, so the
// library holds this lock shared around every such call and exclusively
// around each capture (begin, launches, end, instantiate): no unsafe call of
// the library overlaps a capture of another of its threads.
std::shared_mutex &cfd_capture_mutex();
struct CfdRuntimeGuard {
    std::shared_lock<std::shared_mutex> lk{cfd_capture_mutex()};
};
hipError_t cfd_memcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);  // hipMemcpy, guarded
hipError_t cfd_host_alloc(void **p, size_t bytes);                                     // hipHostMalloc, guarded
void *cfd_dalloc(cfd_mesh *m, size_t bytes);  // device allocation tracked on the mesh
void cfd_dfree(cfd_mesh *m, void *p);          // frees and untracks an allocation of cfd_dalloc (NULL: nothing)
int cfd_ensure_fp32(cfd_mesh *m);
// Builds the work-edge patches of the flux-once residual from the device
// incidence lists (once per mesh) and the FP32 copy of their area vectors
// when fp32 is set.
int cfd_ensure_patches(cfd_mesh *m, bool fp32, int order);  // order: kPatchNodeOrder or kPatchEdgeOrder
int cfd_ensure_ilu_map(cfd_mesh *m);  // natural-order ILU(0) merge map, on first use
// Whether the residual of these parameters runs the flux-once patch kernel:
// the fast build at second order, unless CFD_FLUX_ONCE=0 or 1 forces it.
bool cfd_residual_uses_patches(const cfd_params *p);
// Upper bound of the device bytes of a mesh's work-edge patches (n nodes,
// e edges, b boundary edges, split into `cases` cases).
int64_t cfd_patch_footprint_bytes(int64_t n, int64_t e, int64_t b, int64_t cases);

#define CFD_HIP(m, call)                                                   \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess) return cfd_hip_check((m)->ctx, e_, #call);   \
    } while (0)

// Launch geometry: blocks are remapped so each of the 8 XCDs (dispatched
// round-robin by blockIdx.x) works on a contiguous range of nodes, keeping the
// neighbour gathers of a block in its own XCD's L2.
constexpr int kXcd = 8;
inline int xcd_grid(int nblocks) { return ((nblocks + kXcd - 1) / kXcd) * kXcd; }
__device__ __forceinline__ int xcd_block() {
    const int per = gridDim.x / kXcd;
    return (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
}
constexpr int kSlice = 64;  // nodes per SELL slice (one wavefront)
// Streaming load: read once per kernel, so it is not kept in L2 at the
// expense of the node records that neighbour gathers reuse.
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// kernel launchers implemented in residual.hip / linsolve.hip
// ---------------------------------------------------------------------------
// residual.hip is compiled twice: res_exact holds every launcher (the
// reference's operation order, CFD_ARITH_EXACT); res_fast
// (residual_fast.hip: FMA contraction, shared reciprocals, CFD_ARITH_FAST)
// holds the gradient, limiter and residual launchers.
#define CFD_RES_LAUNCHERS                                                             \
    template <class T>                                                                \
    int launch_gradient(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p);        \
    template <class T>                                                                \
    int launch_limiter(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p);         \
    template <class T>                                                                \
    int launch_gradient_limiter(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p); \
    template <class T>                                                                \
    int launch_residual(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, bool stats);
namespace res_exact {
CFD_RES_LAUNCHERS
int launch_lsq_coef(cfd_mesh *m);
template <class T>
int launch_bc(cfd_mesh *m, ResidualBufs<T> &b, const cfd_params *p, int iter);
int launch_delta_t(cfd_mesh *m, double cfl);
int launch_rms(cfd_mesh *m, double *sums5);
int launch_update(cfd_mesh *m, double relax);
int launch_forces(cfd_mesh *m, const cfd_params *p, int wall_tag, double *out4);
int launch_convert_state_to_fp32(cfd_mesh *m);
int launch_convert_geometry_fp32(cfd_mesh *m);
int launch_residual_to_fp64(cfd_mesh *m);
}  // namespace res_exact
namespace res_fast {
CFD_RES_LAUNCHERS
}  // namespace res_fast
#undef CFD_RES_LAUNCHERS
using namespace res_exact;

int launch_jacobian(cfd_mesh *m, const cfd_params *p, int rhs_mode, int mixed);
// The Jacobian of one step of the steady loop (cfd_solve_steady, batches).
int solve_jacobian(cfd_mesh *m, const cfd_params &p, const cfd_solve_opts *o, int it);
// batched solves (batch.hip): cases that may share one union mesh, and the
// solve itself with each case's final state copied to q_out[c] (may be NULL);
// allow_graphs false launches the Arnoldi steps directly (callers whose other
// threads use the runtime meanwhile: stream capture is not safe beside them).
// done(done_arg, c), when given, is called once per case as soon as its
// result and state are final (finished cases are harvested whenever the union
// is compacted, the rest at the end); on an error return the cases not yet
// reported carry the error status and are not reported.
bool cfd_same_scheme(const cfd_params &a, const cfd_params &b);
bool cfd_same_solve(const cfd_solve_opts &a, const cfd_solve_opts &b);
typedef void (*cfd_case_done_fn)(void *arg, int c);
int cfd_batch_solve_into(cfd_ctx *ctx, int32_t K, const cfd_mesh_desc *meshes, const cfd_params *params,
                         const cfd_solve_opts *opts, cfd_solve_result *results, double *const *q_out,
                         int64_t *device_bytes, bool allow_graphs, cfd_case_done_fn done, void *done_arg);
int launch_jacobian_fd(cfd_mesh *m, const cfd_params *p, int method, int iter, int add_time, int rhs_mode,
                       int mixed);
int run_linsolve(cfd_mesh *m, const cfd_linsolve_opts *o, cfd_linsolve_stats *st);
// The block Gauss-Seidel solve of run_linsolve with its scratch in the
// file-scope array (cfd_linsolve_gs_noalloc): allocates nothing.
int run_gs_noalloc(cfd_mesh *m, int sweeps, cfd_linsolve_stats *st);
// Batched FGMRES over a union mesh: solves the cases with solve[c] != 0,
// writes each case's iterations, final residual estimate and status (a case
// whose Krylov residual turns non-finite gets CFD_E_NONFINITE and no
// correction; the others are unaffected).
int run_linsolve_batch(cfd_mesh *m, const cfd_linsolve_opts *o, const std::vector<char> &solve,
                       std::vector<int> &iters, std::vector<double> &final_res, std::vector<int> &status);
void drop_krylov_graphs(cfd_mesh *m);
// Autotuning (tune.hip): times the same-bits kernel variants on this mesh and
// keeps the fastest; linsolve runs one linear solve of the assembled system.
int autotune_mesh(cfd_mesh *m, const cfd_params *p, const cfd_linsolve_opts *lin, cfd_tune_report *rep,
                  const std::function<int()> &linsolve);
// Scratch of cfd_linsolve_gs_noalloc (linsolve.hip): the device address of
// the current GPU's copy of the file-scope array, or NULL with *err set.
constexpr int kNoAllocMaxNodes = 32768;
double *noalloc_scratch_address(hipError_t *err);
