// serve.hip - continuous-batching steady-solve server over the C ABI.
//
// Requests queue in arrival order.  Worker threads (max_inflight of them)
// each take the next request that fits the HBM budget, run it to convergence
// on its own mesh handle and HIP stream, and post a completion; a request
// that arrives while others are on the GPU is admitted as soon as a worker
// and memory free up, so solves of different sizes overlap on the device.
// Device meshes (state, CRS structure, Jacobian and Krylov workspace) stay
// resident after a solve and are reused by later requests with the same mesh
// key; idle meshes are evicted least-recently-used when memory is needed.
//
// With batching on, a worker that admits a request also takes the compatible
// requests in the look-ahead window (same scheme and solver options) and
// solves them as one union mesh: every kernel launch then covers all of them,
// and the host waits once per step for the whole batch instead of once per
// request.  Batched meshes are built for the batch and freed after it.
//
// Every kernel on the solve path is deterministic (no atomics in sums,
// fixed-shape reductions), and a batched case uses the reductions it would
// use alone, so a request's result does not depend on what else runs
// concurrently or shares its batch.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <vector>

#include "cfd_internal.h"

namespace {
using Clock = std::chrono::steady_clock;

struct Pending {
    cfd_request req;
    double t_submit;
    int64_t need;    // bytes to reserve if no resident mesh can be reused
    int bypassed;    // later requests admitted ahead of this one
    bool batchable;  // FGMRES solve on a small mesh in the batched layout
};

struct Resident {
    int64_t key;
    cfd_mesh *mesh;
    int64_t bytes;  // bytes reserved against the budget
    int32_t krylov_m, mixed;
    bool patches;   // the reservation covers the flux-once residual's patches
    bool busy;
    uint64_t last_use;
};

int krylov_of(const cfd_solve_opts &o) { return o.lin.solver == CFD_SOLVER_FGMRES ? o.lin.krylov_m : 0; }

// Requests admitted together: one request with its mesh slot, or a batch
// solved as one union mesh holding `reserve` bytes of the budget.
struct Job {
    std::vector<Pending> reqs;
    Resident *slot = nullptr;
    int64_t reserve = 0;
};
}  // namespace

struct cfd_server {
    cfd_ctx *ctx = nullptr;
    cfd_server_opts opts{};
    Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Pending> queue;
    std::deque<cfd_completion> done;
    std::list<Resident> meshes;
    int64_t in_use = 0, peak = 0;
    int running = 0;
    uint64_t tick = 0;
    int64_t admitted = 0, reused = 0, evicted = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

// Frees idle resident meshes, least recently used first, until `need` more
// bytes fit the budget.  Called with the server lock held.
bool make_room(cfd_server *s, int64_t need) {
    while (s->in_use + need > s->opts.hbm_budget) {
        auto victim = s->meshes.end();
        for (auto it = s->meshes.begin(); it != s->meshes.end(); ++it)
            if (!it->busy && (victim == s->meshes.end() || it->last_use < victim->last_use)) victim = it;
        if (victim == s->meshes.end()) return false;
        cfd_mesh_destroy(victim->mesh);
        s->in_use -= victim->bytes;
        s->meshes.erase(victim);
        s->evicted++;
    }
    return true;
}

// Tries to start a batch led by queue entry q: the compatible requests of the
// look-ahead window, up to batch_target requests and batch_nodes nodes, when
// there are at least two and their union mesh fits the budget.  Called with
// the server lock held.
// Requests per batch while requests outnumber the idle workers: enough for
// the queue to fit the idle workers, at most batch_max (1: run alone).
int batch_target(const cfd_server *s) {
    const int64_t idle = std::max(1, s->opts.max_inflight - s->running);
    const int64_t want = ((int64_t)s->queue.size() + idle - 1) / idle;
    return (int)std::min<int64_t>(want, s->opts.batch_max);
}

bool admit_batch(cfd_server *s, size_t q, size_t look, Job &job) {
    const Pending &p = s->queue[q];
    const int64_t cap = s->opts.batch_nodes > 0 ? s->opts.batch_nodes : 2000000;
    std::vector<size_t> take{q};
    int64_t nn = p.req.mesh->n_node, ne = p.req.mesh->n_edge, nb = p.req.mesh->n_bedge;
    const int target = batch_target(s);
    for (size_t t = q + 1; t < look && (int)take.size() < target; t++) {
        const Pending &o = s->queue[t];
        if (!o.batchable || !cfd_same_scheme(p.req.params, o.req.params) || !cfd_same_solve(p.req.opts, o.req.opts))
            continue;
        if (nn + o.req.mesh->n_node > cap) continue;
        take.push_back(t);
        nn += o.req.mesh->n_node;
        ne += o.req.mesh->n_edge;
        nb += o.req.mesh->n_bedge;
    }
    if (take.size() < 2 || nn > INT32_MAX / 4 || ne > INT32_MAX / 4 || nb > INT32_MAX / 4) return false;
    const int64_t need = cfd_footprint_bytes((int32_t)nn, (int32_t)ne, (int32_t)nb, krylov_of(p.req.opts),
                                             p.req.opts.lin.mixed) +
                         (cfd_residual_uses_patches(&p.req.params)
                              ? cfd_patch_footprint_bytes(nn, ne, nb, (int64_t)take.size())
                              : 0);
    if (need > s->opts.hbm_budget || !make_room(s, need)) return false;
    s->in_use += need;
    s->peak = std::max(s->peak, s->in_use);
    // every entry passed over ahead of a taken one counts as bypassed once
    for (size_t k = 0, t = 0; k < take.back(); k++) {
        if (t < take.size() && take[t] == k) {
            t++;
            continue;
        }
        s->queue[k].bypassed++;
    }
    job.reqs.clear();
    for (size_t t : take) job.reqs.push_back(s->queue[t]);
    for (size_t t = take.size(); t-- > 0;) s->queue.erase(s->queue.begin() + (long)take[t]);
    job.slot = nullptr;
    job.reserve = need;
    s->admitted += (int64_t)take.size();
    return true;
}

// Picks the first request of the look-ahead window that can start now and
// reserves its memory (with the requests batched alongside it when batching
// is on).  A request that has been bypassed max_bypass times blocks the
// window until it is admitted.  Called with the server lock held.
// Compatible requests queued in the look-ahead window with the one at q
// (itself included).
int companions(const cfd_server *s, size_t q, size_t look) {
    const Pending &p = s->queue[q];
    int n = 0;
    for (size_t t = q; t < look; t++) {
        const Pending &o = s->queue[t];
        if (o.batchable && cfd_same_scheme(p.req.params, o.req.params) && cfd_same_solve(p.req.opts, o.req.opts)) n++;
    }
    return n;
}

bool admit(cfd_server *s, Job &job, double &retry_at) {
    const size_t look = std::min<size_t>(s->queue.size(), (size_t)std::max(1, s->opts.lookahead));
    retry_at = 0.0;
    for (size_t q = 0; q < look; q++) {
        Pending &p = s->queue[q];
        const int target = batch_target(s);
        if (s->opts.batch_max >= 2 && p.batchable && target >= 2) {
            // batching delay: while the GPU is busy with other solves, a short
            // batch waits for companions (up to batch_wait_us after arrival)
            const double until = p.t_submit + 1e-6 * s->opts.batch_wait_us;
            if (s->running > 0 && s->opts.batch_wait_us > 0 && p.bypassed < s->opts.max_bypass && s->now() < until &&
                companions(s, q, look) < target) {
                retry_at = retry_at > 0.0 ? std::min(retry_at, until) : until;
                continue;
            }
            if (admit_batch(s, q, look, job)) return true;
        }
        const int km = krylov_of(p.req.opts);
        const bool patches = cfd_residual_uses_patches(&p.req.params);
        Resident *r = nullptr;
        for (auto &cm : s->meshes)
            if (!cm.busy && cm.key == p.req.mesh_key && cm.mixed == p.req.opts.lin.mixed && cm.krylov_m >= km &&
                (cm.patches || !patches)) {
                r = &cm;
                break;
            }
        if (r) {
            s->reused++;
        } else if (make_room(s, p.need)) {
            s->meshes.push_back(Resident{p.req.mesh_key, nullptr, p.need, km, p.req.opts.lin.mixed, patches, false, 0});
            r = &s->meshes.back();
            s->in_use += p.need;
            s->peak = std::max(s->peak, s->in_use);
        }
        if (r) {
            r->busy = true;
            r->last_use = ++s->tick;
            for (size_t k = 0; k < q; k++) s->queue[k].bypassed++;
            job.reqs.assign(1, p);
            job.slot = r;
            job.reserve = 0;
            s->queue.erase(s->queue.begin() + (long)q);
            s->admitted++;
            return true;
        }
        if (p.bypassed >= s->opts.max_bypass) break;
    }
    return false;
}

// A batch in flight: its requests' results and which have been posted.
struct BatchRun {
    cfd_server *s;
    const Job *job;
    std::vector<cfd_solve_result> res;
    std::vector<char> posted;
    double t_admit;
    int64_t bytes = 0;
};

cfd_completion batch_completion(const BatchRun &b, int c, int status) {
    cfd_completion o{};
    o.id = b.job->reqs[c].req.id;
    o.status = status;
    o.batch_size = (int32_t)b.job->reqs.size();
    o.result = b.res[c];
    o.t_submit = b.job->reqs[c].t_submit;
    o.t_admit = b.t_admit;
    o.t_done = b.s->now();
    o.footprint = b.job->reserve;
    o.device_bytes = b.bytes;
    return o;
}

// Called by the batch solve when a request's result is final: posts its
// completion at once, while the rest of the batch keeps iterating.
// Text of a failure status of one batched case (the solve alone reports the
// same status with cfd_last_error's text).
const char *case_failure_text(int status) {
    switch (status) {
    case CFD_E_NONFINITE: return "non-finite residual";
    case CFD_E_NEGATIVE_PRESSURE: return "negative pressure in the boundary condition";
    case CFD_E_HLLC: return "no HLLC wave configuration";
    case CFD_E_ARG: return "no boundary-condition instance applies";
    default: return "solve failed";
    }
}

void post_batch_case(void *arg, int c) {
    BatchRun &b = *(BatchRun *)arg;
    cfd_completion o = batch_completion(b, c, b.res[c].status);
    if (o.status) snprintf(o.message, sizeof(o.message), "%s", case_failure_text(o.status));
    std::lock_guard<std::mutex> lk(b.s->mu);
    b.s->done.push_back(o);
    b.posted[c] = 1;
    b.s->cv_done.notify_all();
}

// Solves a batch as one union mesh; each request's completion is posted as
// soon as its case is final.  Returns the completions not yet posted (after
// an error: every request not already reported).
std::vector<cfd_completion> run_batch(cfd_server *s, const Job &job) {
    const int K = (int)job.reqs.size();
    std::vector<cfd_mesh_desc> descs(K);
    std::vector<cfd_params> params(K);
    std::vector<cfd_solve_opts> opts(K);
    std::vector<double *> qo(K);
    for (int c = 0; c < K; c++) {
        descs[c] = *job.reqs[c].req.mesh;
        params[c] = job.reqs[c].req.params;
        opts[c] = job.reqs[c].req.opts;
        qo[c] = job.reqs[c].req.q_out;
    }
    BatchRun b{s, &job, std::vector<cfd_solve_result>(K), std::vector<char>(K, 0), s->now()};
    // captured Arnoldi steps, as for served meshes: the library serialises
    // each capture against the allocations, frees and synchronous copies of
    // the other workers (cfd_capture_mutex)
    const int rc = cfd_batch_solve_into(s->ctx, K, descs.data(), params.data(), opts.data(), b.res.data(), qo.data(),
                                        &b.bytes, true, post_batch_case, &b);
    std::vector<cfd_completion> out;
    std::lock_guard<std::mutex> lk(s->mu);  // posted[] is written under the lock
    for (int c = 0; c < K; c++) {
        if (b.posted[c]) continue;
        cfd_completion o = batch_completion(b, c, rc ? rc : b.res[c].status);
        if (o.status)
            snprintf(o.message, sizeof(o.message), "%s", rc ? cfd_last_error(s->ctx) : case_failure_text(o.status));
        out.push_back(o);
    }
    return out;
}

cfd_completion run_one(cfd_server *s, const Pending &p, Resident *slot) {
    cfd_completion c{};
    c.id = p.req.id;
    c.t_submit = p.t_submit;
    c.t_admit = s->now();
    c.footprint = slot->bytes;
    c.batch_size = 1;
    int rc;
    if (!slot->mesh) {
        rc = cfd_mesh_create(s->ctx, p.req.mesh, &p.req.params, &slot->mesh);
    } else {
        c.reused = 1;
        rc = cfd_mesh_reset(slot->mesh, &p.req.params);
    }
    if (!rc) rc = cfd_solve_steady(slot->mesh, &p.req.params, &p.req.opts, &c.result, p.req.rms_hist);
    if (!rc && p.req.q_out) rc = cfd_field_get(slot->mesh, CFD_FIELD_Q, p.req.q_out);
    c.status = rc;
    if (rc) snprintf(c.message, sizeof(c.message), "%s", cfd_last_error(s->ctx));
    c.t_done = s->now();
    return c;
}

void worker(cfd_server *s) {
    hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        Job job;
        bool got = false;
        double retry_at = 0.0;
        while (!s->stop && !(got = admit(s, job, retry_at))) {
            if (retry_at > 0.0)  // a request is waiting for batch companions
                s->cv_work.wait_for(lk, std::chrono::duration<double>(std::max(1e-4, retry_at - s->now())));
            else
                s->cv_work.wait(lk);
        }
        if (!got) break;
        s->running++;
        lk.unlock();
        if (job.reqs.size() > 1) {
            std::vector<cfd_completion> cs = run_batch(s, job);
            lk.lock();
            s->running--;
            s->in_use -= job.reserve;
            for (auto &c : cs) s->done.push_back(c);
            s->cv_done.notify_all();
            s->cv_work.notify_all();
            continue;
        }
        Resident *slot = job.slot;
        cfd_completion c = run_one(s, job.reqs[0], slot);
        const int64_t actual = slot->mesh ? cfd_mesh_device_bytes(slot->mesh) : 0;
        lk.lock();
        s->running--;
        if (c.status != CFD_OK || !s->opts.keep_resident) {
            if (slot->mesh) cfd_mesh_destroy(slot->mesh);
            s->in_use -= slot->bytes;
            for (auto it = s->meshes.begin(); it != s->meshes.end(); ++it)
                if (&*it == slot) {
                    s->meshes.erase(it);
                    break;
                }
        } else {
            if (actual > slot->bytes) {  // estimate exceeded: account what is really held
                s->in_use += actual - slot->bytes;
                slot->bytes = actual;
                s->peak = std::max(s->peak, s->in_use);
            }
            c.footprint = slot->bytes;
            slot->busy = false;
            slot->last_use = ++s->tick;
        }
        c.device_bytes = actual;
        s->done.push_back(c);
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" int cfd_server_create(cfd_ctx *ctx, const cfd_server_opts *o, cfd_server **out) {
    if (!ctx || !o || !out || o->max_inflight < 1 || o->max_inflight > 64 || o->hbm_budget <= 0 || o->batch_max > 4096)
        return cfd_set_error(ctx, CFD_E_ARG, "cfd_server_create: invalid options");
    cfd_server *s = new (std::nothrow) cfd_server();
    if (!s) return cfd_set_error(ctx, CFD_E_OOM, "cfd_server_create: out of host memory");
    s->ctx = ctx;
    s->opts = *o;
    if (s->opts.lookahead < 1) s->opts.lookahead = 16;
    if (s->opts.max_bypass < 1) s->opts.max_bypass = 8;
    s->t0 = Clock::now();
    for (int w = 0; w < o->max_inflight; w++) s->workers.emplace_back(worker, s);
    *out = s;
    return CFD_OK;
}

extern "C" int cfd_server_submit(cfd_server *s, const cfd_request *r) {
    if (!s || !r || !r->mesh) return cfd_set_error(s ? s->ctx : nullptr, CFD_E_ARG, "cfd_server_submit: null argument");
    const cfd_mesh_desc *d = r->mesh;
    const int64_t need = cfd_footprint_bytes(d->n_node, d->n_edge, d->n_bedge, krylov_of(r->opts), r->opts.lin.mixed) +
                         (cfd_residual_uses_patches(&r->params)
                              ? cfd_patch_footprint_bytes(d->n_node, d->n_edge, d->n_bedge, 1)
                              : 0);
    if (need > s->opts.hbm_budget)
        return cfd_set_error(s->ctx, CFD_E_OOM, "request %lld needs %lld bytes, budget is %lld", (long long)r->id,
                             (long long)need, (long long)s->opts.hbm_budget);
    // batchable: FGMRES, and ghosts numbered after the nodes in boundary-edge
    // order (the layout the union mesh concatenates)
    const int32_t case_cap = s->opts.batch_case_nodes > 0 ? s->opts.batch_case_nodes : kSellMinRows;
    bool batchable = r->opts.lin.solver == CFD_SOLVER_FGMRES && d->n_node <= case_cap;
    for (int32_t b = 0; batchable && b < d->n_bedge; b++) batchable = d->bedge[2 * b + 1] == d->n_node + b;
    std::lock_guard<std::mutex> g(s->mu);
    s->queue.push_back(Pending{*r, s->now(), need, 0, batchable});
    s->cv_work.notify_one();
    return CFD_OK;
}

extern "C" int cfd_server_poll(cfd_server *s, cfd_completion *out, int32_t max, double wait_seconds) {
    if (!s || !out || max < 1) return CFD_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_seconds > 0)
        s->cv_done.wait_for(lk, std::chrono::duration<double>(wait_seconds), [s] { return !s->done.empty(); });
    int n = 0;
    while (n < max && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    return n;
}

extern "C" int cfd_server_get_stats(cfd_server *s, cfd_server_stats *st) {
    if (!s || !st) return CFD_E_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    st->bytes_in_use = s->in_use;
    st->peak_bytes = s->peak;
    st->queued = (int32_t)s->queue.size();
    st->running = s->running;
    st->resident = (int32_t)s->meshes.size();
    st->pad_ = 0;
    st->admitted = s->admitted;
    st->reused = s->reused;
    st->evicted = s->evicted;
    st->clock = s->now();
    return CFD_OK;
}

extern "C" void cfd_server_destroy(cfd_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    hipSetDevice(s->ctx->device);
    for (auto &cm : s->meshes)
        if (cm.mesh) cfd_mesh_destroy(cm.mesh);
    delete s;
}
