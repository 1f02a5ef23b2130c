"""Serving loop and single-launch triangular sweeps on the GPU.

- The sync-free ILU(0) factorization and sweeps give bit-identical FGMRES
  results to the one-launch-per-level path (same arithmetic, same order).
- Every request solved by the server, concurrently with others and possibly
  on a reused resident mesh, returns exactly the state, iteration counts and
  forces of the same solve run alone.
- Admission keeps the reserved HBM within the budget, and the footprint
  estimate bounds what a mesh really holds.
"""
import os

import numpy as np
import pytest

from cfdsolutions_amd import cfd
from cfdsolutions_amd import mesh as cmesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cfd.Context(0)
    yield c
    c.close()


def _jacobian_system(ctx, m, p, sync_free, mixed):
    old = os.environ.get("CFD_SYNC_FREE")
    os.environ["CFD_SYNC_FREE"] = "1" if sync_free else "0"
    try:
        dm = cfd.DeviceMesh(ctx, m, p)
    finally:
        if old is None:
            del os.environ["CFD_SYNC_FREE"]
        else:
            os.environ["CFD_SYNC_FREE"] = old
    rng = np.random.default_rng(7)
    q = dm.get(cfd.FIELD_Q)
    q[:, : m.n_node] *= 1.0 + 0.02 * rng.standard_normal((5, m.n_node))
    dm.set(cfd.FIELD_Q, q)
    dm.bc(p, 0)
    dm.residual(p)
    dm.delta_t(p, 3)
    dm.jacobian(p, rhs_mode=1, mixed=mixed)
    return dm


@pytest.mark.parametrize("mixed", [False, True])
def test_sync_free_sweeps_match_level_launches(ctx, mixed):
    print("sync-free vs levels, mixed", mixed, flush=True)
    m = cmesh.bump_channel(16, 6, 6)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    out = []
    for sf in (False, True):
        dm = _jacobian_system(ctx, m, p, sf, mixed)
        st = dm.linsolve(cfd.SOLVER_FGMRES, cfd.PREC_ILU0, krylov_m=30, max_restarts=3, rtol=1e-8, mixed=mixed)
        out.append((st.iterations, st.final_res, dm.get(cfd.FIELD_X)))
        dm.close()
    assert out[0][0] == out[1][0]
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("precond", [cfd.PREC_ILU0_MC, cfd.PREC_ILU0, cfd.PREC_BJACOBI])
def test_block_sell_matches_bsr(ctx, mixed, precond):
    """The block-SELL SpMV and multicolour sweeps (one row per lane, used for large systems) give the bits of the
    BSR kernels (five lanes per row): same FGMRES iterations, residual and solution; and a steady solve on them
    gives the same state."""
    print("block-SELL vs BSR, mixed", mixed, "precond", precond, flush=True)
    m = cmesh.bump_channel(16, 6, 6)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    out = []
    old = os.environ.get("CFD_SELL")
    try:
        for sell in ("0", "1"):
            os.environ["CFD_SELL"] = sell
            dm = _jacobian_system(ctx, m, p, True, mixed)
            st = dm.linsolve(cfd.SOLVER_FGMRES, precond, krylov_m=30, max_restarts=3, rtol=1e-8, mixed=mixed)
            out.append((st.iterations, st.final_res, dm.get(cfd.FIELD_X)))
            dm.close()
    finally:
        if old is None:
            del os.environ["CFD_SELL"]
        else:
            os.environ["CFD_SELL"] = old
    assert out[0][0] == out[1][0]
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_block_sell_steady_solve_matches_bsr(ctx):
    print("block-SELL vs BSR steady solve", flush=True)
    m = cmesh.bump_channel(12, 4, 4)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    res = []
    old = os.environ.get("CFD_SELL")
    try:
        for sell in ("0", "1"):
            os.environ["CFD_SELL"] = sell
            dm = cfd.DeviceMesh(ctx, m, p)
            r, _ = dm.solve_steady(p, mixed=True, precond=cfd.PREC_ILU0_MC, max_iter=40, tol=1e-10, check=False)
            res.append((r.iterations, r.krylov_total, dm.get(cfd.FIELD_Q)))
            dm.close()
    finally:
        if old is None:
            del os.environ["CFD_SELL"]
        else:
            os.environ["CFD_SELL"] = old
    assert res[0][:2] == res[1][:2]
    assert np.array_equal(res[0][2], res[1][2])


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_replays_match_eager(ctx, mixed):
    """Captured Arnoldi-step graphs, replayed over three Newton steps (new matrix values each time, eager work and
    synchronisations in between), give the bits of the eager launches at every replay."""
    print("graph replays vs eager, mixed", mixed, flush=True)
    m = cmesh.bump_channel(16, 6, 6)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    runs = {}
    old = os.environ.get("CFD_GRAPHS")
    try:
        for graphs in (False, True):
            os.environ["CFD_GRAPHS"] = "1" if graphs else "0"
            dm = cfd.DeviceMesh(ctx, m, p)
            out = []
            for it in range(3):
                dm.bc(p, it)
                dm.residual(p)
                dm.delta_t(p, it)
                dm.jacobian(p, rhs_mode=1, mixed=mixed)
                st = dm.linsolve(cfd.SOLVER_FGMRES, cfd.PREC_ILU0, krylov_m=20, max_restarts=2, rtol=1e-6,
                                 mixed=mixed)
                out.append((st.iterations, st.final_res, dm.get(cfd.FIELD_X)))
                dm.update(1.0)
            runs[graphs] = out
            dm.close()
    finally:
        if old is None:
            os.environ.pop("CFD_GRAPHS", None)
        else:
            os.environ["CFD_GRAPHS"] = old
    for (i0, r0, x0), (i1, r1, x1) in zip(runs[False], runs[True]):
        assert i0 == i1 and r0 == r1
        assert np.array_equal(x0, x1)


REQS = [  # (mesh key, dims, flux, order, mach, alpha)
    (1, (12, 4, 4), "roe", 1, 0.5, 0.0),
    (2, (16, 6, 4), "hllc", 1, 0.6, 2.0),
    (1, (12, 4, 4), "roe", 1, 0.7, -1.0),
    (3, (20, 6, 6), "roe", 1, 0.4, 3.0),
    (2, (16, 6, 4), "roe", 1, 0.55, 0.5),
    (1, (12, 4, 4), "hllc", 1, 0.45, 1.0),
]
SOLVE = dict(max_iter=200, tol=1e-10, krylov_m=20, max_restarts=4, rtol=1e-3, precond=cfd.PREC_ILU0, mixed=True)


@pytest.fixture(scope="module")
def alone_runs(ctx):
    """Each request of REQS solved by itself: (iterations, Krylov total, forces, final state)."""
    meshes = {k: cmesh.bump_channel(*d) for k, d, *_ in REQS}
    alone = []
    for k, d, flux, order, mach, alpha in REQS:
        p = cfd.make_params(mach=mach, alpha_deg=alpha, flux=flux, order=order)
        dm = cfd.DeviceMesh(ctx, meshes[k], p)
        r, _ = dm.solve_steady(p, alpha_deg=alpha, **SOLVE)
        alone.append((r.iterations, r.krylov_total, list(r.force), dm.get(cfd.FIELD_Q)))
        dm.close()
    return meshes, alone


def test_server_results_match_alone(ctx, alone_runs):
    meshes, alone = alone_runs
    srv = cfd.Server(ctx, hbm_budget=8 << 30, max_inflight=3)
    for rid, (k, d, flux, order, mach, alpha) in enumerate(REQS):
        p = cfd.make_params(mach=mach, alpha_deg=alpha, flux=flux, order=order)
        srv.submit(rid, k, meshes[k], p, want_state=True, alpha_deg=alpha, **SOLVE)
    done = []
    while len(done) < len(REQS):
        got = srv.poll(wait=60.0)
        assert got, "server made no progress in 60 s"
        done += got
    # with every mesh idle, a request on a known geometry runs on its resident mesh
    k, d, flux, order, mach, alpha = REQS[0]
    srv.submit(len(REQS), k, meshes[k], cfd.make_params(mach=mach, alpha_deg=alpha, flux=flux, order=order),
               want_state=True, alpha_deg=alpha, **SOLVE)
    again = srv.poll(wait=60.0)
    assert len(again) == 1 and again[0]["reused"]
    again[0]["id"] = 0
    done += again
    st = srv.stats()
    srv.close()
    assert st["reused"] >= 1
    for c in done:
        it, kry, force, q = alone[c["id"]]
        assert c["status"] == 0, c["message"]
        assert c["result"].converged
        assert c["result"].iterations == it and c["result"].krylov_total == kry
        assert list(c["result"].force) == force
        assert np.array_equal(c["q"], q), "request %d differs from its alone run" % c["id"]
        assert c["t_submit"] <= c["t_admit"] <= c["t_done"]


class _env:
    """Sets environment variables for the duration of a with block (read by the library at each call)."""

    def __init__(self, **kv):
        self.kv, self.old = kv, {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = v

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("compact", ["0", "1"])
def test_server_batches_match_alone(ctx, alone_runs, compact):
    """Continuous batching: with one worker busy on the first request, the queued compatible requests are solved
    together as union meshes, and every result equals the same request solved alone, bit for bit -- with each
    batch compacted to its unfinished cases as soon as they hold half of its nodes or less (the default, the
    finished ones being posted at that point) and without compaction (CFD_BATCH_COMPACT=0)."""
    meshes, alone = alone_runs
    with _env(CFD_BATCH_COMPACT=compact):
        srv = cfd.Server(ctx, hbm_budget=8 << 30, max_inflight=1, batch_max=4)
        for rid, (k, d, flux, order, mach, alpha) in enumerate(REQS):
            p = cfd.make_params(mach=mach, alpha_deg=alpha, flux=flux, order=order)
            srv.submit(rid, k, meshes[k], p, want_state=True, alpha_deg=alpha, **SOLVE)
        done = []
        while len(done) < len(REQS):
            got = srv.poll(wait=60.0)
            assert got, "server made no progress in 60 s"
            done += got
        st = srv.stats()
        srv.close()
    assert max(c["batch_size"] for c in done) >= 2
    assert st["admitted"] == len(REQS) and st["peak_bytes"] <= 8 << 30
    for c in done:
        it, kry, force, q = alone[c["id"]]
        assert c["status"] == 0, c["message"]
        assert c["result"].converged
        assert c["result"].iterations == it and c["result"].krylov_total == kry
        assert list(c["result"].force) == force
        assert np.array_equal(c["q"], q), "request %d differs from its alone run" % c["id"]
        assert 0 < c["device_bytes"] <= c["footprint"], (c["device_bytes"], c["footprint"])


def test_server_batches_only_when_requests_outnumber_workers(ctx, alone_runs):
    """Adaptive batching: with as many workers as requests every request runs alone; with one worker the queued
    requests are batched, except those on meshes above batch_case_nodes.  Results stay bit-identical."""
    meshes, alone = alone_runs
    big_key = max(meshes, key=lambda k: meshes[k].n_node)
    for workers, case_nodes in ((len(REQS), 0), (1, meshes[big_key].n_node - 1)):
        srv = cfd.Server(ctx, hbm_budget=8 << 30, max_inflight=workers, batch_max=4, batch_case_nodes=case_nodes)
        for rid, (k, d, flux, order, mach, alpha) in enumerate(REQS):
            p = cfd.make_params(mach=mach, alpha_deg=alpha, flux=flux, order=order)
            srv.submit(rid, k, meshes[k], p, want_state=True, alpha_deg=alpha, **SOLVE)
        done = []
        while len(done) < len(REQS):
            got = srv.poll(wait=60.0)
            assert got, "server made no progress in 60 s"
            done += got
        srv.close()
        for c in done:
            it, kry, force, q = alone[c["id"]]
            assert c["status"] == 0, c["message"]
            assert (c["result"].iterations, c["result"].krylov_total) == (it, kry)
            assert np.array_equal(c["q"], q), "request %d differs from its alone run" % c["id"]
            if workers == len(REQS) or REQS[c["id"]][0] == big_key:
                assert c["batch_size"] == 1, (workers, c["id"], c["batch_size"])
        if workers == 1:
            assert max(c["batch_size"] for c in done) >= 2


def test_server_admission_respects_budget(ctx):
    m = cmesh.bump_channel(16, 6, 6)
    need = cfd.footprint_bytes(m.n_node, m.n_edge, m.n_bedge, krylov_m=20, mixed=True)
    budget = int(2.5 * need)  # room for two meshes at a time
    srv = cfd.Server(ctx, hbm_budget=budget, max_inflight=4, keep_resident=False)
    for rid in range(6):
        p = cfd.make_params(mach=0.3 + 0.05 * rid, alpha_deg=0.5 * rid, flux="roe", order=1)
        srv.submit(rid, 100 + rid, m, p, alpha_deg=0.5 * rid, **SOLVE)
    done = []
    while len(done) < 6:
        got = srv.poll(wait=60.0)
        assert got, "server made no progress in 60 s"
        done += got
    st = srv.stats()
    srv.close()
    assert st["peak_bytes"] <= budget
    for c in done:
        assert c["status"] == 0, c["message"]
        assert 0 < c["device_bytes"] <= c["footprint"], (c["device_bytes"], c["footprint"])


def test_server_rejects_request_larger_than_budget(ctx):
    m = cmesh.bump_channel(12, 4, 4)
    srv = cfd.Server(ctx, hbm_budget=1 << 20, max_inflight=1)
    p = cfd.make_params(mach=0.5)
    with pytest.raises(cfd.CfdError) as e:
        srv.submit(0, 1, m, p, **SOLVE)
    assert e.value.code == cfd.E_OOM
    srv.close()


@pytest.mark.parametrize("precond", [cfd.PREC_ILU0, cfd.PREC_ILU0_MC])
def test_batch_solve_matches_alone(ctx, precond):
    """Cases solved together in one union mesh reproduce, bit for bit, the iteration counts, residual norms and
    forces of each case solved alone."""
    cases = [((12, 4, 4), 0.08, 0.5, 0.0), ((16, 6, 4), 0.12, 0.6, 2.0), ((14, 5, 5), 0.1, 0.45, -1.0),
             ((12, 4, 4), 0.1, 0.7, 3.0)]
    meshes = [cmesh.bump_channel(*d, bump_h=h) for d, h, _, _ in cases]
    params = [cfd.make_params(mach=mach, alpha_deg=a, flux="roe", order=1) for _, _, mach, a in cases]
    kw = dict(SOLVE)
    kw["precond"] = precond
    alone = []
    for m, p, (_, _, _, a) in zip(meshes, params, cases):
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, **kw)
        alone.append(r)
        dm.close()
    rc, batch = cfd.batch_solve(ctx, meshes, params, alphas=[a for *_, a in cases], **kw)
    assert rc == 0, lib_error(ctx)
    for r0, r1 in zip(alone, batch):
        assert r1.converged and r1.status == 0
        assert (r1.iterations, r1.krylov_total) == (r0.iterations, r0.krylov_total)
        assert r1.first_rms == r0.first_rms and r1.final_rms == r0.final_rms
        assert list(r1.force) == list(r0.force)
        assert r1.lift == r0.lift and r1.drag == r0.drag


def test_batch_long_krylov_matches_alone(ctx):
    """FGMRES(40) with a linear tolerance the FP32 basis cannot reach (1e-12): the linear solves average 26-33
    Krylov steps, and a case averaging more than 32 has solves past 32 steps, where the per-case Gram-Schmidt
    passes use their widest FP32 instantiation that keeps the basis values in LDS (33-41 vectors).  The batched
    cases reproduce their alone solves bit for bit."""
    print("batch vs alone, FGMRES(40), linear tolerance 1e-12", flush=True)
    cases = [((12, 4, 4), 0.08, 0.5, 0.0), ((14, 5, 5), 0.1, 0.45, -1.0), ((16, 6, 4), 0.12, 0.6, 2.0)]
    meshes = [cmesh.bump_channel(*d, bump_h=h) for d, h, _, _ in cases]
    params = [cfd.make_params(mach=mach, alpha_deg=a, flux="roe", order=1) for _, _, mach, a in cases]
    kw = dict(max_iter=60, tol=1e-10, krylov_m=40, max_restarts=1, rtol=1e-12, precond=cfd.PREC_ILU0_MC,
              mixed=True)
    alone = []
    for m, p, (_, _, _, a) in zip(meshes, params, cases):
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, **kw)
        alone.append(r)
        dm.close()
    assert any(r.krylov_total > 32 * r.iterations for r in alone)  # some linear solve past 32 steps
    rc, batch = cfd.batch_solve(ctx, meshes, params, alphas=[a for *_, a in cases], **kw)
    assert rc == 0, lib_error(ctx)
    for r0, r1 in zip(alone, batch):
        print("  alone: %d Newton, %d Krylov" % (r0.iterations, r0.krylov_total), flush=True)
        assert r0.krylov_total > 20 * r0.iterations  # long linear solves
        assert r1.status == r0.status and r1.converged == r0.converged
        assert (r1.iterations, r1.krylov_total) == (r0.iterations, r0.krylov_total)
        assert r1.first_rms == r0.first_rms and r1.final_rms == r0.final_rms
        assert list(r1.force) == list(r0.force)


def test_batch_solve_compaction_matches_alone(ctx):
    """Compaction: whenever the cases still iterating hold half of the union's nodes or less, they move to a new
    union of their own.  Every case still reproduces its alone solve bit for bit; the finished cases are
    harvested at the compaction, so their reported times are shorter than the batch's."""
    cases = [((20, 6, 6), 0.08, 0.35, 0.0), ((16, 6, 4), 0.12, 0.45, 2.0), ((14, 5, 5), 0.1, 0.5, -1.0),
             ((12, 4, 4), 0.1, 0.75, 3.0), ((12, 4, 4), 0.06, 0.7, 1.0)]
    meshes = [cmesh.bump_channel(*d, bump_h=h) for d, h, _, _ in cases]
    params = [cfd.make_params(mach=mach, alpha_deg=a, flux="roe", order=1) for _, _, mach, a in cases]
    kw = dict(SOLVE)
    kw["precond"] = cfd.PREC_ILU0_MC
    alone = []
    for m, p, (_, _, _, a) in zip(meshes, params, cases):
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, **kw)
        alone.append(r)
        dm.close()
    with _env(CFD_BATCH_COMPACT="1"):
        rc, batch = cfd.batch_solve(ctx, meshes, params, alphas=[a for *_, a in cases], **kw)
    assert rc == 0, lib_error(ctx)
    for r0, r1 in zip(alone, batch):
        assert r1.converged and r1.status == 0
        assert (r1.iterations, r1.krylov_total) == (r0.iterations, r0.krylov_total)
        assert r1.first_rms == r0.first_rms and r1.final_rms == r0.final_rms
        assert list(r1.force) == list(r0.force)
        assert r1.lift == r0.lift and r1.drag == r0.drag
    assert len({r.iterations for r in alone}) > 1
    assert min(r.seconds for r in batch) < max(r.seconds for r in batch), "no compaction took place"


def lib_error(ctx):
    return cfd.lib().cfd_last_error(ctx.h).decode()
