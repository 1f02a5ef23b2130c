"""Failure paths on the GPU, forced from the host (every kernel still exits
normally; the failures are numerical, raised by the library's own checks).

- A steady solve reports CFD_E_NONFINITE for a non-finite residual and
  CFD_E_NEGATIVE_PRESSURE for a negative ghost pressure, as the reference's
  NaN abort (Euler3D/src/Solver_Steady_Implicit.cpp:127-133) and its
  boundary-condition error do, instead of exiting.
- A failing case of a batched solve fails alone, with the status and
  iteration count of its solve alone, and every other case of the batch
  still equals its alone solve bit for bit.
- The server posts a failing request's error and keeps serving the others,
  batched or not.
- Alternating FP64 and mixed solves on one mesh keeps its footprint flat.
"""
import numpy as np
import pytest

from cfdsolutions_amd import cfd
from cfdsolutions_amd import mesh as cmesh

pytestmark = pytest.mark.gpu

SOLVE = dict(max_iter=200, tol=1e-10, krylov_m=20, max_restarts=4, rtol=1e-3, precond=cfd.PREC_ILU0_MC, mixed=True)


@pytest.fixture(scope="module")
def ctx():
    c = cfd.Context(0)
    yield c
    c.close()


def bump_fast():
    """Mach 2 over a bump of height 0.15: the characteristic wall condition on the bump's lee side drives the
    ghost pressure below zero in the first iteration (p_b = p + rho c u_n with u_n < -c / gamma there)."""
    return cmesh.bump_channel(12, 4, 3, bump_h=0.15), cfd.make_params(mach=2.0, alpha_deg=0.0, flux="roe", order=1)


def nan_params():
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    p.inf_rho = float("nan")
    return p


def test_steady_solve_reports_nonfinite_residual(ctx):
    print("failure: NaN state in a steady solve", flush=True)
    m = cmesh.bump_channel(12, 4, 4)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    dm = cfd.DeviceMesh(ctx, m, p)
    q = dm.get(cfd.FIELD_Q)
    interior = int(np.setdiff1d(np.arange(m.n_node), m.bedge[:, 0])[0])
    q[0, interior] = np.nan
    dm.set(cfd.FIELD_Q, q)
    r, _ = dm.solve_steady(p, check=False, **SOLVE)
    assert r.status == cfd.E_NONFINITE and r.iterations == 0
    # the handle stays usable: a reset mesh solves normally
    dm.reset(p)
    r2, _ = dm.solve_steady(p, check=False, **SOLVE)
    assert r2.status == 0 and r2.converged
    dm.close()


def test_steady_solve_reports_negative_ghost_pressure(ctx):
    print("failure: negative ghost pressure in a steady solve", flush=True)
    m, p = bump_fast()
    dm = cfd.DeviceMesh(ctx, m, p)
    r, _ = dm.solve_steady(p, check=False, **SOLVE)
    assert r.status == cfd.E_NEGATIVE_PRESSURE and r.iterations == 0
    assert "negative pressure" in cfd.lib().cfd_last_error(ctx.h).decode()
    dm.close()


def _alone(ctx, m, p, alpha):
    dm = cfd.DeviceMesh(ctx, m, p)
    r, _ = dm.solve_steady(p, alpha_deg=alpha, check=False, **SOLVE)
    dm.close()
    return r


def test_batch_isolates_failing_cases(ctx):
    """Two failing cases (a negative ghost pressure, a non-finite freestream) among three healthy ones: the batch
    succeeds, each failing case carries the status and iteration count of its own solve, and each healthy case
    equals its alone solve bit for bit."""
    print("failure: batch with failing cases", flush=True)
    fast_m, fast_p = bump_fast()
    cases = [(cmesh.bump_channel(12, 4, 4, bump_h=0.08), cfd.make_params(mach=0.5, alpha_deg=0.0), 0.0),
             (fast_m, fast_p, 0.0),
             (cmesh.bump_channel(14, 5, 5, bump_h=0.1), cfd.make_params(mach=0.45, alpha_deg=-1.0), -1.0),
             (cmesh.bump_channel(12, 4, 4, bump_h=0.1), nan_params(), 1.0),
             (cmesh.bump_channel(16, 6, 4, bump_h=0.12), cfd.make_params(mach=0.6, alpha_deg=2.0), 2.0)]
    alone = [_alone(ctx, m, p, a) for m, p, a in cases]
    assert [r.status for r in alone] == [0, cfd.E_NEGATIVE_PRESSURE, 0, cfd.E_ARG, 0]
    rc, batch = cfd.batch_solve(ctx, [c[0] for c in cases], [c[1] for c in cases], alphas=[c[2] for c in cases],
                                **SOLVE)
    assert rc == 0, cfd.lib().cfd_last_error(ctx.h).decode()
    for r0, r1 in zip(alone, batch):
        assert r1.status == r0.status
        assert r1.iterations == r0.iterations
        if r0.status:
            continue
        assert r1.converged
        assert r1.krylov_total == r0.krylov_total
        assert r1.first_rms == r0.first_rms and r1.final_rms == r0.final_rms
        assert list(r1.force) == list(r0.force)
        assert r1.lift == r0.lift and r1.drag == r0.drag


@pytest.mark.parametrize("batched", [False, True])
def test_server_isolates_failing_request(ctx, batched):
    """A request that fails completes with its error, the server keeps serving, and every other request's result
    equals its alone solve bit for bit (requests solved one per worker, or batched behind a busy worker)."""
    print("failure: server with a failing request, batched", batched, flush=True)
    fast_m, fast_p = bump_fast()
    reqs = [(1, cmesh.bump_channel(12, 4, 4), cfd.make_params(mach=0.5, alpha_deg=0.0), 0.0),
            (2, fast_m, fast_p, 0.0),
            (3, cmesh.bump_channel(14, 5, 4), cfd.make_params(mach=0.55, alpha_deg=1.0), 1.0),
            (4, cmesh.bump_channel(12, 4, 4, bump_h=0.08), nan_params(), 1.0),
            (5, cmesh.bump_channel(16, 5, 4), cfd.make_params(mach=0.4, alpha_deg=-1.0), -1.0)]
    alone = {}
    for k, m, p, a in reqs:
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, check=False, **SOLVE)
        alone[k] = (r.status, r.iterations, r.krylov_total, list(r.force), dm.get(cfd.FIELD_Q) if not r.status else None)
        dm.close()
    srv = cfd.Server(ctx, hbm_budget=8 << 30, max_inflight=1 if batched else 2, batch_max=4 if batched else 0)
    for k, m, p, a in reqs:
        srv.submit(k, k, m, p, want_state=True, alpha_deg=a, **SOLVE)
    done = []
    while len(done) < len(reqs):
        got = srv.poll(wait=60.0)
        assert got, "server made no progress in 60 s"
        done += got
    srv.close()
    if batched:
        assert max(c["batch_size"] for c in done) >= 2
    for c in done:
        st, it, kry, force, q = alone[c["id"]]
        assert c["status"] == st, (c["id"], c["status"], c["message"])
        if st:
            assert c["message"]
            continue
        assert (c["result"].iterations, c["result"].krylov_total) == (it, kry)
        assert list(c["result"].force) == force
        assert np.array_equal(c["q"], q), "request %d differs from its alone run" % c["id"]


def test_footprint_flat_when_precision_alternates(ctx):
    """Each precision keeps one Krylov basis: FP64 and mixed solves in turn allocate nothing after the first of
    each, and a longer restart replaces the basis instead of adding one."""
    print("footprint with alternating precision", flush=True)
    m = cmesh.bump_channel(12, 4, 4)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0)
    dm = cfd.DeviceMesh(ctx, m, p)
    sizes = []
    for mixed in (False, True, False, True, False):
        dm.reset(p)
        r, _ = dm.solve_steady(p, **dict(SOLVE, mixed=mixed))
        assert r.converged
        sizes.append(dm.device_bytes())
    assert sizes[1] == sizes[2] == sizes[3] == sizes[4], sizes
    dm.reset(p)
    dm.solve_steady(p, **dict(SOLVE, mixed=True, krylov_m=40))
    grown = dm.device_bytes()
    L = 5 * m.n_node
    assert grown - sizes[-1] == 4 * L * 2 * (40 - 20), (grown, sizes[-1])
    dm.close()
