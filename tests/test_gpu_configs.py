"""The batched and serving workloads at their real case sizes, on the GPU.

- 32 cases drawn from the batched-small-cases distribution (bump channels of
  8k-40k nodes, log-uniform; bump height U(0.05, 0.15); first-order Roe;
  Mach U(0.3, 0.8); AoA U(-2, 4) deg; RMS < 1e-10): the union-mesh batch
  reproduces each case's alone solve bit for bit (iterations, Krylov total,
  residual norms, forces, lift and drag).
- A request of more than 200,000 nodes served alongside small requests
  returns exactly the state, iteration counts and forces of its solve alone,
  and so does every small request.
"""
import numpy as np
import pytest

from cfdsolutions_amd import cfd
from cfdsolutions_amd import mesh as cmesh

pytestmark = pytest.mark.gpu

SOLVE = dict(max_iter=400, tol=1e-10, krylov_m=30, max_restarts=4, rtol=1e-3, precond=cfd.PREC_ILU0_MC, mixed=True)


@pytest.fixture(scope="module")
def ctx():
    c = cfd.Context(0)
    yield c
    c.close()


def bump_dims(n_nodes):
    """Hex counts (3n, n, n) of the bump channel whose node count is closest to n_nodes."""
    n = max(4, int(round((n_nodes / 3.0) ** (1.0 / 3.0))))
    return 3 * n, n, n


def test_batch_of_32_cases_matches_alone(ctx):
    print("batch of 32 cases vs alone", flush=True)
    rng = np.random.default_rng(11)
    n = 32
    sizes = np.exp(rng.uniform(np.log(8000), np.log(40000), n))
    meshes = [cmesh.bump_channel(*bump_dims(s), bump_h=h) for s, h in zip(sizes, rng.uniform(0.05, 0.15, n))]
    machs, alphas = rng.uniform(0.3, 0.8, n), rng.uniform(-2.0, 4.0, n)
    params = [cfd.make_params(mach=mk, alpha_deg=a, flux="roe", order=1) for mk, a in zip(machs, alphas)]
    assert min(m.n_node for m in meshes) >= 6000 and max(m.n_node for m in meshes) <= 50000
    alone = []
    for m, p, a in zip(meshes, params, alphas):
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, **SOLVE)
        alone.append(r)
        dm.close()
    rc, batch = cfd.batch_solve(ctx, meshes, params, alphas=list(alphas), **SOLVE)
    assert rc == 0, cfd.lib().cfd_last_error(ctx.h).decode()
    for k, (r0, r1) in enumerate(zip(alone, batch)):
        assert r0.converged and r1.converged and r1.status == 0, k
        assert (r1.iterations, r1.krylov_total) == (r0.iterations, r0.krylov_total), k
        assert r1.first_rms == r0.first_rms and r1.final_rms == r0.final_rms, k
        assert list(r1.force) == list(r0.force), k
        assert r1.lift == r0.lift and r1.drag == r0.drag, k


def test_large_request_served_alone_and_in_the_mix(ctx):
    print("serving a 203k-node request among small ones", flush=True)
    big = cmesh.bump_channel(120, 40, 40, bump_h=0.1)
    assert big.n_node > 200000
    reqs = [(1, big, cfd.make_params(mach=0.6, alpha_deg=2.0, flux="roe", order=1), 2.0),
            (2, cmesh.bump_channel(24, 8, 8), cfd.make_params(mach=0.5, alpha_deg=0.0, flux="roe", order=1), 0.0),
            (3, cmesh.bump_channel(30, 10, 10, bump_h=0.08),
             cfd.make_params(mach=0.7, alpha_deg=-1.0, flux="hllc", order=1), -1.0),
            (4, cmesh.bump_channel(24, 8, 8), cfd.make_params(mach=0.4, alpha_deg=3.0, flux="roe", order=1), 3.0)]
    alone = {}
    for k, m, p, a in reqs:
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, **SOLVE)
        assert r.converged, k
        alone[k] = (r.iterations, r.krylov_total, list(r.force), dm.get(cfd.FIELD_Q))
        dm.close()
    srv = cfd.Server(ctx, hbm_budget=16 << 30, max_inflight=3, batch_max=4)
    for k, m, p, a in reqs:
        srv.submit(k, k, m, p, want_state=True, alpha_deg=a, **SOLVE)
    done = []
    while len(done) < len(reqs):
        got = srv.poll(wait=120.0)
        assert got, "server made no progress in 120 s"
        done += got
    srv.close()
    for c in done:
        it, kry, force, q = alone[c["id"]]
        assert c["status"] == 0, c["message"]
        assert (c["result"].iterations, c["result"].krylov_total) == (it, kry), c["id"]
        assert list(c["result"].force) == force, c["id"]
        assert np.array_equal(c["q"], q), "request %d differs from its alone run" % c["id"]
