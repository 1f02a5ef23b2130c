#!/usr/bin/env python3
"""Residual-evaluation throughput and converged-solve throughput on MI355X
for the implicit unstructured finite-volume Euler hot path.

Headline (`value`): the 160^3 jittered-cube residual microbenchmark.  One
step = Green-Gauss gradient + Venkatakrishnan limiter (K = 5, one fused
kernel) + second-order Roe residual with the Conv/Diss split, FP64, on a
160^3 hex grid split into 24.6M tetrahedra (4.17M nodes), interior nodes
jittered by U(-0.2h, 0.2h), RCM-ordered, with a random per-node state (Mach
U(0.3, 2.0) along +x, density and pressure with 5% Gaussian noise).  Inputs
are resident in HBM before timing starts.  The headline runs the fast build
(FMA contraction, shared reciprocals, the flux-once residual over work-edge
patches), whose FP64 results stay within 1e-12 relative of the reference;
the exact build (the reference's operation order, bit-identical results) is
reported beside it.  value = algorithmic bytes of G+L+R2 (each array once
per kernel, as if gradient and limiter ran separately) / wall time per step.
The roofline block reports the residual kernel alone from HIP events on the
stream it runs on.  cpu_baseline times the CPU oracle (single-threaded
restatement of the reference) on the same mesh and state and checks both
GPU builds against it.

Also measured: HLLC and AUSM+, FP32, unfused gradient/limiter and first-order
variants ("variants"); converged steady solves of the 48x16x16 and 192x64x64
bump channels, plain FP64 vs mixed precision ("solves", "solve_192x64x64");
the serving loop against the plain FP64 path on a stream of solve requests
("serving"); and 512 small cases solved as one batch ("batched_small_cases").
The serving and batch legs' mixed solves use mixed = 2 by default (FP32
Jacobian, factors and Krylov basis, with the block-SELL copies of the
multicolour ILU(0) factors in FP16); --mixed-level 1 keeps FP32, 3 also puts
the Jacobian's block-SELL copy in FP16.
"""
import argparse
import ctypes
import json
import os
import sys
import time

# The serving leg overlaps many solves, one HIP stream each; HIP maps a
# process's streams onto GPU_MAX_HW_QUEUES hardware queues (4 unless set).
# The runtime's default of 4 is kept: 8 or 16 queues served more slowly.

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


class HipEvents:
    """hipEvent timing on an explicit stream (the mesh handle's stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")

    def create(self):
        e = ctypes.c_void_p()
        assert self.h.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = ctypes.c_float()
        assert self.h.hipEventElapsedTime(ctypes.byref(t), a, b) == 0
        return t.value

    def destroy(self, e):
        self.h.hipEventDestroy(e)


def algorithmic_bytes(N, E, B, s, i=4):
    """Compulsory traffic per evaluation, each array counted once per kernel."""
    R1 = E * (2 * i + 3 * s) + N * (5 * s + 5 * s + s) + B * (3 * i + 3 * s + 5 * s)
    G = E * (2 * i + 3 * s) + N * (5 * s + s + 15 * s) + B * (2 * i + 3 * s + 5 * s)
    L = E * (2 * i) + N * (5 * s + 15 * s + 3 * s + s + 5 * s)
    R2 = E * (2 * i + 3 * s) + N * (5 * s + 15 * s + 5 * s + 3 * s + 5 * s + s) + B * (3 * i + 3 * s + 5 * s)
    return dict(R1=R1, R1split=R1 + N * 5 * s, G=G, L=L, R2=R2, R2split=R2 + N * 5 * s)


def make_state(m, p, seed=20261202):
    """Random per-node state; ghosts at the freestream of p."""
    g = p.gamma
    rng = np.random.default_rng(seed)
    n = m.n_node
    mach = rng.uniform(0.3, 2.0, n)
    rho = 1.0 + 0.05 * np.clip(rng.standard_normal(n), -3, 3)
    pr = p.inf_p * (1.0 + 0.05 * np.clip(rng.standard_normal(n), -3, 3))
    c = np.sqrt(g * pr / rho)
    u = mach * c
    q = np.zeros((5, m.n_total))
    q[0, :n] = rho
    q[1, :n] = rho * u
    q[4, :n] = pr / (g - 1.0) + 0.5 * rho * u * u
    et = p.inf_p / ((g - 1.0) * p.inf_rho) + 0.5 * (p.inf_u ** 2 + p.inf_v ** 2 + p.inf_w ** 2)
    q[0, n:] = p.inf_rho
    q[1, n:] = p.inf_rho * p.inf_u
    q[2, n:] = p.inf_rho * p.inf_v
    q[3, n:] = p.inf_rho * p.inf_w
    q[4, n:] = p.inf_rho * et
    return q


def time_variant(dm, ev, p, steps, warmup, second_order=True, fused=True):
    """Returns (wall ms per step, mean ms of G, L, R kernels, median ms of one step from its events).  With
    fused=True the gradient and limiter run as
    one kernel (cfd_gradient_limiter) and its time is reported under G, with L = 0."""
    stream = dm.stream()

    def one(evs=None):
        if evs:
            ev.record(evs[0], stream)
        if second_order and fused:
            dm.gradient_limiter(p)
            if evs:
                ev.record(evs[1], stream)
                ev.record(evs[2], stream)
        elif second_order:
            dm.gradient(p)
            if evs:
                ev.record(evs[1], stream)
            dm.limiter(p)
            if evs:
                ev.record(evs[2], stream)
        elif evs:
            ev.record(evs[1], stream)
            ev.record(evs[2], stream)
        dm.residual(p)
        if evs:
            ev.record(evs[3], stream)

    for _ in range(warmup):
        one()
    dm.sync()
    evsets = [[ev.create() for _ in range(4)] for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        one(evsets[k])
    dm.sync()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    g = l = r = 0.0
    per_step = []
    for e in evsets:
        g += ev.elapsed_ms(e[0], e[1])
        l += ev.elapsed_ms(e[1], e[2])
        r += ev.elapsed_ms(e[2], e[3])
        per_step.append(ev.elapsed_ms(e[0], e[3]))
        for x in e:
            ev.destroy(x)
    return wall, g / steps, l / steps, r / steps, float(np.median(per_step))


def cpu_baseline(m, p, q, build="parity"):
    """One G+L+R2 Roe evaluation by the oracle on one core: build "parity" (-O2 -ffp-contract=off, the checker) or
    "speed" (the same source at -O3 -march=x86-64-v3)."""
    from oracle import oracle as orc

    old = orc.use_build(build)
    try:
        om = orc.Mesh(m.xyz, m.edge, m.edge_area, m.bedge, m.bedge_area, m.bedge_type, m.vol, m.n2n_ia, m.n2n_ja,
                      m.bedge_tag)
        op = orc.params({"rho": p.inf_rho, "u": p.inf_u, "v": p.inf_v, "w": p.inf_w, "p": p.inf_p,
                         "outflow_p": p.outflow_p}, flux=orc.ROE, order=2, limiter=orc.LIM_VENKAT,
                        venkat_k=p.venkat_k)
        t0 = time.perf_counter()
        g = orc.gg_gradient(om, q)
        phi = orc.limiter(om, op, q, g)
        rc, rd, _, _ = orc.residual(om, op, q, g, phi)
        return time.perf_counter() - t0, rc, rd
    finally:
        orc.use_build(old)


def cpu_solve(m, p, build="speed"):
    """The 48x16x16 steady solve by the oracle on one core (same algorithm as the GPU FP64 path: backward Euler,
    approximate Roe Jacobian, FGMRES(30) + natural-order ILU(0))."""
    from oracle import oracle as orc

    old = orc.use_build(build)
    try:
        om = orc.Mesh(m.xyz, m.edge, m.edge_area, m.bedge, m.bedge_area, m.bedge_type, m.vol, m.n2n_ia, m.n2n_ja,
                      m.bedge_tag)
        op = orc.params({"rho": p.inf_rho, "u": p.inf_u, "v": p.inf_v, "w": p.inf_w, "p": p.inf_p,
                         "outflow_p": p.outflow_p}, flux=orc.ROE, order=1, limiter=orc.LIM_NONE,
                        venkat_k=p.venkat_k, cfl_min=p.cfl_min, cfl_max=p.cfl_max, cfl_ramp=p.cfl_ramp)
        et = p.inf_p / ((p.gamma - 1) * p.inf_rho) + 0.5 * (p.inf_u ** 2 + p.inf_v ** 2 + p.inf_w ** 2)
        q = np.zeros((5, om.n_total))
        q[:] = np.array([p.inf_rho, p.inf_rho * p.inf_u, p.inf_rho * p.inf_v, p.inf_rho * p.inf_w,
                         p.inf_rho * et])[:, None]
        t0 = time.perf_counter()
        _, r, _, _, _ = orc.solve_steady(om, op, q, m.bedge_tag, max_iter=SOLVE_KW["max_iter"], tol=SOLVE_KW["tol"],
                                         krylov_m=SOLVE_KW["krylov_m"], max_restarts=SOLVE_KW["max_restarts"],
                                         krylov_rtol=SOLVE_KW["rtol"], precond=2, gradient=orc.GRAD_LSQ)
        return time.perf_counter() - t0, r
    finally:
        orc.use_build(old)


def measured_copy_gbs(mib=1024, reps=20):
    """Device-to-device copy bandwidth of one MIB MiB buffer, (read + write bytes) / time, through the HIP runtime
    (ctypes), timed with HIP events around `reps` copies."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    hip.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    hip.hipFree.argtypes = [vp]
    hip.hipMemset.argtypes = [vp, ctypes.c_int, sz]
    hip.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
    hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
    hip.hipEventRecord.argtypes = [vp, vp]
    hip.hipEventSynchronize.argtypes = [vp]
    hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
    hip.hipEventDestroy.argtypes = [vp]
    n = mib << 20
    a, b, e0, e1 = vp(), vp(), vp(), vp()
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        return None
    try:
        hip.hipMemset(a, 0, n)
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        for _ in range(3):
            hip.hipMemcpyAsync(b, a, n, 3, None)  # hipMemcpyDeviceToDevice
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            hip.hipMemcpyAsync(b, a, n, 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
        return 2.0 * n * reps / (ms.value * 1e-3) / 1e9
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 160^3 benchmark, from the committed rocprofv3 counter summary
    (2 x FETCH_SIZE + WRITE_SIZE, calibrated against 1 GiB device copies); None if it is not there."""
    import re
    path = os.path.join(ROOT, "profiles", "residual_cube160_fp64_hbm_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"- %s .*= ([0-9.]+) GB against" % re.escape(kernel), line)
        if m:
            return int(float(m.group(1)) * 1e9)
    return None


def bump_dims(n_nodes):
    """Bump-channel hex counts (3n, n, n) whose node count is closest to n_nodes."""
    n = max(4, int(round((n_nodes / 3.0) ** (1.0 / 3.0))))
    return 3 * n, n, n


SOLVE_KW = dict(max_iter=400, tol=1e-10, krylov_m=30, max_restarts=4, rtol=1e-3)
MIXED_LEVEL = 2  # mixed= of the serving and batch legs' mixed solves (--mixed-level)


def solve_leg(ctx, cpu=True):
    """Bump channel 48x16x16 (14,161 nodes), first-order Roe, M 0.5, AoA 0, backward Euler, CFL 1 -> 1000,
    FGMRES(30) + ILU(0), to RMS 1e-10: plain FP64 vs mixed precision, each alone on the GPU."""
    from cfdsolutions_amd import cfd, mesh

    m = mesh.bump_channel(48, 16, 16, bump_h=0.1)
    p = cfd.make_params(mach=0.5, alpha_deg=0.0, flux="roe", order=1)
    runs = {"fp64": (cfd.PREC_ILU0, False), "mixed": (cfd.PREC_ILU0, True),
            "mixed_multicolour_ilu": (cfd.PREC_ILU0_MC, True)}
    out = {}
    for name, (precond, mixed) in runs.items():
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, precond=precond, mixed=mixed, **SOLVE_KW)  # warm-up (allocation)
        dm.reset(p)
        r, hist = dm.solve_steady(p, precond=precond, mixed=mixed, **SOLVE_KW)
        out[name] = (r, dm.get(cfd.FIELD_Q))
        dm.close()
    r0, q0 = out["fp64"]
    res = {"workload": "bump channel 48x16x16 (%d nodes), 1st-order Roe, M 0.5, FGMRES(30)+ILU(0), CFL 1 -> 1000, "
                       "RMS < 1e-10" % m.n_node}
    for name, (r, q) in out.items():
        res[name] = {"solves_per_s": 1.0 / r.seconds, "seconds": r.seconds, "newton": r.iterations,
                     "krylov": r.krylov_total, "converged": bool(r.converged),
                     "force_rel_err_vs_fp64": float(np.max(np.abs(np.array(r.force) - np.array(r0.force))) /
                                                    np.max(np.abs(r0.force))),
                     "state_rel_err_vs_fp64": float(np.max(np.abs(q - q0)) / np.max(np.abs(q0))),
                     "speedup_vs_fp64": r0.seconds / r.seconds}
        print("bench: solve 48x16x16 %-22s %.3f s (%d newton / %d krylov), force err %.2e" % (
            name, r.seconds, r.iterations, r.krylov_total, res[name]["force_rel_err_vs_fp64"]),
            file=sys.stderr, flush=True)
    if cpu:
        secs, rc = cpu_solve(m, p)
        res["cpu_oracle"] = {"seconds": secs, "solves_per_s": 1.0 / secs, "cores": 1, "kind": "port",
                             "build": "oracle.c -O3 -march=x86-64-v3", "newton": rc.iterations,
                             "krylov": rc.krylov_total, "converged": bool(rc.converged),
                             "force_rel_err_vs_gpu_fp64": float(np.max(np.abs(np.array(rc.force[:]) -
                                                                              np.array(r0.force))) /
                                                                np.max(np.abs(r0.force))),
                             "gpu_fp64_speedup": secs / r0.seconds,
                             "gpu_mixed_multicolour_speedup": secs / out["mixed_multicolour_ilu"][0].seconds}
        print("bench: solve 48x16x16 cpu oracle (1 core)        %.3f s (%d newton / %d krylov)" % (
            secs, rc.iterations, rc.krylov_total), file=sys.stderr, flush=True)
    return res


def solve3_leg(ctx):
    """Bump channel 192x64x64 (815,425 nodes), second-order HLLC with Green-Gauss gradients and the
    Venkatakrishnan limiter, M 0.7, AoA 2 deg, FGMRES(40) with the multicolour ILU(0), to RMS 1e-10: plain FP64 vs
    mixed precision (FP32 Jacobian, ILU(0) and Krylov basis), each alone on the GPU, with the approximate Roe
    Jacobian; the mixed solve with the block-SELL copies of the ILU(0) factors stored in FP16 (mixed = 2), and of
    the factors and the Jacobian (mixed = 3); and the mixed solve with the finite-difference Jacobian of the HLLC
    flux."""
    from cfdsolutions_amd import cfd, mesh

    m = mesh.bump_channel(192, 64, 64, bump_h=0.1)
    p = cfd.make_params(mach=0.7, alpha_deg=2.0, flux="hllc", order=2, limiter=cfd.LIM_VENKAT, gradient=cfd.GRAD_GG)
    kw = dict(SOLVE_KW, krylov_m=40, alpha_deg=2.0, precond=cfd.PREC_ILU0_MC)
    out = {}
    runs = (("fp64", False, cfd.SOLVE_JAC_APPROX), ("mixed", True, cfd.SOLVE_JAC_APPROX),
            ("mixed_fp16", cfd.MIXED_FP16, cfd.SOLVE_JAC_APPROX),
            ("mixed_fp16_all", cfd.MIXED_FP16_ALL, cfd.SOLVE_JAC_APPROX), ("mixed_fd", True, cfd.SOLVE_JAC_FD))
    for name, mixed, jac in runs:
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, mixed=mixed, check=False, jacobian=jac, **kw)
        out[name] = (r, dm.get(cfd.FIELD_Q))
        print("bench: solve 192x64x64 %-6s %.2f s (%d newton / %d krylov), converged %d, rms %.2e" % (
            name, r.seconds, r.iterations, r.krylov_total, r.converged, r.final_rms), file=sys.stderr, flush=True)
        dm.close()
    (r0, q0), (r1, q1), (r2, q2) = out["fp64"], out["mixed"], out["mixed_fd"]

    def level(name):
        r, q = out[name]
        return {"seconds": r.seconds, "newton": r.iterations, "krylov": r.krylov_total, "converged": bool(r.converged),
                "lift": r.lift, "drag": r.drag,
                "force_rel_err_vs_fp64": float(np.max(np.abs(np.array(r.force) - np.array(r0.force))) /
                                               np.max(np.abs(r0.force))),
                "state_rel_err_vs_fp64": float(np.max(np.abs(q - q0)) / np.max(np.abs(q0))),
                "speedup_vs_fp64": r0.seconds / r.seconds}
    return {"workload": "bump channel 192x64x64 (%d nodes), 2nd-order HLLC + GG + Venkatakrishnan, M 0.7, AoA 2, "
                        "FGMRES(40) + multicolour ILU(0), RMS < 1e-10" % m.n_node,
            "fp64": {"seconds": r0.seconds, "newton": r0.iterations, "krylov": r0.krylov_total,
                     "converged": bool(r0.converged), "lift": r0.lift, "drag": r0.drag},
            "mixed": {"seconds": r1.seconds, "newton": r1.iterations, "krylov": r1.krylov_total,
                      "converged": bool(r1.converged), "lift": r1.lift, "drag": r1.drag,
                      "force_rel_err_vs_fp64": float(np.max(np.abs(np.array(r1.force) - np.array(r0.force))) /
                                                     np.max(np.abs(r0.force))),
                      "state_rel_err_vs_fp64": float(np.max(np.abs(q1 - q0)) / np.max(np.abs(q0)))},
            "mixed_fd_jacobian": {"seconds": r2.seconds, "newton": r2.iterations, "krylov": r2.krylov_total,
                                  "converged": bool(r2.converged), "lift": r2.lift, "drag": r2.drag,
                                  "force_rel_err_vs_fp64": float(np.max(np.abs(np.array(r2.force) -
                                                                              np.array(r0.force))) /
                                                                 np.max(np.abs(r0.force))),
                                  "state_rel_err_vs_fp64": float(np.max(np.abs(q2 - q0)) / np.max(np.abs(q0)))},
            "mixed_fp16_factors": level("mixed_fp16"), "mixed_fp16_factors_and_matrix": level("mixed_fp16_all"),
            "speedup_mixed": r0.seconds / r1.seconds}


def serve_leg(ctx, n_req, n_meshes, nmin, nmax, inflight, batch_max=8, seed=20261202, mach=(0.3, 0.85),
              aoa=(-2.0, 6.0), fluxes=("roe", "hllc"), stream=True, batch_wait_us=50000, batch_case_nodes=0):
    """Steady-solve requests on bump channels of log-uniform size in [nmin, nmax] nodes, bump height U(0.05,
    0.15), Mach U(0.3, 0.85), AoA U(-2, 6) deg, Roe or HLLC (first order), to RMS 1e-10, FGMRES(30) with the
    multicolour ILU(0).  The plain FP64 GPU path solves them one at a time (meshes resident).  The server solves
    them in mixed precision, first with one request per solve (meshes resident, `inflight` at once) as a burst,
    then with continuous batching (up to batch_max compatible requests per union-mesh solve, a request waiting
    at most batch_wait_us for companions while the GPU is busy) as a burst (capacity) and as a Poisson stream at 80% of
    that capacity (latency)."""
    from cfdsolutions_amd import cfd, mesh

    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_meshes))
    heights = rng.uniform(0.05, 0.15, n_meshes)
    pool = [mesh.bump_channel(*bump_dims(n), bump_h=h) for n, h in zip(sizes, heights)]
    reqs = []
    for _ in range(n_req):
        reqs.append((int(rng.integers(n_meshes)), float(rng.uniform(*mach)), float(rng.uniform(*aoa)),
                     fluxes[int(rng.integers(len(fluxes)))]))

    def params(r):
        return cfd.make_params(mach=r[1], alpha_deg=r[2], flux=r[3], order=1)

    # plain FP64 GPU path, one request at a time, meshes already resident
    dms = [cfd.DeviceMesh(ctx, m, params(reqs[0])) for m in pool]
    t64, res64 = 0.0, []
    for r in reqs:
        p = params(r)
        dms[r[0]].reset(p)
        out, _ = dms[r[0]].solve_steady(p, alpha_deg=r[2], precond=cfd.PREC_ILU0_MC, mixed=False, check=False,
                                        **SOLVE_KW)
        t64 += out.seconds
        res64.append(out)
        print("bench: serve baseline request %d/%d: %d nodes, %.2f s" % (len(res64), n_req, pool[r[0]].n_node,
                                                                         out.seconds), file=sys.stderr, flush=True)
    for d in dms:
        d.close()
    print("bench: serve baseline fp64 %d solves in %.2f s" % (n_req, t64), file=sys.stderr, flush=True)

    def run(srv, arrivals):
        t_start = srv.stats()["clock"]
        done, k, t0 = [], 0, time.perf_counter()
        while len(done) < n_req:
            now = time.perf_counter() - t0
            while k < n_req and arrivals[k] <= now:
                r = reqs[k]
                srv.submit(k, r[0], pool[r[0]], params(r), alpha_deg=r[2], precond=cfd.PREC_ILU0_MC,
                           mixed=MIXED_LEVEL, **SOLVE_KW)
                k += 1
            wait = 0.0 if k >= n_req else max(0.0, arrivals[k] - (time.perf_counter() - t0))
            got = srv.poll(wait=min(wait, 0.05) if k < n_req else 1.0)
            for c in got:
                print("bench: served request %d (%.2f s latency, batch of %d)" % (
                    c["id"], c["t_done"] - c["t_submit"], c["batch_size"]), file=sys.stderr, flush=True)
            done += got
        t_end = max(c["t_done"] for c in done)
        lat = np.array([c["t_done"] - c["t_submit"] for c in done])
        return done, t_end - t_start, lat

    # one request per solve, meshes resident before timing (one solve per mesh)
    srv = cfd.Server(ctx, hbm_budget=200 << 30, max_inflight=inflight)
    for k, m in enumerate(pool):
        srv.submit(-1 - k, k, m, params(reqs[0]), alpha_deg=reqs[0][2], precond=cfd.PREC_ILU0_MC, mixed=MIXED_LEVEL,
                   **SOLVE_KW)
    got = 0
    while got < n_meshes:
        got += len(srv.poll(wait=30.0))
    single, span_1, lat_1 = run(srv, np.zeros(n_req))
    srv.close()
    cap1 = n_req / span_1
    print("bench: serve without batching %.2f solves/s" % cap1, file=sys.stderr, flush=True)

    # continuous batching
    srv = cfd.Server(ctx, hbm_budget=200 << 30, max_inflight=inflight, batch_max=batch_max,
                     batch_wait_us=batch_wait_us, batch_case_nodes=batch_case_nodes)
    burst, span_b, lat_b = run(srv, np.zeros(n_req))
    cap = n_req / span_b
    if stream:
        gaps = rng.exponential(1.0 / (0.8 * cap), n_req)
        stream, span_s, lat_s = run(srv, np.cumsum(gaps) - gaps[0])
    else:
        stream, span_s, lat_s = burst, span_b, lat_b
    st = srv.stats()
    srv.close()
    # a request that the plain FP64 path does not converge within max_iter either is a property of the
    # workload (counted apart); a failure is an error status, or no convergence where FP64 converged
    hard = [c for c in single + burst + stream
            if c["status"] == 0 and not c["result"].converged and not res64[c["id"]].converged]
    bad = [c for c in single + burst + stream
           if c["status"] != 0 or (not c["result"].converged and res64[c["id"]].converged)]
    for c in bad + hard:
        r = reqs[c["id"]]
        print("bench: request %d (%d nodes, %s, M %.3f, AoA %.2f) failed: status %d, %d iterations, rms %.3e, %s; "
              "FP64 alone: status %d, converged %d, %d iterations" % (
                  c["id"], pool[r[0]].n_node, r[3], r[1], r[2], c["status"], c["result"].iterations,
                  c["result"].final_rms, c["message"] or "not converged", res64[c["id"]].status,
                  res64[c["id"]].converged, res64[c["id"]].iterations), file=sys.stderr, flush=True)
    ferr = 0.0  # over the requests both paths converged
    for c in single + burst + stream:
        if not (c["status"] == 0 and c["result"].converged and res64[c["id"]].converged):
            continue
        f64 = np.array(res64[c["id"]].force)
        ferr = max(ferr, float(np.max(np.abs(np.array(c["result"].force) - f64)) / np.max(np.abs(f64))))
    base = n_req / t64
    res = {"workload": "%d requests on %d bump channels of %d-%d nodes (log-uniform), 1st-order %s, "
                       "M %.2f-%.2f, AoA %g..%g, RMS < 1e-10, %d in flight, batches of up to %d" % (
                           n_req, n_meshes, min(m.n_node for m in pool), max(m.n_node for m in pool),
                           "/".join(fluxes), mach[0], mach[1], aoa[0], aoa[1], inflight, batch_max),
           "fp64_sequential_solves_per_s": base,
           "serving_capacity_solves_per_s": cap,
           "serving_without_batching_solves_per_s": cap1,
           "mean_batch_size": float(np.mean([c["batch_size"] for c in burst + stream])),
           "serving_vs_fp64": cap / base,
           "mixed_level": int(MIXED_LEVEL),
           "stream_rate_per_s": 0.8 * cap,
           "stream_solves_per_s": n_req / span_s,
           "latency_p50_s": float(np.percentile(lat_s, 50)), "latency_p99_s": float(np.percentile(lat_s, 99)),
           "burst_latency_p50_s": float(np.percentile(lat_b, 50)),
           "burst_latency_p99_s": float(np.percentile(lat_b, 99)),
           "force_rel_err_mixed_vs_fp64": ferr, "failed": len(bad),
           "not_converged_in_fp64_either": len(hard),
           "fp64_failed": sum(1 for r in res64 if r.status != 0 or not r.converged),
           "peak_hbm_bytes": int(st["peak_bytes"]),
           "mesh_reuse": int(st["reused"])}
    print("bench: serve fp64 %.2f solves/s, serving %.2f solves/s (%.2fx), stream p50 %.3f s p99 %.3f s, "
          "force err %.2e, failed %d" % (base, cap, cap / base, res["latency_p50_s"], res["latency_p99_s"], ferr,
                                          len(bad)), file=sys.stderr, flush=True)
    return res


def batch_leg(ctx, n_case, n_sample, nmin=8000, nmax=40000, seed=4):
    """Bump channels of log-uniform size in [nmin, nmax] nodes, bump height U(0.05, 0.15), Mach U(0.3, 0.8), AoA
    U(-2, 4) deg, first-order Roe, to RMS 1e-10, FGMRES(30) with the multicolour ILU(0).  The plain FP64 GPU path
    solves the first n_sample of them one at a time (the cases are drawn independently, so these are a random
    sample of the same distribution); cfd_batch_solve solves all of them in one union mesh with the
    mixed-precision linear solve."""
    from cfdsolutions_amd import cfd, mesh

    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax), n_case))
    meshes = [mesh.bump_channel(*bump_dims(n), bump_h=h) for n, h in zip(sizes, rng.uniform(0.05, 0.15, n_case))]
    machs, alphas = rng.uniform(0.3, 0.8, n_case), rng.uniform(-2.0, 4.0, n_case)
    params = [cfd.make_params(mach=mk, alpha_deg=a, flux="roe", order=1) for mk, a in zip(machs, alphas)]
    t64, r64 = 0.0, []
    n_sample = min(n_sample, n_case)
    for m, p, a in zip(meshes[:n_sample], params[:n_sample], alphas[:n_sample]):
        dm = cfd.DeviceMesh(ctx, m, p)
        r, _ = dm.solve_steady(p, alpha_deg=a, precond=cfd.PREC_ILU0_MC, mixed=False, check=False, **SOLVE_KW)
        t64 += r.seconds
        r64.append(r)
        dm.close()
        print("bench: batch baseline case %d/%d: %d nodes, %.2f s" % (len(r64), n_sample, m.n_node, r.seconds),
              file=sys.stderr, flush=True)
    print("bench: batch of %d cases (%d nodes in all)" % (n_case, sum(m.n_node for m in meshes)), file=sys.stderr,
          flush=True)
    rc, rb = cfd.batch_solve(ctx, meshes, params, alphas=list(alphas), precond=cfd.PREC_ILU0_MC, mixed=MIXED_LEVEL,
                             **SOLVE_KW)
    # each case reports the time until its own result was final (finished cases are harvested when the union
    # compacts), so the batch took as long as its slowest case
    tb = max(r.seconds for r in rb) if rc == 0 else float("nan")
    ferr = max((float(np.max(np.abs(np.array(b.force) - np.array(a.force))) / np.max(np.abs(a.force)))
                for a, b in zip(r64, rb)), default=float("nan")) if rc == 0 else float("nan")
    nodes = [m.n_node for m in meshes]
    rate64 = n_sample / t64 if n_sample else float("nan")
    res = {"workload": "%d bump channels of %d-%d nodes (log-uniform, %d in all), 1st-order Roe, M 0.3-0.8, "
                       "AoA -2..4, RMS < 1e-10" % (n_case, min(nodes), max(nodes), sum(nodes)),
           "fp64_sequential_solves_per_s": rate64,
           "fp64_sample_cases": n_sample,
           "batch_mixed_solves_per_s": n_case / tb, "mixed_level": int(MIXED_LEVEL),
           "batch_vs_fp64": (n_case / tb) / rate64,
           "batch_seconds": tb, "status": int(rc),
           "converged": int(sum(1 for b in rb if b.converged)),
           "force_rel_err_mixed_vs_fp64": ferr}
    print("bench: batch %d cases fp64 sequential %.2f solves/s (%d-case sample), batch mixed %.2f solves/s (%.2fx) "
          "in %.2f s, force err %.2e" % (n_case, rate64, n_sample, n_case / tb, (n_case / tb) / rate64, tb, ferr),
          file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=160, help="hexes per side of the cube")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--only", action="store_true",
                    help="run only the headline (fast build) and its exact-build twin (for profiling)")
    ap.add_argument("--no-solves", action="store_true", help="skip the steady-solve and serving legs")
    ap.add_argument("--no-cpu-solve", action="store_true", help="skip the oracle's CPU timing of the 48x16x16 solve")
    ap.add_argument("--legs", type=str, default="residual,solve,solve3,serve,batch",
                    help="comma list of legs to run: residual, solve (48x16x16), solve3 (192x64x64), serve, batch")
    ap.add_argument("--serve-requests", type=int, default=16,
                    help="requests of the serving leg (the survey's serving stream has 2000; the bench takes a "
                         "sample of the same distribution)")
    ap.add_argument("--serve-meshes", type=int, default=6)
    ap.add_argument("--serve-nodes", type=str, default="20000,1500000",
                    help="min,max nodes of the serving meshes (log-uniform)")
    ap.add_argument("--serve-inflight", type=int, default=4,
                    help="server worker threads; 4 matches the HIP runtime's 4 hardware queues per process")
    ap.add_argument("--serve-batch", type=int, default=8, help="requests per union-mesh batch in the server")
    ap.add_argument("--serve-wait-us", type=int, default=50000,
                    help="longest wait of a batchable request for companions while the GPU is busy")
    ap.add_argument("--serve-case-nodes", type=int, default=0,
                    help="largest mesh (nodes) the server batches (0: the library default, 131072)")
    ap.add_argument("--batch-cases", type=int, default=512,
                    help="small cases (8k-40k nodes, Roe, M 0.3-0.8, AoA -2..4) solved as one batch (0: skip)")
    ap.add_argument("--batch-fp64-sample", type=int, default=16,
                    help="cases of the batch also solved one at a time in FP64 for the baseline rate (0: none)")
    ap.add_argument("--mixed-level", type=int, default=2, choices=(1, 2, 3),
                    help="mixed= of the serving and batch legs' mixed solves: 1 FP32 Jacobian, factors and Krylov "
                         "basis; 2 (default) as 1 with the block-SELL copies of the multicolour ILU(0) factors in "
                         "FP16; 3 as 2 with the Jacobian's block-SELL copy in FP16 too")
    ap.add_argument("--autotune", action="store_true",
                    help="solve legs tune each mesh's same-bits kernel variants at its first Newton iteration "
                         "(FP64 and mixed paths alike)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (scale factor for HBM counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU; the residual path shards by node partition with one halo "
              "exchange per stage, which is not built here. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from cfdsolutions_amd import cfd, mesh

    if args.autotune:
        SOLVE_KW["autotune"] = True
    global MIXED_LEVEL
    MIXED_LEVEL = args.mixed_level
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("bench: calibration copy of %d MiB done" % args.calibrate_copy, file=sys.stderr, flush=True)
        del a, b

    legs = set(x.strip() for x in args.legs.split(",") if x.strip())
    if "residual" not in legs:
        ctx = cfd.Context(0)
        out = {"metric": "converged solves/s & p99 latency (mixed vs FP64 GPU); residual-eval GB/s",
               "legs": sorted(legs), "autotune": bool(args.autotune)}
        if "solve" in legs:
            out["solves"] = solve_leg(ctx, cpu=not args.no_cpu_solve)
        if "solve3" in legs:
            out["solve_192x64x64"] = solve3_leg(ctx)
        if "serve" in legs:
            nmin, nmax = (int(x) for x in args.serve_nodes.split(","))
            out["serving"] = serve_leg(ctx, args.serve_requests, args.serve_meshes, nmin, nmax, args.serve_inflight,
                                       args.serve_batch, batch_wait_us=args.serve_wait_us,
                                batch_case_nodes=args.serve_case_nodes)
        if "batch" in legs and args.batch_cases:
            out["batched_small_cases"] = batch_leg(ctx, args.batch_cases, args.batch_fp64_sample)
        ctx.close()
        print(json.dumps(out))
        return

    t0 = time.perf_counter()
    m = mesh.jittered_cube(args.n, jitter=0.2, seed=20261202)
    t_mesh = time.perf_counter() - t0
    N, E, B = m.n_node, m.n_edge, m.n_bedge
    print("bench: mesh %d^3 N=%d E=%d B=%d (%.1fs)" % (args.n, N, E, B, t_mesh), file=sys.stderr, flush=True)
    p = cfd.make_params(mach=0.8, flux="roe", order=2, limiter=cfd.LIM_VENKAT, limiter_order=1, venkat_k=5.0,
                        gradient=cfd.GRAD_GG)
    q = make_state(m, p)
    ctx = cfd.Context(0)
    dm = cfd.DeviceMesh(ctx, m, p)
    dm.set(cfd.FIELD_Q, q)
    ev = HipEvents()

    bytes64 = algorithmic_bytes(N, E, B, 8)
    bytes32 = algorithmic_bytes(N, E, B, 4)
    variants = {}
    EX, FA = cfd.ARITH_EXACT, cfd.ARITH_FAST
    plan = [("roe", cfd.FP64, True, True, FA), ("roe", cfd.FP64, True, True, EX)]
    if not args.no_variants and not args.only:
        plan += [("roe", cfd.FP64, True, False, EX), ("hllc", cfd.FP64, True, True, EX),
                 ("ausm+", cfd.FP64, True, True, EX), ("roe", cfd.FP32, True, True, EX),
                 ("hllc", cfd.FP32, True, True, EX), ("ausm+", cfd.FP32, True, True, EX),
                 ("roe", cfd.FP64, False, True, EX), ("hllc", cfd.FP64, True, True, FA),
                 ("ausm+", cfd.FP64, True, True, FA), ("roe", cfd.FP32, True, True, FA),
                 ("hllc", cfd.FP32, True, True, FA), ("ausm+", cfd.FP32, True, True, FA),
                 ("roe", cfd.FP64, False, True, FA)]
    for flux, prec, second, fused, arith in plan:
        pv = cfd.make_params(mach=0.8, flux=flux, order=2 if second else 1, limiter=cfd.LIM_VENKAT if second else 0,
                             venkat_k=5.0, gradient=cfd.GRAD_GG, precision=prec, arith=arith)
        head_case = (flux, prec, second, fused, arith) == plan[0]
        steps = args.steps if head_case else max(10, args.steps // 4)
        wall, tg, tl, tr, tmed = time_variant(dm, ev, pv, steps, args.warmup if head_case else 5, second, fused)
        by = bytes64 if prec == cfd.FP64 else bytes32
        rkey = ("R2" if second else "R1") + ("split" if flux == "roe" else "")
        total = (by["G"] + by["L"] + by[rkey]) if second else by[rkey]
        name = "%s-%s%s-%s" % (flux, "fp64" if prec == cfd.FP64 else "fp32", "fast" if arith == FA else "",
                               ("GL+R2" if fused else "G+L+R2") if second else "R1")
        variants[name] = dict(ms_per_step=wall, median_step_ms=tmed, gbs=total / wall / 1e6, g_ms=tg, l_ms=tl, r_ms=tr,
                              r_gbs=by[rkey] / tr / 1e6 if tr > 0 else None,
                              g_gbs=(by["G"] + (by["L"] if fused else 0)) / tg / 1e6 if tg > 0 else None,
                              l_gbs=by["L"] / tl / 1e6 if tl > 0 else None, bytes=total)
        print("bench: %-22s %8.3f ms/step %8.1f GB/s  (G %.3f L %.3f R %.3f ms)" %
              (name, wall, total / wall / 1e6, tg, tl, tr), file=sys.stderr, flush=True)
    # the exact build with the flux-once patch kernel (work edges in the reference's edge order: the same bits as
    # the node-centred kernel; the exact build runs it when CFD_FLUX_ONCE=1)
    once_name = "roe-fp64-GL+R2-flux-once"
    os.environ["CFD_FLUX_ONCE"] = "1"
    try:
        pv = cfd.make_params(mach=0.8, flux="roe", order=2, limiter=cfd.LIM_VENKAT, venkat_k=5.0,
                             gradient=cfd.GRAD_GG, arith=EX)
        wall, tg, tl, tr, tmed = time_variant(dm, ev, pv, max(10, args.steps // 4), 5, True, True)
    finally:
        del os.environ["CFD_FLUX_ONCE"]
    total = bytes64["G"] + bytes64["L"] + bytes64["R2split"]
    variants[once_name] = dict(ms_per_step=wall, median_step_ms=tmed, gbs=total / wall / 1e6, g_ms=tg, l_ms=tl,
                               r_ms=tr, r_gbs=bytes64["R2split"] / tr / 1e6, bytes=total)
    print("bench: %-22s %8.3f ms/step %8.1f GB/s  (G %.3f L %.3f R %.3f ms)" %
          (once_name, wall, total / wall / 1e6, tg, tl, tr), file=sys.stderr, flush=True)
    head_name, exact_name = "roe-fp64fast-GL+R2", "roe-fp64-GL+R2"
    head = variants[head_name]
    # parity spot check of both builds of the measured configuration against the oracle: the exact build to the
    # bit, the fast build within 1e-12 relative (max over nodes and components of |gpu - oracle| / max |oracle|)
    cpu = None
    if not args.no_cpu_baseline and not args.only:
        gpu = {}
        for arith, once in ((FA, None), (EX, None), (EX, "1")):
            pv = cfd.make_params(mach=0.8, flux="roe", order=2, limiter=cfd.LIM_VENKAT, venkat_k=5.0,
                                 gradient=cfd.GRAD_GG, arith=arith)
            if once:
                os.environ["CFD_FLUX_ONCE"] = once
            try:
                dm.gradient_limiter(pv)
                dm.residual(pv)
            finally:
                os.environ.pop("CFD_FLUX_ONCE", None)
            gpu[(arith, once)] = (dm.get(cfd.FIELD_RES_CONV), dm.get(cfd.FIELD_RES_DISS))

        def rel_to(rc, rd, arith, once=None):
            gc, gd = gpu[(arith, once)]
            return float(max(np.max(np.abs(gc - rc)) / np.max(np.abs(rc)),
                             np.max(np.abs(gd - rd)) / np.max(np.abs(rd))))

        secs, rc, rd = cpu_baseline(m, pv, q)
        rel, rel_ex, rel_once = rel_to(rc, rd, FA), rel_to(rc, rd, EX), rel_to(rc, rd, EX, "1")
        secs_f, rc, rd = cpu_baseline(m, pv, q, build="speed")
        rel_f = rel_to(rc, rd, FA)
        cpu = dict(value=head["bytes"] / secs / 1e9, unit="GB/s", cores=1, kind="port",
                   sample="one full G+L+R2 Roe FP64 evaluation of the same 160^3 mesh and state (%.1f s)" % secs,
                   build="oracle.c -O2 -ffp-contract=off (parity build)",
                   parity_max_rel_diff=rel, parity_tolerance=1e-12, exact_build_max_rel_diff=rel_ex,
                   exact_flux_once_max_rel_diff=rel_once,
                   speed_build=dict(value=head["bytes"] / secs_f / 1e9, seconds=secs_f,
                                    build="oracle.c -O3 -march=x86-64-v3", max_rel_diff_vs_gpu=rel_f))
        print("bench: cpu oracle %.2f s (speed build %.2f s), GPU-vs-oracle max rel diff: fast %.3e, exact %.3e, "
              "exact flux-once %.3e (speed build vs fast %.1e)" % (secs, secs_f, rel, rel_ex, rel_once, rel_f),
              file=sys.stderr, flush=True)
        if rel > 1e-12 or rel_ex != 0.0 or rel_once != 0.0:
            print("bench: PARITY FAILURE of the measured configuration", file=sys.stderr, flush=True)
            sys.exit(1)
    dm.close()
    solves = solve3 = serving = batched = None
    if not args.no_solves and not args.only:
        del dm
        if "solve" in legs:
            solves = solve_leg(ctx, cpu=not args.no_cpu_solve)
        if "solve3" in legs:
            solve3 = solve3_leg(ctx)
        if "serve" in legs:
            nmin, nmax = (int(x) for x in args.serve_nodes.split(","))
            serving = serve_leg(ctx, args.serve_requests, args.serve_meshes, nmin, nmax, args.serve_inflight,
                                args.serve_batch, batch_wait_us=args.serve_wait_us,
                                batch_case_nodes=args.serve_case_nodes)
        if "batch" in legs and args.batch_cases:
            batched = batch_leg(ctx, args.batch_cases, args.batch_fp64_sample)
    ctx.close()
    r_achieved = head["r_gbs"]
    copy_gbs = measured_copy_gbs()
    if copy_gbs:
        print("bench: device-to-device copy of 1 GiB: %.0f GB/s" % copy_gbs, file=sys.stderr, flush=True)
    traffic_bytes = measured_traffic("k_residual_patch") if args.n == 160 else None
    exact = variants.get(exact_name)
    out = {
        "metric": "converged solves/s & p99 latency (mixed vs FP64 GPU); residual-eval GB/s",
        "value": round(head["gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (jittered cube mesh, random per-node state)",
        "config": {"workload": "residual microbenchmark: %d^3 jittered cube, %d tets, %d nodes, RCM; "
                               "GG gradient + Venkatakrishnan limiter + 2nd-order Roe residual (Conv/Diss), FP64, "
                               "fast build (within 1e-12 of the reference)" % (args.n, 6 * args.n ** 3, N),
                   "nodes": N, "edges": E, "ghosts": B, "parallelism": "single GPU",
                   "solve_autotune": bool(args.autotune)},
        "exact_build": {"value": round(exact["gbs"], 2), "ms_per_step": round(exact["ms_per_step"], 4),
                        "residual_kernel": "k_residual<double,Roe,2nd order,limited>",
                        "residual_gbs": round(exact["r_gbs"], 1), "residual_ms": round(exact["r_ms"], 4),
                        "flux_once": {"value": round(variants[once_name]["gbs"], 2),
                                      "ms_per_step": round(variants[once_name]["ms_per_step"], 4),
                                      "residual_kernel": "k_residual_patch<double,Roe,2nd order,limited> "
                                                         "(work edges in the reference's edge order, CFD_FLUX_ONCE=1)",
                                      "residual_ms": round(variants[once_name]["r_ms"], 4)}}
        if exact else None,
        "roofline": {"bound": "hbm", "kernel": "k_residual_patch<double,Roe,2nd order,limited> (fast build)",
                     "achieved": round(r_achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(r_achieved / PEAK_HBM_GBS, 4), "traffic": traffic_bytes,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, profiles/residual_cube160_fp64_hbm_counters.md"
                     if traffic_bytes else None,
                     "algorithmic_bytes_per_launch": bytes64["R2split"],
                     "measured_copy_peak": round(copy_gbs, 1) if copy_gbs else None,
                     "frac_of_measured_copy": round(r_achieved / copy_gbs, 4) if copy_gbs else None,
                     "avg_launch_ms": round(head["r_ms"], 4)},
        "cpu_baseline": cpu,
        "solves": solves,
        "solve_192x64x64": solve3,
        "serving": serving,
        "batched_small_cases": batched,
        "variants": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in variants.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
