#!/usr/bin/env python3
"""Residual-evaluation throughput on MI355X: the 160^3 jittered-cube
microbenchmark of the implicit unstructured finite-volume Euler hot path.

Workload (one "step"): Green-Gauss gradient (G) + Venkatakrishnan limiter (L,
K = 5) + second-order Roe residual with the Conv/Diss split (R2), FP64, on a
160^3 hex grid split into 24.6M tetrahedra (4.17M nodes), interior nodes
jittered by U(-0.2h, 0.2h), RCM-ordered, with a random per-node state
(Mach U(0.3, 2.0) along +x, density and pressure with 5% Gaussian noise).
Inputs are resident in HBM before timing starts.

value = algorithmic bytes of G+L+R2 / wall time per step (GB/s).  The
roofline block reports the dominant kernel (R2) from HIP events on the
stream it runs on.  cpu_baseline times the CPU oracle (single-threaded port
of the reference) on the same mesh and state.

Also measured (in "variants"): HLLC and AUSM+ second order, FP32 versions,
and the first-order residual R1.
"""
import argparse
import ctypes
import json
import os
import sys
import time

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
    G = E * (2 * i + 3 * s) + N * (5 * s + s + 15 * s) + B * (2 * i + 3 * s)
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


def time_variant(dm, ev, p, steps, warmup, second_order=True):
    """Returns (wall ms per step, mean ms of G, L, R kernels)."""
    stream = dm.stream()

    def one(evs=None):
        if evs:
            ev.record(evs[0], stream)
        if second_order:
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
    for e in evsets:
        g += ev.elapsed_ms(e[0], e[1])
        l += ev.elapsed_ms(e[1], e[2])
        r += ev.elapsed_ms(e[2], e[3])
        for x in e:
            ev.destroy(x)
    return wall, g / steps, l / steps, r / steps


def cpu_baseline(m, p, q):
    from oracle import oracle as orc

    om = orc.Mesh(m.xyz, m.edge, m.edge_area, m.bedge, m.bedge_area, m.bedge_type, m.vol, m.n2n_ia, m.n2n_ja,
                  m.bedge_tag)
    op = orc.params({"rho": p.inf_rho, "u": p.inf_u, "v": p.inf_v, "w": p.inf_w, "p": p.inf_p,
                     "outflow_p": p.outflow_p}, flux=orc.ROE, order=2, limiter=orc.LIM_VENKAT, venkat_k=p.venkat_k)
    t0 = time.perf_counter()
    g = orc.gg_gradient(om, q)
    phi = orc.limiter(om, op, q, g)
    rc, rd, _, _ = orc.residual(om, op, q, g, phi)
    return time.perf_counter() - t0, rc, rd


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=160, help="hexes per side of the cube")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--only", action="store_true", help="run only the headline variant (for profiling)")
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
    plan = [("roe", cfd.FP64, True)]
    if not args.no_variants and not args.only:
        plan += [("hllc", cfd.FP64, True), ("ausm+", cfd.FP64, True), ("roe", cfd.FP32, True),
                 ("hllc", cfd.FP32, True), ("ausm+", cfd.FP32, True), ("roe", cfd.FP64, False)]
    for flux, prec, second in plan:
        pv = cfd.make_params(mach=0.8, flux=flux, order=2 if second else 1, limiter=cfd.LIM_VENKAT if second else 0,
                             venkat_k=5.0, gradient=cfd.GRAD_GG, precision=prec)
        steps = args.steps if (flux, prec, second) == plan[0] else max(10, args.steps // 4)
        wall, tg, tl, tr = time_variant(dm, ev, pv, steps, args.warmup if (flux, prec, second) == plan[0] else 5,
                                        second)
        by = bytes64 if prec == cfd.FP64 else bytes32
        rkey = ("R2" if second else "R1") + ("split" if flux == "roe" else "")
        total = (by["G"] + by["L"] + by[rkey]) if second else by[rkey]
        name = "%s-%s-%s" % (flux, "fp64" if prec == cfd.FP64 else "fp32", "G+L+R2" if second else "R1")
        variants[name] = dict(ms_per_step=wall, gbs=total / wall / 1e6, g_ms=tg, l_ms=tl, r_ms=tr,
                              r_gbs=by[rkey] / tr / 1e6 if tr > 0 else None,
                              g_gbs=by["G"] / tg / 1e6 if tg > 0 else None,
                              l_gbs=by["L"] / tl / 1e6 if tl > 0 else None, bytes=total)
        print("bench: %-22s %8.3f ms/step %8.1f GB/s  (G %.3f L %.3f R %.3f ms)" %
              (name, wall, total / wall / 1e6, tg, tl, tr), file=sys.stderr, flush=True)
    head = variants["roe-fp64-G+L+R2"]
    # parity spot check of the measured configuration against the oracle
    cpu = None
    if not args.no_cpu_baseline and not args.only:
        pv = cfd.make_params(mach=0.8, flux="roe", order=2, limiter=cfd.LIM_VENKAT, venkat_k=5.0,
                             gradient=cfd.GRAD_GG)
        dm.gradient(pv)
        dm.limiter(pv)
        dm.residual(pv)
        gc, gd = dm.get(cfd.FIELD_RES_CONV), dm.get(cfd.FIELD_RES_DISS)
        secs, rc, rd = cpu_baseline(m, pv, q)
        rel = max(np.max(np.abs(gc - rc)) / np.max(np.abs(rc)), np.max(np.abs(gd - rd)) / np.max(np.abs(rd)))
        cpu = dict(value=head["bytes"] / secs / 1e9, unit="GB/s", cores=1, kind="port",
                   sample="one full G+L+R2 Roe FP64 evaluation of the same 160^3 mesh and state (%.1f s)" % secs,
                   parity_max_rel_diff=float(rel))
        print("bench: cpu oracle %.2f s, GPU-vs-oracle max rel diff %.3e" % (secs, rel), file=sys.stderr, flush=True)
    dm.close()
    ctx.close()
    r_achieved = variants["roe-fp64-G+L+R2"]["r_gbs"]
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
                               "GG gradient + Venkatakrishnan limiter + 2nd-order Roe residual (Conv/Diss), FP64"
                               % (args.n, 6 * args.n ** 3, N),
                   "nodes": N, "edges": E, "ghosts": B, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_residual<double,Roe,2nd order,limited>",
                     "achieved": round(r_achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(r_achieved / PEAK_HBM_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": bytes64["R2split"],
                     "avg_launch_ms": round(variants["roe-fp64-G+L+R2"]["r_ms"], 4)},
        "cpu_baseline": cpu,
        "variants": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                     for k, v in variants.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
