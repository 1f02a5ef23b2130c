"""Writes the golden results of the reference's steady implicit loop for the
drop-in test (tests/test_gpu_shim.py).

Run in the development container, where the reference sources exist:
    python tests/golden/make_shim_golden.py
It builds oracle/_ref/steady_ref (the reference's own objects and its
Solver_Steady_Implicit loop, oracle/steady_driver.cpp), generates each case's
UGRID mesh with cfdsolutions_amd.mesh, runs a fixed number of pseudo-time
iterations from the freestream, and stores the final state and the bytes of
the reference's Residual.res in tests/golden/shim/<case>.npz.  The fixtures
hold data only.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cfdsolutions_amd import mesh as cmesh  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "steady_ref")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shim")

# name: bump channel (nx, ny, nz, h), driver arguments
CASES = {
    "bump10_roe1_approx": ((10, 3, 3, 0.1), dict(flux="roe", order=1, lim="none", mach=0.5, alpha=0.0, iters=8,
                                                 sweeps=6, jac="approx")),
    "bump10_roe2_venkat_approx": ((10, 3, 3, 0.1), dict(flux="roe", order=2, lim="venkat", mach=0.5, alpha=2.0,
                                                        iters=8, sweeps=6, jac="approx")),
    "bump12_roe2_bj_central": ((12, 4, 3, 0.08), dict(flux="roe", order=2, lim="bj", mach=0.6, alpha=-1.0, iters=5,
                                                      sweeps=4, jac="central")),
}


def driver_args(mesh_path, outdir, a):
    return [mesh_path, outdir, a["flux"], str(a["order"]), a["lim"], repr(a["mach"]), repr(a["alpha"]),
            str(a["iters"]), str(a["sweeps"]), a["jac"]]


def make_mesh(spec, path):
    nx, ny, nz, h = spec
    cmesh.bump_channel(nx, ny, nz, bump_h=h, ugrid_path=path)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "shim", "-j8"], check=True)
    os.makedirs(OUT, exist_ok=True)
    for name, (mspec, a) in CASES.items():
        with tempfile.TemporaryDirectory() as td:
            ug = os.path.join(td, "mesh.ugrid")
            make_mesh(mspec, ug)
            subprocess.run([DRIVER] + driver_args(ug, td, a), check=True, stdout=subprocess.DEVNULL)
            q = np.load(os.path.join(td, "q_final.npy"))
            with open(os.path.join(td, "Residual.res"), "rb") as fh:
                res = np.frombuffer(fh.read(), dtype=np.uint8)
        out = os.path.join(OUT, name + ".npz")
        np.savez_compressed(out, q_final=q, residual_res=res,
                            meta=np.array(json.dumps({"mesh": mspec, "args": a})))
        print(name, q.shape, os.path.getsize(out))


if __name__ == "__main__":
    main()
