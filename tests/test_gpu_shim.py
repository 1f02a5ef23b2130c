"""The drop-in test: the reference's own steady implicit loop
(Solver_Steady_Implicit, Euler3D/src/Solver_Steady_Implicit.cpp:33-168)
linked with include/ref_shim/cfd_reference_shim.cpp, so its boundary
condition, gradient, limiter, residual, time step, Jacobian and block
Gauss-Seidel solve run on the MI355X through the C ABI, while the loop, the
RMS and the update stay the reference's host code.

oracle/_ref/steady_shim is built from the reference's objects in the
development container (oracle/Makefile target `shim`) and travels with the
tree.  Its final state must equal, bit for bit, the state the unmodified
reference reaches on the CPU (tests/golden/shim, written by
tests/golden/make_shim_golden.py), and the reference's Residual.res must be
byte-identical.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEADY_SHIM = os.path.join(ROOT, "oracle", "_ref", "steady_shim")
SHIM_GOLDEN = os.path.join(ROOT, "tests", "golden", "shim")

pytestmark = pytest.mark.gpu


def shim_cases():
    return sorted(f[:-4] for f in os.listdir(SHIM_GOLDEN) if f.endswith(".npz"))


@pytest.mark.skipif(not os.path.exists(STEADY_SHIM), reason="shim binary not built (needs the reference sources)")
@pytest.mark.parametrize("name", shim_cases())
def test_reference_loop_through_shim_matches_reference(name):
    from cfdsolutions_amd import mesh as cmesh

    g = dict(np.load(os.path.join(SHIM_GOLDEN, name + ".npz")))
    meta = json.loads(str(g["meta"]))
    a = meta["args"]
    nx, ny, nz, h = meta["mesh"]
    print("shim case", name, flush=True)
    sys.stdout.flush()
    with tempfile.TemporaryDirectory() as td:
        ug = os.path.join(td, "mesh.ugrid")
        cmesh.bump_channel(nx, ny, nz, bump_h=h, ugrid_path=ug)
        args = [STEADY_SHIM, ug, td, a["flux"], str(a["order"]), a["lim"], repr(a["mach"]), repr(a["alpha"]),
                str(a["iters"]), str(a["sweeps"]), a["jac"]]
        r = subprocess.run(args, capture_output=True, text=True, timeout=100)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        q = np.load(os.path.join(td, "q_final.npy"))
        with open(os.path.join(td, "Residual.res"), "rb") as fh:
            res = fh.read()
    assert np.array_equal(q, g["q_final"]), np.max(np.abs(q - g["q_final"]))
    assert res == g["residual_res"].tobytes()
