"""Restart and VTK files of states the GPU produced.

- For every golden case, the state after the characteristic boundary
  condition is computed on the GPU (from the case's input state) and written
  by the restart and VTK writers; the files equal, byte for byte, the ones
  the compiled reference wrote for its own state after the boundary
  condition (Restart_Writer, Euler3D/src/RestartIO.cpp:45-109; VTK_Writer,
  Euler3D/src/MeshIO.cpp:219-399).
- A converged GPU solve written to a restart file and read back into a new
  device mesh gives the same bits, and a solve started from it is already
  converged, with the residual RMS of the original handle continuing from
  the same state (the characteristic condition reads the ghost states, so
  the first RMS of a continued solve is that of one more boundary pass, in
  the reference as here).
"""
import numpy as np
import pytest

from cfdsolutions_amd import cfd
from cfdsolutions_amd import mesh as cmesh
from tests.golden_util import cases, load
from tests.test_gpu_parity import FixtureMesh, gpu_params
from tests.test_io_golden import regen_ugrid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cfd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", cases())
def test_gpu_state_files_match_reference(ctx, name, tmp_path):
    print("restart/VTK of the GPU state:", name, flush=True)
    d = load(name)
    p = gpu_params(d)
    dm = cfd.DeviceMesh(ctx, FixtureMesh(d), p)
    dm.set(cfd.FIELD_Q, d["q_in"])
    it = d["meta"]["args"]["iter"]
    dm.bc(p, it)
    q = dm.get(cfd.FIELD_Q)
    dm.close()
    n, b = d["xyz"].shape[0], d["bedge"].shape[0]
    rst = tmp_path / "restart.srt"
    cmesh.write_restart(rst, q, n, b, iteration=it)
    assert rst.read_bytes() == d["restart_file"].tobytes()
    ug, vtk = tmp_path / "mesh.ugrid", tmp_path / "solution.vtk"
    regen_ugrid(d, ug)
    inf = d["inf"]
    cmesh.write_vtk(ug, vtk, q, gamma=float(inf[8]), r_gas=float(inf[9]))
    assert vtk.read_bytes() == d["vtk_file"].tobytes()


def test_restart_of_a_converged_solve(ctx, tmp_path):
    print("restart of a converged GPU solve", flush=True)
    m = cmesh.bump_channel(24, 8, 8)
    p = cfd.make_params(mach=0.5, alpha_deg=1.0, flux="roe", order=1)
    kw = dict(max_iter=300, tol=1e-10, krylov_m=30, max_restarts=4, rtol=1e-3, precond=cfd.PREC_ILU0)
    dm = cfd.DeviceMesh(ctx, m, p)
    r, hist = dm.solve_steady(p, alpha_deg=1.0, **kw)
    assert r.converged
    q = dm.get(cfd.FIELD_Q)
    r1, hist1 = dm.solve_steady(p, alpha_deg=1.0, **kw)  # the original handle, continued
    dm.close()
    assert r1.converged and r1.iterations == 0
    rst = tmp_path / "restart.srt"
    cmesh.write_restart(rst, q, m.n_node, m.n_bedge, iteration=r.iterations)
    it, q2 = cmesh.read_restart(rst, m.n_node, m.n_bedge)
    assert it == r.iterations
    assert np.array_equal(q2.view(np.uint64), np.ascontiguousarray(q).view(np.uint64))
    dm2 = cfd.DeviceMesh(ctx, m, p)
    dm2.set(cfd.FIELD_Q, q2)
    assert np.array_equal(dm2.get(cfd.FIELD_Q), q)
    r2, hist2 = dm2.solve_steady(p, alpha_deg=1.0, **kw)
    dm2.close()
    assert r2.converged and r2.iterations == 0
    assert hist2[0] == hist1[0]
