"""CPU checks of the measurement helpers in bench.py and of the oracle's speed build used for CPU timings."""
import numpy as np

import bench
from oracle import oracle as orc
from tests.golden_util import cases, load, oracle_mesh, oracle_params


def test_measured_traffic_reads_the_committed_counter_summary():
    # 2 x FETCH_SIZE + WRITE_SIZE of k_residual, in bytes, from profiles/
    t = bench.measured_traffic("k_residual")
    assert t is not None and 5e9 < t < 12e9
    assert bench.measured_traffic("no_such_kernel") is None


def test_bump_dims_is_close_to_the_requested_node_count():
    for n in (8000, 14161, 40000, 1500000):
        nx, ny, nz = bench.bump_dims(n)
        assert nx == 3 * ny and ny == nz
        assert abs((nx + 1) * (ny + 1) * (nz + 1) - n) / n < 0.25


def test_oracle_speed_build_is_selectable_and_close_to_the_parity_build():
    """The speed build (-O3, FMA contraction allowed) times the CPU path; it must compute the same residual up to
    rounding, and selecting it must not change what the parity build returns."""
    name = next(c for c in cases() if "roe2" in c)
    d = load(name)
    m, p = oracle_mesh(d, orc), oracle_params(d, orc)
    g = d["grad"] if "grad" in d else None
    phi = d["phi"] if "phi" in d else None
    rc0, rd0, _, _ = orc.residual(m, p, d["q_bc"], g, phi)
    old = orc.use_build("speed")
    try:
        rc1, rd1, _, _ = orc.residual(m, p, d["q_bc"], g, phi)
    finally:
        assert orc.use_build(old) == "speed"
    assert np.max(np.abs(rc1 - rc0)) <= 1e-12 * np.max(np.abs(rc0))
    assert np.max(np.abs(rd1 - rd0)) <= 1e-12 * np.max(np.abs(rd0))
    rc2, _, _, _ = orc.residual(m, p, d["q_bc"], g, phi)
    assert np.array_equal(rc2, rc0)
    assert np.array_equal(rc0, np.asarray(d["res_conv"]).reshape(rc0.shape))
