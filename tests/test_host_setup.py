"""Host-side mesh set-up: the parallel ILU(0) merge map and the multicolour
level schedules built colour by colour in parallel equal a sequential pass
(tests/host/setup_levels_check.cpp, compiled here with g++)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "cfdsolutions_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_parallel_setup_matches_sequential(tmp_path):
    exe = str(tmp_path / "setup_levels_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "host", "setup_levels_check.cpp")], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.strip().splitlines()
    assert len(lines) == 6 and all(l.endswith("same") for l in lines), out.stdout
