"""The C-ABI libraries load and export every symbol their headers declare."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    lines = open(os.path.join(ROOT, "include", header)).read().splitlines()
    code = "\n".join(l for l in lines if not l.lstrip().startswith(("*", "/*", "//")))
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(cfd\w*|cfdm\w*)\s*\(", code, re.M)))


@pytest.mark.parametrize("header,lib", [("cfd_hip.h", "libcfd_hip.so"), ("cfd_mesh.h", "libcfd_mesh.so")])
def test_exports(header, lib):
    path = os.path.join(ROOT, "cfdsolutions_amd", lib)
    if not os.path.exists(path):
        pytest.fail("%s is not built" % lib)
    L = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts_match_ctypes(tmp_path):
    """Every struct of cfd_hip.h has the same size and field offsets in C and in the ctypes binding."""
    import shutil
    import subprocess
    import sys

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    sys.path.insert(0, ROOT)
    from cfdsolutions_amd import cfd

    structs = {"cfd_mesh_desc": cfd.MeshDesc, "cfd_params": cfd.Params, "cfd_linsolve_opts": cfd.LinsolveOpts,
               "cfd_linsolve_stats": cfd.LinsolveStats, "cfd_solve_opts": cfd.SolveOpts,
               "cfd_solve_result": cfd.SolveResult, "cfd_server_opts": cfd.ServerOpts, "cfd_request": cfd.Request,
               "cfd_completion": cfd.Completion, "cfd_server_stats": cfd.ServerStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cfd_hip.h"', "int main(void) {"]
    expect = []
    for cname, py in structs.items():
        lines.append('printf("%%zu\\n", sizeof(%s));' % cname)
        expect.append(ctypes.sizeof(py))
        for f, _ in py._fields_:
            lines.append('printf("%%zu\\n", offsetof(%s, %s));' % (cname, f))
            expect.append(getattr(py, f).offset)
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got == expect
