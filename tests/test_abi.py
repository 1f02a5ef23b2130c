"""The C-ABI libraries load and export every symbol their headers declare."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(cfd\w*|cfdm\w*)\s*\(", text, re.M)))


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
