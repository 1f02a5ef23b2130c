"""In-tree build of the native libraries.

libcfd_mesh.so  host C++ (g++): mesh generation and preprocessing
libcfd_hip.so   HIP for gfx950 (hipcc): device kernels + the C-ABI
Both are written next to this file so they travel with the source tree.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["cfd_api.hip", "residual.hip", "residual_fast.hip", "linsolve.hip", "serve.hip", "batch.hip", "prep.hip",
               "tune.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # parity build: no FMA contraction, so FP64 results match the CPU
    # reference bit for bit
    "-ffp-contract=off",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def lib_path(name):
    return os.path.join(HERE, name)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_mesh():
    out = lib_path("libcfd_mesh.so")
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I" + INC, "-o", out,
          os.path.join(CSRC, "mesh.cpp")])
    return out


def build_hip():
    out = lib_path("libcfd_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return None
    tmp = out + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc")] + HIP_FLAGS + ["-I" + INC, "-I" + CSRC, "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def build_all():
    return [build_mesh(), build_hip()]
