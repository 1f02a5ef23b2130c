"""MI355X-native implicit unstructured finite-volume Euler hot path.

The compute path is libcfd_hip.so (hand-written HIP for gfx950 behind a C ABI,
include/cfd_hip.h); mesh preprocessing is libcfd_mesh.so (include/cfd_mesh.h).
"""
