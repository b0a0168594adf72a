"""MI355X (gfx950) big-integer multiply engine: 3-prime CRT-NTT in hand-written HIP.

The product is the C-ABI shared library ``libntt_amd.so`` built in-tree from
``cryptography_amd/csrc`` (``make -C cryptography_amd``).  This package is a
thin ctypes harness over that ABI for tests and benchmarks: it has no compute
path of its own and raises if the library is missing.
"""
from .engine import (  # noqa: F401
    NttError,
    lib,
    lib_path,
    init,
    mul,
    sqr,
    DeviceBuffer,
    mul_device,
    sqr_device,
    log_len,
    poly_mul,
)
