"""ctypes bindings to libntt_amd.so (see include/ntt_engine.h, include/multiply.h).

No fallback: if the library is not built, or the GPU is absent, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libntt_amd.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_lib = None

STATUS = {0: "NTT_OK", 1: "NTT_ERR_NOT_INIT", 2: "NTT_ERR_TOO_LARGE", 3: "NTT_ERR_HIP",
          4: "NTT_ERR_ARG", 5: "NTT_ERR_INTEGRITY", 6: "NTT_ERR_NOMEM"}


class NttError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise RuntimeError(f"{lib_path} is not built: run `make -C cryptography_amd` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(lib_path)
    sz, u64, vp = ctypes.c_size_t, ctypes.c_uint64, ctypes.c_void_p
    L.ntt_engine_last_error.restype = ctypes.c_char_p
    L.ntt_engine_init.argtypes = [ctypes.c_int]
    L.ntt_engine_is_init.restype = ctypes.c_int
    L.ntt_engine_log_len.argtypes = [sz, sz]
    L.ntt_mul_limbs.argtypes = [_u64p, _u64p, sz, _u64p, sz]
    L.ntt_sqr_limbs.argtypes = [_u64p, _u64p, sz]
    for f in ("ntt_mul_device", "ntt_mul_device_async"):
        getattr(L, f).argtypes = [vp, vp, sz, vp, sz]
    for f in ("ntt_sqr_device", "ntt_sqr_device_async"):
        getattr(L, f).argtypes = [vp, vp, sz]
    L.ntt_engine_set_threshold.argtypes = [ctypes.c_int]
    L.ntt_engine_dispatch_count.restype = ctypes.c_ulong
    L.ntt_engine_set_timing.argtypes = [ctypes.c_int]
    L.ntt_engine_last_ntt_ms.restype = ctypes.c_double
    L.ntt_engine_last_total_ms.restype = ctypes.c_double
    L.ntt_device_alloc.restype = vp
    L.ntt_device_alloc.argtypes = [sz]
    L.ntt_device_free.argtypes = [vp]
    L.ntt_memcpy_h2d.argtypes = [vp, vp, sz]
    L.ntt_memcpy_d2h.argtypes = [vp, vp, sz]
    L.ntt_engine_read_twiddles.argtypes = [_u64p, sz]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise NttError(rc, lib().ntt_engine_last_error().decode())


def init(device=0):
    check(lib().ntt_engine_init(device))


def log_len(na, nb):
    return lib().ntt_engine_log_len(na, nb)


def _as_u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _p(a):
    return a.ctypes.data_as(_u64p)


def mul(a, b):
    """Product of two u64 limb arrays through the host-buffer ABI (na+nb limbs)."""
    init()
    a, b = _as_u64(a), _as_u64(b)
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    check(lib().ntt_mul_limbs(_p(out), _p(a), len(a), _p(b), len(b)))
    return out


def sqr(a):
    init()
    a = _as_u64(a)
    out = np.zeros(2 * len(a), dtype=np.uint64)
    check(lib().ntt_sqr_limbs(_p(out), _p(a), len(a)))
    return out


class DeviceBuffer:
    """Device memory (u64 words) owned by the caller, freed on close()."""

    def __init__(self, n_words):
        init()
        self.n = int(n_words)
        self.ptr = lib().ntt_device_alloc(max(1, self.n) * 8)
        if not self.ptr:
            raise NttError(6, f"ntt_device_alloc({self.n * 8}) failed")

    @classmethod
    def from_array(cls, a):
        a = _as_u64(a)
        buf = cls(len(a))
        buf.upload(a)
        return buf

    def upload(self, a):
        a = _as_u64(a)
        assert len(a) <= self.n
        if len(a):
            check(lib().ntt_memcpy_h2d(self.ptr, a.ctypes.data, len(a) * 8))

    def download(self, n=None):
        n = self.n if n is None else n
        out = np.zeros(n, dtype=np.uint64)
        if n:
            check(lib().ntt_memcpy_d2h(out.ctypes.data, self.ptr, n * 8))
        return out

    def close(self):
        if self.ptr:
            lib().ntt_device_free(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def mul_device(dc, da, na, db, nb, sync=True):
    f = lib().ntt_mul_device if sync else lib().ntt_mul_device_async
    check(f(dc.ptr, da.ptr, na, db.ptr, nb))


def sqr_device(dc, da, na, sync=True):
    f = lib().ntt_sqr_device if sync else lib().ntt_sqr_device_async
    check(f(dc.ptr, da.ptr, na))
