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


class MulDesc(ctypes.Structure):
    """ntt_mul_desc: c[0..n_out) = a*b + add, device pointers."""
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("add", ctypes.c_void_p), ("na", ctypes.c_uint32), ("nb", ctypes.c_uint32),
                ("nadd", ctypes.c_uint32), ("n_out", ctypes.c_uint32)]


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
    L.ntt_engine_kernel_times.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.ntt_mul_batch_device.argtypes = [ctypes.POINTER(MulDesc), sz]
    L.ntt_mul_batch_device_async.argtypes = [ctypes.POINTER(MulDesc), sz]
    L.ntt_batch_graph_create.argtypes = [ctypes.POINTER(MulDesc), sz, ctypes.POINTER(ctypes.c_void_p)]
    L.ntt_batch_graph_launch.argtypes = [ctypes.c_void_p]
    L.ntt_batch_graph_destroy.argtypes = [ctypes.c_void_p]
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


class DevicePool:
    """One device allocation carved into u64 sub-buffers (offsets in words)."""

    def __init__(self, n_words):
        self.buf = DeviceBuffer(n_words)
        self.used = 0

    def take(self, n):
        off = self.used
        self.used += max(1, int(n))
        assert self.used <= self.buf.n, "pool exhausted"
        return off

    def ptr(self, off):
        return self.buf.ptr + 8 * off

    def upload(self, off, a):
        a = _as_u64(a)
        if len(a):
            check(lib().ntt_memcpy_h2d(self.ptr(off), a.ctypes.data, len(a) * 8))

    def download(self, off, n):
        out = np.zeros(n, dtype=np.uint64)
        if n:
            check(lib().ntt_memcpy_d2h(out.ctypes.data, self.ptr(off), n * 8))
        return out

    def close(self):
        self.buf.close()


def make_descs(items):
    """items: (a_ptr, na, b_ptr, nb, c_ptr, n_out, add_ptr, nadd) tuples -> ctypes array."""
    arr = (MulDesc * len(items))()
    for i, (pa, na, pb, nb, pc, n_out, padd, nadd) in enumerate(items):
        arr[i] = MulDesc(pa, pb, pc, padd, na, nb, nadd, n_out)
    return arr


def mul_batch_device(descs, sync=True):
    f = lib().ntt_mul_batch_device if sync else lib().ntt_mul_batch_device_async
    check(f(descs, len(descs)))


class BatchGraph:
    """A batch captured into a hipGraph; launch() replays it."""

    def __init__(self, descs):
        h = ctypes.c_void_p()
        check(lib().ntt_batch_graph_create(descs, len(descs), ctypes.byref(h)))
        self.h = h

    def launch(self):
        check(lib().ntt_batch_graph_launch(self.h))

    def close(self):
        if self.h:
            lib().ntt_batch_graph_destroy(self.h)
            self.h = None


def sync():
    check(lib().ntt_engine_sync())


def split_tree(a, b, use_graph=True, runs=1):
    """(P, B) of sum_{k=a}^{b-1} 1/k! from the level-batched device tree."""
    L = lib()
    init()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.ntt_split_tree_create.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.POINTER(vp)]
    L.ntt_split_tree_run.argtypes = [vp]
    L.ntt_split_tree_root_caps.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
    L.ntt_split_tree_result.argtypes = [vp, _u64p, sz, ctypes.POINTER(sz), _u64p, sz, ctypes.POINTER(sz)]
    L.ntt_split_tree_destroy.argtypes = [vp]
    h = vp()
    check(L.ntt_split_tree_create(a, b, 1 if use_graph else 0, ctypes.byref(h)))
    try:
        for _ in range(runs):
            check(L.ntt_split_tree_run(h))
        cp, cb = sz(), sz()
        check(L.ntt_split_tree_root_caps(h, ctypes.byref(cp), ctypes.byref(cb)))
        P = np.zeros(cp.value, np.uint64)
        B = np.zeros(cb.value, np.uint64)
        n_p, n_b = sz(), sz()
        check(L.ntt_split_tree_result(h, _p(P), cp.value, ctypes.byref(n_p), _p(B), cb.value, ctypes.byref(n_b)))
        return P[:n_p.value].copy(), B[:n_b.value].copy()
    finally:
        L.ntt_split_tree_destroy(h)


class BigIntC(ctypes.Structure):
    """BigInt of include/bigint.h."""
    _fields_ = [("limbs", _u64p), ("n_limbs", ctypes.c_int), ("cap", ctypes.c_int), ("managed", ctypes.c_int)]


class SplitResultC(ctypes.Structure):
    _fields_ = [("P", BigIntC), ("B", BigIntC)]


def _bigint_to_np(b):
    return np.ctypeslib.as_array(b.limbs, shape=(b.n_limbs,)).copy() if b.n_limbs else np.zeros(0, np.uint64)


def binary_split_ref_api(a, b, batched=False):
    """(P, B) through the reference-compatible binary_split / binary_split_batched."""
    L = lib()
    init()
    L.binary_split.argtypes = [ctypes.POINTER(SplitResultC), ctypes.c_long, ctypes.c_long, ctypes.c_void_p]
    L.binary_split_batched.argtypes = [ctypes.POINTER(SplitResultC), ctypes.c_long, ctypes.c_long]
    L.split_free.argtypes = [ctypes.POINTER(SplitResultC)]
    sr = SplitResultC()
    if batched:
        L.binary_split_batched(ctypes.byref(sr), a, b)
    else:
        L.binary_split(ctypes.byref(sr), a, b, None)
    P, B = _bigint_to_np(sr.P), _bigint_to_np(sr.B)
    L.split_free(ctypes.byref(sr))
    return P, B


class EStats(ctypes.Structure):
    _fields_ = [("t_tree", ctypes.c_double), ("t_recip", ctypes.c_double), ("t_div", ctypes.c_double),
                ("t_radix", ctypes.c_double), ("t_total", ctypes.c_double), ("n_terms", ctypes.c_long),
                ("products", ctypes.c_ulong)]


def compute_e(n_digits, batched_tree=True):
    """'2.718...' with n_digits significant digits (reference compute_e output value)."""
    L = lib()
    init()
    L.ntt_compute_e.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                ctypes.POINTER(EStats)]
    # an uninitialised buffer and one ASCII decode of it: at 10^8 digits a zeroed ctypes
    # buffer and its .value copy cost more host time than the division
    buf = np.empty(n_digits + 2, dtype=np.uint8)
    st = EStats()
    check(L.ntt_compute_e(n_digits, 1 if batched_tree else 0, buf.ctypes.data_as(ctypes.c_char_p),
                          n_digits + 2, ctypes.byref(st)))
    # the library writes "2." and n_digits - 1 digits, then a NUL
    return str(memoryview(buf)[:n_digits + 1], "ascii"), {k: getattr(st, k) for k, _ in EStats._fields_}


def get_plan(log_len):
    """The pass plan in effect for length 2^log_len: a tuple of pass sizes (log2), first pass first."""
    L = lib()
    L.ntt_engine_get_plan.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    out = (ctypes.c_int * 8)()
    n = L.ntt_engine_get_plan(log_len, out, 8)
    return tuple(out[i] for i in range(n))


def set_pass_split(log_len, split):
    """An explicit pass split for length 2^log_len (() clears it)."""
    L = lib()
    L.ntt_engine_set_pass_split.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    arr = (ctypes.c_int * max(1, len(split)))(*split)
    check(L.ntt_engine_set_pass_split(log_len, arr, len(split)))


def autotune(min_log=11, max_log=27, save=None, load=None):
    """Run (or load) the autotuner; returns (threshold, {log_len: pass plan})."""
    L = lib()
    init()
    L.ntt_engine_autotune.argtypes = [ctypes.c_int, ctypes.c_int]
    L.ntt_engine_autotune_save.argtypes = [ctypes.c_char_p]
    L.ntt_engine_autotune_load.argtypes = [ctypes.c_char_p]
    if load:
        check(L.ntt_engine_autotune_load(load.encode()))
    else:
        check(L.ntt_engine_autotune(min_log, max_log))
    if save:
        check(L.ntt_engine_autotune_save(save.encode()))
    return L.ntt_engine_get_threshold(), {l: get_plan(l) for l in range(min_log, max_log + 1)}


def poly_mul(f, g, negacyclic=False):
    """f * g mod (X^n - 1) (or X^n + 1): (n, 3) uint64 array, row k = coefficient k as a
    little-endian 192-bit residue modulo p0 p1 p2 (ntt_poly.h)."""
    L = lib()
    init()
    f, g = _as_u64(f), _as_u64(g)
    if len(f) != len(g):
        raise ValueError("f and g must have the same length")
    L.ntt_poly_mul.argtypes = [_u64p, _u64p, _u64p, ctypes.c_size_t, ctypes.c_int]
    out = np.zeros(3 * len(f), dtype=np.uint64)
    check(L.ntt_poly_mul(_p(out), _p(f), _p(g), len(f), 1 if negacyclic else 0))
    return out.reshape(-1, 3)


def set_tile_batch(tiles):
    """Most tiles one block of a forward / inverse pass works through; returns the old value."""
    L = lib()
    L.ntt_engine_set_tile_batch.argtypes = [ctypes.c_int]
    old = L.ntt_engine_get_tile_batch()
    check(L.ntt_engine_set_tile_batch(tiles))
    return old


def set_xtab(on):
    """Cross twiddles of passes with at most 2^20-residue column blocks from
    per-block tables (True) or from in-pass Montgomery chains (False);
    returns the old setting."""
    L = lib()
    L.ntt_engine_set_xtab.argtypes = [ctypes.c_int]
    return bool(L.ntt_engine_set_xtab(1 if on else 0))


def get_tile_batch():
    """Current tiles per block of the long forward / inverse passes."""
    return lib().ntt_engine_get_tile_batch()


class ServeStats(ctypes.Structure):
    _fields_ = [("makespan_s", ctypes.c_double), ("batches", ctypes.c_size_t), ("max_inflight", ctypes.c_size_t),
                ("completed", ctypes.c_size_t), ("spectrum_hits", ctypes.c_size_t),
                ("spectra_made", ctypes.c_size_t)]


def _stats_dict(st):
    return {k: getattr(st, k) for k, _ in ServeStats._fields_}


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_inflight", ctypes.c_int), ("max_batch", ctypes.c_uint32)]


def _serve_lib():
    L = lib()
    vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64
    L.ntt_serve_open.argtypes = [ctypes.POINTER(ServeConfig), ctypes.POINTER(vp)]
    L.ntt_serve_submit.argtypes = [vp, ctypes.POINTER(MulDesc), ctypes.POINTER(u64)]
    L.ntt_serve_poll.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.ntt_serve_wait.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    L.ntt_serve_pin.argtypes = [vp, vp, ctypes.c_uint32]
    L.ntt_serve_unpin.argtypes = [vp, vp, ctypes.c_uint32]
    L.ntt_serve_get_stats.argtypes = [vp, ctypes.POINTER(ServeStats)]
    L.ntt_serve_close.argtypes = [vp]
    L.ntt_serve_replay.argtypes = [ctypes.POINTER(MulDesc), ctypes.POINTER(ctypes.c_double), sz, sz, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ServeStats)]
    return L


class Server:
    """A continuous-batching server (include/ntt_serve.h): submit() from any
    thread, then poll() or wait() on the ticket."""

    def __init__(self, hbm_budget, max_inflight=3, max_batch=0):
        init()
        self.L = _serve_lib()
        cfg = ServeConfig(hbm_budget, max_inflight, max_batch)
        h = ctypes.c_void_p()
        check(self.L.ntt_serve_open(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def submit(self, desc):
        """desc: a MulDesc; returns its ticket."""
        t = ctypes.c_uint64()
        check(self.L.ntt_serve_submit(self.h, ctypes.byref(desc), ctypes.byref(t)))
        return t.value

    def _done(self, f, ticket):
        st, td = ctypes.c_int(), ctypes.c_double()
        r = f(self.h, ticket, ctypes.byref(st), ctypes.byref(td))
        if r < 0:
            check(-r)
        return (r == 1), st.value, td.value

    def poll(self, ticket):
        """(done, status, t_done seconds since open); a done ticket is retired."""
        return self._done(self.L.ntt_serve_poll, ticket)

    def wait(self, ticket):
        """(status, t_done) once the request is complete."""
        _, st, td = self._done(self.L.ntt_serve_wait, ticket)
        return st, td

    def pin(self, ptr, n):
        check(self.L.ntt_serve_pin(self.h, ptr, n))

    def unpin(self, ptr, n):
        check(self.L.ntt_serve_unpin(self.h, ptr, n))

    def stats(self):
        st = ServeStats()
        check(self.L.ntt_serve_get_stats(self.h, ctypes.byref(st)))
        return _stats_dict(st)

    def close(self):
        if self.h:
            self.L.ntt_serve_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def serve_replay(descs, arrival, hbm_budget, slots=3):
    """Replay an arrival stream of batch descriptors through a server; returns (t_done, stats)."""
    L = _serve_lib()
    init()
    arr = np.ascontiguousarray(arrival, dtype=np.float64)
    out = np.zeros(len(descs), dtype=np.float64)
    st = ServeStats()
    check(L.ntt_serve_replay(descs, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(descs), hbm_budget,
                             slots, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(st)))
    return out, _stats_dict(st)


def split_tree_time(a, b, mode, runs=1):
    """Seconds per evaluation of the tree in the given mode (1 graph, 0 eager batches, 2 per product)."""
    import time
    L = lib()
    init()
    vp = ctypes.c_void_p
    L.ntt_split_tree_create.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, ctypes.POINTER(vp)]
    L.ntt_split_tree_run.argtypes = [vp]
    L.ntt_split_tree_destroy.argtypes = [vp]
    h = vp()
    check(L.ntt_split_tree_create(a, b, mode, ctypes.byref(h)))
    try:
        check(L.ntt_split_tree_run(h))       # warm
        t0 = time.perf_counter()
        for _ in range(runs):
            check(L.ntt_split_tree_run(h))
        return (time.perf_counter() - t0) / runs
    finally:
        L.ntt_split_tree_destroy(h)
