"""CPU tests of the C ABI: the library loads and exports every function that
include/*.h declares; host-side BigInt helpers behave like the reference's
(ntt/lib/arith/bigint.c) on Python-int checks.  No GPU call is made."""
import ctypes
import glob
import os
import random
import re
import subprocess

import pytest

import cryptography_amd as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"static inline[^{]*\{[^}]*\}", "", txt)
        for m in re.finditer(r"^[A-Za-z_][\w \t\*]*?\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", txt, re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.add(name)
    return sorted(names)


def test_headers_declare_the_reference_boundary():
    names = set(declared_functions())
    for must in ("bigint_mul", "ntt_bigint_mul", "ntt_mul_init", "ntt_mul_teardown",
                 "bigint_mul_get_threshold", "bigint_mul_set_threshold", "ntt_get_gpu_dispatch_count",
                 "ntt_gpu_emergency_reset", "bigint_alloc", "bigint_free", "bigint_add", "bigint_sqr",
                 "ntt_mul_device", "ntt_mul_limbs"):
        assert must in names, must


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", E.lib_path], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


# ---- host BigInt helpers (pure host code, no device call) -------------------

class BigInt(ctypes.Structure):
    _fields_ = [("limbs", ctypes.POINTER(ctypes.c_uint64)), ("n_limbs", ctypes.c_int),
                ("cap", ctypes.c_int), ("managed", ctypes.c_int)]


@pytest.fixture(scope="module")
def L():
    L = E.lib()
    L.bigint_alloc.restype = BigInt
    L.bigint_alloc.argtypes = [ctypes.c_int]
    for f in ("bigint_add", "bigint_sub"):
        getattr(L, f).argtypes = [ctypes.POINTER(BigInt)] * 3
    for f in ("bigint_shl", "bigint_shr"):
        getattr(L, f).argtypes = [ctypes.POINTER(BigInt), ctypes.POINTER(BigInt), ctypes.c_int]
    L.bigint_mul_u64.argtypes = [ctypes.POINTER(BigInt), ctypes.POINTER(BigInt), ctypes.c_uint64]
    L.bigint_add_u64.argtypes = [ctypes.POINTER(BigInt), ctypes.c_uint64]
    L.bigint_sub_u64.argtypes = [ctypes.POINTER(BigInt), ctypes.c_uint64]
    L.bigint_cmp.argtypes = [ctypes.POINTER(BigInt)] * 2
    L.bigint_bits.argtypes = [ctypes.POINTER(BigInt)]
    L.bigint_copy.argtypes = [ctypes.POINTER(BigInt)] * 2
    L.bigint_free.argtypes = [ctypes.POINTER(BigInt)]
    return L


def mk(L, v):
    n = max(1, (v.bit_length() + 63) // 64)
    b = L.bigint_alloc(n)
    for i in range(n):
        b.limbs[i] = (v >> (64 * i)) & (2 ** 64 - 1)
    b.n_limbs = (v.bit_length() + 63) // 64
    return b


def val(b):
    return sum(int(b.limbs[i]) << (64 * i) for i in range(b.n_limbs))


def test_bigint_helpers_vs_python(L):
    rng = random.Random(9)
    vals = [0, 1, 2 ** 64 - 1, 2 ** 64, 2 ** 128 - 1] + [rng.getrandbits(rng.randrange(1, 700)) for _ in range(60)]
    for x in vals:
        y = rng.choice(vals)
        a, b, c = mk(L, x), mk(L, y), L.bigint_alloc(1)
        L.bigint_add(ctypes.byref(c), ctypes.byref(a), ctypes.byref(b))
        assert val(c) == x + y and (c.n_limbs == 0 or c.limbs[c.n_limbs - 1] != 0)
        hi, lo = (x, y) if x >= y else (y, x)
        A, B = mk(L, hi), mk(L, lo)
        L.bigint_sub(ctypes.byref(c), ctypes.byref(A), ctypes.byref(B))
        assert val(c) == hi - lo
        assert L.bigint_cmp(ctypes.byref(A), ctypes.byref(B)) == (hi > lo) - (hi < lo)
        s = rng.randrange(0, 300)
        L.bigint_shl(ctypes.byref(c), ctypes.byref(a), s)
        assert val(c) == x << s
        L.bigint_shr(ctypes.byref(c), ctypes.byref(a), s)
        assert val(c) == x >> s
        v = rng.getrandbits(64)
        L.bigint_mul_u64(ctypes.byref(c), ctypes.byref(a), v)
        assert val(c) == x * v
        assert L.bigint_bits(ctypes.byref(a)) == x.bit_length()
        # aliasing: c == a
        L.bigint_add(ctypes.byref(a), ctypes.byref(a), ctypes.byref(b))
        assert val(a) == x + y
        L.bigint_shr(ctypes.byref(a), ctypes.byref(a), 7)
        assert val(a) == (x + y) >> 7
        L.bigint_add_u64(ctypes.byref(b), 2 ** 64 - 1)
        assert val(b) == y + 2 ** 64 - 1
        L.bigint_sub_u64(ctypes.byref(b), 5)
        assert val(b) == y + 2 ** 64 - 6
        for z in (a, b, c, A, B):
            L.bigint_free(ctypes.byref(z))
