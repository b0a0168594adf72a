"""newton.h and base_convert.h entry points (the reference's
ntt/lib/arith/newton.h:17-30 and base_convert.h:19-32) against Python integer
arithmetic, which pins them exactly: floor(2^(2b)/Q), N // D, str(N), 10**n.

CPU: the host logic (cryptography_amd/csrc/{bigint,newton,base_convert}.c),
compiled here against a CPU bigint_mul over the oracle
(tests/host_stub/multiply_cpu.c).  GPU: the same functions in libntt_amd.so,
with every product on the device."""
import ctypes
import os
import random
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import cryptography_amd as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_u64p = ctypes.POINTER(ctypes.c_uint64)
if hasattr(sys, "set_int_max_str_digits"):      # Python builds that cap int <-> str conversion
    sys.set_int_max_str_digits(0)


class BigInt(ctypes.Structure):
    _fields_ = [("limbs", _u64p), ("n_limbs", ctypes.c_int), ("cap", ctypes.c_int), ("managed", ctypes.c_int)]


_BP = ctypes.POINTER(BigInt)
_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def _setup(L):
    L.bigint_alloc.restype = BigInt
    L.bigint_alloc.argtypes = [ctypes.c_int]
    L.bigint_free.argtypes = [_BP]
    L.newton_reciprocal.argtypes = [_BP, _BP]
    L.bigint_div_newton.argtypes = [_BP, _BP, _BP, _BP]
    L.base_convert_init.argtypes = [ctypes.c_int]
    L.bigint_to_decimal.argtypes = [_BP]
    L.bigint_to_decimal.restype = ctypes.c_void_p
    L.bigint_write_decimal.argtypes = [_BP, ctypes.c_int]
    L.bigint_write_decimal.restype = ctypes.c_long
    L.base_convert_pow10_exact.argtypes = [_BP, ctypes.c_long]
    return L


def mk(L, v):
    n = max(1, (v.bit_length() + 63) // 64)
    b = L.bigint_alloc(n)
    if v:
        arr = np.frombuffer(v.to_bytes(8 * n, "little"), dtype=np.uint64)
        ctypes.memmove(b.limbs, arr.ctypes.data, 8 * n)
    b.n_limbs = (v.bit_length() + 63) // 64
    return b


def val(b):
    if b.n_limbs == 0:
        return 0
    return int.from_bytes(np.ctypeslib.as_array(b.limbs, shape=(b.n_limbs,)).tobytes(), "little")


def recip(L, q):
    Q, x = mk(L, q), L.bigint_alloc(1)
    L.newton_reciprocal(ctypes.byref(x), ctypes.byref(Q))
    r = val(x)
    L.bigint_free(ctypes.byref(Q))
    L.bigint_free(ctypes.byref(x))
    return r


def div(L, n, d, rec=None):
    N, D, q = mk(L, n), mk(L, d), L.bigint_alloc(1)
    R = mk(L, rec if rec is not None else recip(L, d))
    L.bigint_div_newton(ctypes.byref(q), ctypes.byref(N), ctypes.byref(D), ctypes.byref(R))
    r = val(q)
    for b in (N, D, q, R):
        L.bigint_free(ctypes.byref(b))
    return r


def to_dec(L, n):
    N = mk(L, n)
    p = L.bigint_to_decimal(ctypes.byref(N))
    L.bigint_free(ctypes.byref(N))
    if not p:
        return None
    s = ctypes.string_at(p).decode()
    _libc.free(p)
    return s


def recip_cases(rng, max_bits):
    qs = [1, 2, 3, 5, 10, 2 ** 62, 2 ** 63 - 1, 2 ** 64 - 1, 2 ** 64, 2 ** 64 + 1, 2 ** 100, 3 ** 200,
          10 ** 300, 2 ** 1000 - 1, 2 ** 1000]
    for bits in (61, 62, 63, 64, 65, 127, 128, 129, 200, 1000, 4095, 4096):
        qs.append(rng.getrandbits(bits) | (1 << (bits - 1)))
    while len(qs) < 40:
        bits = rng.randrange(2, max_bits)
        qs.append(rng.getrandbits(bits) | (1 << (bits - 1)))
    return qs


def check_recip(L, rng, max_bits):
    for q in recip_cases(rng, max_bits):
        b = q.bit_length()
        assert recip(L, q) == (1 << (2 * b)) // q, q.bit_length()


def check_div(L, rng, max_bits):
    for _ in range(40):
        bd = rng.randrange(1, max_bits)
        d = rng.getrandbits(bd) | (1 << (bd - 1))
        for bn in (max(1, bd - 3), bd, 2 * bd - 1, 2 * bd, 2 * bd + 1, 3 * bd + 5, 5 * bd):
            n = rng.getrandbits(bn) | (1 << (bn - 1))
            assert div(L, n, d) == n // d, (bn, bd)
        k = rng.getrandbits(bd)
        for n in (d * k, d * k - 1 if d * k else 0, d * k + d - 1):
            assert div(L, n, d) == n // d
    # the contract allows a reciprocal one below the floor
    for _ in range(10):
        bd = rng.randrange(64, max_bits)
        d = rng.getrandbits(bd) | (1 << (bd - 1))
        n = rng.getrandbits(2 * bd)
        r = (1 << (2 * bd)) // d - 1
        assert div(L, n, d, rec=r) == n // d
    # powers of two (the reference's floor-1 case)
    for e in (1, 63, 64, 65, 500):
        d = 1 << e
        n = rng.getrandbits(2 * e + 7)
        assert div(L, n, d) == n >> e


def check_decimal(L, rng, max_digits, level):
    L.base_convert_init(level)
    vals = [0, 1, 9, 10, 11, 99, 100, 10 ** 15, 10 ** 16 - 1, 10 ** 16, 10 ** 16 + 1, 2 ** 64 - 1, 2 ** 64,
            10 ** 1000 + 7, 10 ** 999 - 1, 10 ** 17 * 3]
    for d in (1, 5, 16, 17, 31, 32, 33, 100, 1000, 2047, 2048, 2049, max_digits):
        vals.append(rng.randrange(10 ** (d - 1), 10 ** d))
    for v in vals:
        assert to_dec(L, v) == str(v), len(str(v))
    # beyond the cache: NULL
    assert to_dec(L, 10 ** (2 ** (level + 1))) is None
    assert to_dec(L, 10 ** (2 ** (level + 1)) - 1) == "9" * 2 ** (level + 1)
    # write_decimal
    v = rng.randrange(10 ** 300, 10 ** 301)
    with tempfile.TemporaryFile() as f:
        N = mk(L, v)
        n = L.bigint_write_decimal(ctypes.byref(N), f.fileno())
        L.bigint_free(ctypes.byref(N))
        f.seek(0)
        assert n == 301 and f.read().decode() == str(v)
    # pow10_exact
    for n in (0, 1, 2, 7, 64, 1000, 2 ** level, 2 ** (level + 1) - 1):
        out = L.bigint_alloc(1)
        L.base_convert_pow10_exact(ctypes.byref(out), n)
        assert val(out) == 10 ** n, n
        L.bigint_free(ctypes.byref(out))


# ---- CPU: host logic against a CPU multiply --------------------------------

@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("hostlib")
    so = os.path.join(d, "libhost_arith.so")
    src = os.path.join(ROOT, "cryptography_amd", "csrc")
    subprocess.run(["gcc", "-O2", "-shared", "-fPIC", "-fopenmp", "-std=c11",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"),
                    os.path.join(src, "bigint.c"), os.path.join(src, "newton.c"), os.path.join(src, "base_convert.c"),
                    os.path.join(ROOT, "tests", "host_stub", "multiply_cpu.c"), os.path.join(ROOT, "oracle", "oracle.c"),
                    "-o", so], check=True)
    return _setup(ctypes.CDLL(so))


def test_host_reciprocal_exact(host_lib):
    check_recip(host_lib, random.Random(11), 3000)


def test_host_division_exact(host_lib):
    check_div(host_lib, random.Random(12), 1500)


def test_host_decimal_conversion(host_lib):
    check_decimal(host_lib, random.Random(13), 3000, 11)


# ---- GPU: the library, products on the device ------------------------------

@pytest.fixture(scope="module")
def gpu_lib():
    E.init()
    return _setup(E.lib())


@pytest.mark.gpu
def test_gpu_reciprocal_exact(gpu_lib):
    rng = random.Random(21)
    check_recip(gpu_lib, rng, 20000)
    for bits in (1 << 17, (1 << 20) + 12345):
        q = rng.getrandbits(bits) | (1 << (bits - 1))
        assert recip(gpu_lib, q) == (1 << (2 * bits)) // q


@pytest.mark.gpu
def test_gpu_division_exact(gpu_lib):
    rng = random.Random(22)
    check_div(gpu_lib, rng, 6000)
    d = rng.getrandbits(300000) | (1 << 299999)
    n = rng.getrandbits(900000)
    assert div(gpu_lib, n, d) == n // d


@pytest.mark.gpu
def test_gpu_decimal_device_path(gpu_lib):
    """Above 256 limbs bigint_to_decimal runs ntt_to_decimal (the fraction
    (N + 1/2) / 10^D through the device radix tree): powers of ten and of two
    and their neighbours (all-9 and all-0 digit runs), the threshold, and
    random values, against str(); then a 2*10^6-digit value against GMP."""
    gpu_lib.base_convert_init(20)
    try:
        _device_decimal_cases(gpu_lib)
    finally:
        gpu_lib.base_convert_teardown()      # later tests start from an empty cache


def _device_decimal_cases(gpu_lib):
    rng = random.Random(24)
    cases = []
    for d in (4932, 4933, 4940, 5000, 19 * 300, 19 * 300 + 1, 12345, 77777):
        cases += [10 ** d - 1, 10 ** d, 10 ** d + 1, 5 * 10 ** (d - 1), 10 ** d // 2, rng.randrange(10 ** (d - 1), 10 ** d)]
    for bits in (16384, 16385, 65536, 65537, 200003):
        cases += [(1 << bits) - 1, 1 << bits, (1 << bits) + 1, rng.getrandbits(bits) | (1 << (bits - 1))]
    for v in cases:
        assert to_dec(gpu_lib, v) == str(v), (v.bit_length(), len(str(v)))
    L = gpu_lib
    L.ntt_to_decimal.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_size_t)]
    import _oracle as O
    limbs = O.splitmix(77, 104000)                     # about 2.0e6 digits
    cap = int(64 * len(limbs) * 0.30103) + 2
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_size_t()
    assert L.ntt_to_decimal(limbs.ctypes.data_as(_u64p), len(limbs), buf, cap, ctypes.byref(n)) == 0
    assert buf.raw[:n.value].decode() == O.gmp_to_decimal(limbs)
    zero = np.zeros(3, np.uint64)
    assert L.ntt_to_decimal(zero.ctypes.data_as(_u64p), 3, buf, cap, ctypes.byref(n)) == 0
    assert buf.raw[:n.value] == b"0"


@pytest.mark.gpu
def test_gpu_decimal_conversion(gpu_lib):
    rng = random.Random(23)
    check_decimal(gpu_lib, rng, 20000, 15)
    gpu_lib.base_convert_init(16)                  # 10^(2^17) > any 100000-digit value
    v = rng.randrange(10 ** 99999, 10 ** 100000)
    assert to_dec(gpu_lib, v) == str(v)
