"""ctypes bindings to the CPU oracle (oracle/liboracle.so, oracle/libgmpref.so).

Test infrastructure only: tests, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` load these libraries as the checker.  The product
path never does.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a):
    return a.ctypes.data_as(_u64p)


def _load(name):
    path = os.path.join(ORACLE_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `make -C oracle`")
    return ctypes.CDLL(path)


_orc = None
_gmp = None


def orc():
    global _orc
    if _orc is None:
        L = _load("liboracle.so")
        L.orc_prime.restype = ctypes.c_uint64
        L.orc_prime_k.restype = ctypes.c_int
        L.orc_prime_gen.restype = ctypes.c_uint64
        L.orc_root.restype = ctypes.c_uint64
        L.orc_root.argtypes = [ctypes.c_int, ctypes.c_int]
        for f in ("orc_addmod", "orc_submod", "orc_mulmod", "orc_powmod"):
            getattr(L, f).restype = ctypes.c_uint64
            getattr(L, f).argtypes = [ctypes.c_uint64] * 3
        L.orc_invmod.restype = ctypes.c_uint64
        L.orc_invmod.argtypes = [ctypes.c_uint64] * 2
        L.orc_shoup_pre.restype = ctypes.c_uint64
        L.orc_shoup_pre.argtypes = [ctypes.c_uint64] * 2
        L.orc_shoup_mul.restype = ctypes.c_uint64
        L.orc_shoup_mul.argtypes = [ctypes.c_uint64] * 4
        L.orc_shoup_mul_lazy.restype = ctypes.c_uint64
        L.orc_shoup_mul_lazy.argtypes = [ctypes.c_uint64] * 4
        L.orc_ntt_forward.argtypes = [_u64p, ctypes.c_int, ctypes.c_int]
        L.orc_ntt_inverse.argtypes = [_u64p, ctypes.c_int, ctypes.c_int]
        L.orc_dft_naive.argtypes = [_u64p, _u64p, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_garner3.argtypes = [ctypes.c_uint64] * 3 + [_u64p]
        L.orc_split16.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t]
        L.orc_pack16.restype = ctypes.c_uint64
        L.orc_pack16.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t]
        L.orc_pack16_cla.restype = ctypes.c_int
        L.orc_pack16_cla.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t]
        L.orc_mul_ntt3.restype = ctypes.c_int
        L.orc_mul_ntt3.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p, ctypes.c_int]
        L.orc_mul_ntt3_par.restype = ctypes.c_int
        L.orc_mul_ntt3_par.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p, ctypes.c_int]
        L.orc_mul_schoolbook.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.orc_log_len.restype = ctypes.c_int
        L.orc_log_len.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
        L.orc_binary_split.argtypes = [ctypes.c_long, ctypes.c_long,
                                       ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t),
                                       ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t)]
        L.orc_split_terms_needed.restype = ctypes.c_long
        L.orc_split_terms_needed.argtypes = [ctypes.c_long]
        L.orc_splitmix_fill.argtypes = [ctypes.c_uint64, _u64p, ctypes.c_size_t]
        _orc = L
    return _orc


def gmp():
    global _gmp
    if _gmp is None:
        L = _load("libgmpref.so")
        L.gmpref_mul.restype = ctypes.c_size_t
        L.gmpref_mul.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p]
        L.gmpref_time_mul.restype = ctypes.c_double
        L.gmpref_time_mul.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, ctypes.c_int]
        L.gmpref_binary_split.argtypes = [ctypes.c_long, ctypes.c_long,
                                          ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t),
                                          ctypes.POINTER(_u64p), ctypes.POINTER(ctypes.c_size_t)]
        L.gmpref_e_digits.restype = ctypes.c_int
        L.gmpref_e_digits.argtypes = [ctypes.c_long, ctypes.c_char_p]
        L.gmpref_to_decimal.restype = ctypes.c_long
        L.gmpref_to_decimal.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
        _gmp = L
    return _gmp


# ---- numpy-level helpers -------------------------------------------------

def primes():
    L = orc()
    return [int(L.orc_prime(i)) for i in range(3)]


def splitmix(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    if n:
        orc().orc_splitmix_fill(ctypes.c_uint64(seed), _ptr(out), n)
    return out


def random_operand(seed, n_bits):
    """Uniform random n_bits-bit integer with the top bit set, as u64 limbs."""
    n = (n_bits + 63) // 64
    a = splitmix(seed, n)
    top = (n_bits - 1) % 64
    mask = np.uint64((1 << (top + 1)) - 1) if top < 63 else np.uint64(0xFFFFFFFFFFFFFFFF)
    a[-1] = (a[-1] & mask) | np.uint64(1 << top)
    return a


def limbs_to_int(a):
    return int.from_bytes(np.ascontiguousarray(a, dtype="<u8").tobytes(), "little")


def int_to_limbs(x, n):
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").copy()


def ntt_forward(a, log_n, pi):
    a = np.ascontiguousarray(a, dtype=np.uint64).copy()
    orc().orc_ntt_forward(_ptr(a), log_n, pi)
    return a


def ntt_inverse(a, log_n, pi):
    a = np.ascontiguousarray(a, dtype=np.uint64).copy()
    orc().orc_ntt_inverse(_ptr(a), log_n, pi)
    return a


def garner3(r0, r1, r2):
    out = np.zeros(3, dtype=np.uint64)
    orc().orc_garner3(r0, r1, r2, _ptr(out))
    return int(out[0]) | (int(out[1]) << 64) | (int(out[2]) << 128)


def mul_ntt3(a, b, threads=1):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    rc = orc().orc_mul_ntt3(_ptr(a), len(a), _ptr(b), len(b), _ptr(out), threads)
    if rc != 0:
        raise RuntimeError("orc_mul_ntt3 failed")
    return out


def mul_ntt3_par(a, b, threads):
    """The oracle's 3-prime NTT multiply with every transform on `threads` host threads."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    if orc().orc_mul_ntt3_par(_ptr(a), len(a), _ptr(b), len(b), _ptr(out), threads) != 0:
        raise RuntimeError("orc_mul_ntt3_par failed")
    return out


def mul_schoolbook(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    orc().orc_mul_schoolbook(_ptr(a), len(a), _ptr(b), len(b), _ptr(out))
    return out


def gmp_mul(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(len(a) + len(b), dtype=np.uint64)
    gmp().gmpref_mul(_ptr(a), len(a), _ptr(b), len(b), _ptr(out))
    return out


def _take(ptr, n):
    arr = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].copy() if n else np.zeros(0, np.uint64)
    ctypes.CDLL(None).free(ctypes.cast(ptr, ctypes.c_void_p))
    return arr


def binary_split(a, b, use_gmp=False):
    P, B = _u64p(), _u64p()
    np_, nb = ctypes.c_size_t(), ctypes.c_size_t()
    f = gmp().gmpref_binary_split if use_gmp else orc().orc_binary_split
    f(a, b, ctypes.byref(P), ctypes.byref(np_), ctypes.byref(B), ctypes.byref(nb))
    return _take(P, np_.value), _take(B, nb.value)


def gmp_e_digits(n_digits):
    buf = ctypes.create_string_buffer(n_digits + 1)
    if gmp().gmpref_e_digits(n_digits, buf) != 0:
        raise RuntimeError("gmpref_e_digits failed")
    return buf.value.decode()


def gmp_to_decimal(limbs):
    """Decimal string of the integer in the u64 limb array (GMP mpz_get_str)."""
    a = np.ascontiguousarray(limbs, dtype=np.uint64)
    cap = len(a) * 20 + 2
    buf = ctypes.create_string_buffer(cap)
    n = gmp().gmpref_to_decimal(_ptr(a), len(a), buf, cap)
    if n < 0:
        raise RuntimeError("gmpref_to_decimal: buffer too small")
    return buf.raw[:n].decode()
