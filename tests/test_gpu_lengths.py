"""GPU parity at every transform length and at the headline size.

Every length L = 2^3 .. 2^27 (one pass up to 2^12 points, then two, then
three passes of at most 2^10 points), with unequal operands whose limb counts
are not powers of two: the product and a square, exact against GMP up to
L = 2^23, and above that through the product modulo three primes (a wrong
product would have to be a multiple of all three).  Then the 2^30-bit x
2^30-bit product and square of the bench's headline (L = 2^27), checked the
same way.
"""
import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu

# 2^61 - 1 and 2^89 - 1 (Mersenne primes), 2^64 - 2^32 + 1
MOD_PRIMES = ((1 << 61) - 1, (1 << 89) - 1, (1 << 64) - (1 << 32) + 1)
GMP_MAX_LOG = 23


def _int(x):
    return int.from_bytes(np.ascontiguousarray(x, dtype=np.uint64).tobytes(), "little")


def _check_mod(a, b, c):
    A, B, C = _int(a), _int(b), _int(c)
    for q in MOD_PRIMES:
        assert C % q == (A % q) * (B % q) % q, f"product wrong modulo {q:#x}"


def _sizes(log_len):
    """Limb counts na, nb >= 1 with 4 (na + nb) in (2^(log_len-1), 2^log_len],
    unequal from L = 2^4 up (L = 2^3 holds only 1 x 1)."""
    hi = 1 << (log_len - 2)
    total = max((1 << (log_len - 3)) + 1, hi - max(1, hi // 5))
    na = max(1, (total * 5) // 9)
    return na, total - na


@pytest.mark.parametrize("log_len", range(3, 28))
def test_every_transform_length(log_len):
    na, nb = _sizes(log_len)
    assert E.log_len(na, nb) == log_len
    a = O.splitmix(1000 + log_len, na)
    b = O.splitmix(2000 + log_len, nb)
    E.init()
    L = E.lib()
    saved = L.ntt_engine_get_threshold()
    L.ntt_engine_set_threshold(0)          # the transform path, also for the smallest operands
    try:
        c = E.mul(a, b)
        s = E.sqr(a)
    finally:
        L.ntt_engine_set_threshold(saved)
    assert len(c) == na + nb and len(s) == 2 * na
    if log_len <= GMP_MAX_LOG:
        assert np.array_equal(c, O.gmp_mul(a, b))
        assert np.array_equal(s, O.gmp_mul(a, a))
    else:
        _check_mod(a, b, c)
        _check_mod(a, a, s)


@pytest.mark.parametrize("split", [
    (7, 6), (3, 10), (10, 3), (8, 6), (9, 6), (6, 9), (10, 6), (6, 10), (9, 8), (10, 9), (8, 6, 6), (3, 3, 10),
    (9, 6, 6), (6, 6, 9), (6, 9, 6), (9, 7, 6), (6, 8, 9), (9, 9, 6), (6, 9, 9), (9, 7, 9), (9, 8, 9),
    (10, 10, 5), (5, 10, 10), (3, 9, 3, 9), (9, 10, 7),
])
def test_explicit_pass_splits(split):
    """Uneven pass splits (ntt_engine_set_pass_split, chosen by the autotuner):
    first passes of 2^3..2^10 points (radix-8 rounds or not, with and without
    cross-twiddle tables for the inner passes), short last passes; product and
    square exact against GMP (modulo three primes above 2^23), the
    plan restored afterwards."""
    log_len = sum(split)
    na, nb = _sizes(log_len)
    a = O.splitmix(5000 + log_len, na)
    b = O.splitmix(6000 + log_len, nb)
    E.init()
    before = EN.get_plan(log_len)
    EN.set_pass_split(log_len, split)
    try:
        assert EN.get_plan(log_len) == split
        c = E.mul(a, b)
        s = E.sqr(a)
    finally:
        EN.set_pass_split(log_len, ())
    assert EN.get_plan(log_len) == before
    if log_len <= GMP_MAX_LOG:
        assert np.array_equal(c, O.gmp_mul(a, b))
        assert np.array_equal(s, O.gmp_mul(a, a))
    else:
        _check_mod(a, b, c)
        _check_mod(a, a, s)


@pytest.mark.parametrize("log_len,cap", [(13, 4), (16, 4), (18, 6), (19, 5), (20, 4), (22, 5), (24, 4)])
def test_pass_cap_plans(log_len, cap):
    """A pass cap below the default (ntt_engine_set_pass_cap, or a loaded tuning
    file) splits the transform into more, smaller passes: 3 to 6 passes of 2^3
    to 2^6 points here, with several middle forward and inverse passes.
    Product and square exact against GMP; the cap is restored afterwards."""
    na, nb = _sizes(log_len)
    a = O.splitmix(3000 + log_len, na)
    b = O.splitmix(4000 + log_len, nb)
    E.init()
    L = E.lib()
    E.engine.check(L.ntt_engine_set_pass_cap(log_len, cap))
    try:
        assert L.ntt_engine_get_pass_cap(log_len) == cap
        c = E.mul(a, b)
        s = E.sqr(a)
    finally:
        E.engine.check(L.ntt_engine_set_pass_cap(log_len, 0))
    assert np.array_equal(c, O.gmp_mul(a, b))
    assert np.array_equal(s, O.gmp_mul(a, a))


def test_headline_2p30_bit_product_and_square():
    """2^30-bit x 2^30-bit (top bits set), L = 2^27: the product and the
    square modulo three primes, on device-resident operands."""
    a = O.random_operand(1, 1 << 30)
    b = O.random_operand(2, 1 << 30)
    assert len(a) == len(b) == 1 << 24 and E.log_len(len(a), len(b)) == 27
    da, db = E.DeviceBuffer.from_array(a), E.DeviceBuffer.from_array(b)
    dc = E.DeviceBuffer(2 << 24)
    E.mul_device(dc, da, len(a), db, len(b))
    _check_mod(a, b, dc.download())
    E.sqr_device(dc, da, len(a))
    _check_mod(a, a, dc.download())


@pytest.mark.timeout(600)
@pytest.mark.parametrize("log_len", [28, 29, 30])
def test_longest_transform_lengths(log_len):
    """The lengths above the headline, up to the largest the engine accepts
    (2^30 coefficients: 48 GiB of residue rows): unequal device-resident
    operands, the product and a square modulo three primes."""
    na, nb = _sizes(log_len)
    assert E.log_len(na, nb) == log_len
    print(f"log_len {log_len}: {na} x {nb} limbs", flush=True)
    a = O.splitmix(5000 + log_len, na)
    b = O.splitmix(6000 + log_len, nb)
    # the square of a has 2 na > na + nb limbs: it gets its own output buffer
    with E.DeviceBuffer.from_array(a) as da, E.DeviceBuffer.from_array(b) as db, \
            E.DeviceBuffer(na + nb) as dc, E.DeviceBuffer(2 * na) as ds:
        E.mul_device(dc, da, na, db, nb)
        _check_mod(a, b, dc.download())
        E.sqr_device(ds, da, na)
        _check_mod(a, a, ds.download())
