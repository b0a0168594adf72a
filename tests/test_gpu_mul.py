"""GPU parity: the HIP CRT-NTT multiply against the oracle, GMP and Python ints.

Every call goes through the C ABI of libntt_amd.so.  Results must be bit-exact.
"""
import random

import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu


def pyint_mul(a, b):
    return O.int_to_limbs(O.limbs_to_int(a) * O.limbs_to_int(b), len(a) + len(b))


def test_twiddle_tables_on_device():
    E.init()
    n = 3 * 2 * (1 << 15) * 2
    buf = np.zeros(n, dtype=np.uint64)
    E.engine.check(E.lib().ntt_engine_read_twiddles(buf.ctypes.data_as(E.engine._u64p), n))
    t = buf.reshape(3, 2, 1 << 15, 2)
    rng = random.Random(3)
    for pi, p in enumerate(O.primes()):
        w = O.orc().orc_root(pi, 30)
        for j in [0, 1, 2, 12345, (1 << 15) - 1] + [rng.randrange(1 << 15) for _ in range(50)]:
            lo, hi = pow(w, j, p), pow(w, j << 15, p)
            assert int(t[pi, 0, j, 0]) == lo and int(t[pi, 0, j, 1]) == (lo << 64) // p
            assert int(t[pi, 1, j, 0]) == hi and int(t[pi, 1, j, 1]) == (hi << 64) // p


SIZES = [(1, 1), (1, 2), (2, 1), (3, 3), (64, 64), (65, 64), (64, 65), (100, 7), (7, 100),
         (128, 128), (255, 1), (1000, 1000), (1024, 1024), (1025, 1023), (4096, 17), (5000, 4999),
         (16384, 16384), (40000, 3)]


@pytest.mark.parametrize("na,nb", SIZES)
def test_mul_random_vs_oracle(na, nb):
    a = O.splitmix(na * 1000 + nb, na)
    b = O.splitmix(nb * 7919 + na + 1, nb)
    got = E.mul(a, b)
    want = O.gmp_mul(a, b)
    assert np.array_equal(got, want)
    if na + nb <= 4096:
        assert np.array_equal(got, O.mul_ntt3(a, b))


@pytest.mark.parametrize("na,nb", [(65, 65), (300, 2000), (4096, 4096)])
def test_ntt_path_forced_small(na, nb):
    """Threshold 0 forces the NTT path even for small operands."""
    E.init()
    L = E.lib()
    saved = L.ntt_engine_get_threshold()
    L.ntt_engine_set_threshold(0)
    try:
        for na_, nb_ in [(1, 1), (2, 3), (na, nb)]:
            a = O.splitmix(na_ + 5, na_)
            b = O.splitmix(nb_ + 6, nb_)
            assert np.array_equal(E.mul(a, b), pyint_mul(a, b))
    finally:
        L.ntt_engine_set_threshold(saved)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 1024, 4096, 65536])
def test_all_ones_operands(n):
    a = np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)
    want = pyint_mul(a, a)
    assert np.array_equal(E.mul(a, a), want)
    assert np.array_equal(E.sqr(a), want)


@pytest.mark.parametrize("n", [1, 5, 64, 1000, 8192, 70000])
def test_square_vs_mul(n):
    a = O.splitmix(n + 99, n)
    got = E.sqr(a)
    assert np.array_equal(got, O.gmp_mul(a, a))


def test_zero_and_sparse_operands():
    z = np.zeros(1000, np.uint64)
    a = O.splitmix(1, 1000)
    assert np.array_equal(E.mul(z, a), np.zeros(2000, np.uint64))
    # powers of two and single-bit operands: long carry chains after the convolution
    x = np.zeros(3000, np.uint64); x[-1] = 1 << 63
    y = np.full(2500, 0xFFFFFFFFFFFFFFFF, np.uint64)
    assert np.array_equal(E.mul(x, y), pyint_mul(x, y))
    assert np.array_equal(E.mul(np.zeros(0, np.uint64), a), np.zeros(1000, np.uint64))


def test_carry_chain_across_blocks():
    # (2^(64n) - 1) * (2^(64m) + 1): ripples a carry over > 2048 limbs (several carry blocks)
    n = 9000
    a = np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)
    b = np.zeros(n, np.uint64); b[0] = 1; b[-1] = 1
    assert np.array_equal(E.mul(a, b), pyint_mul(a, b))


def test_config1_2_22_bit_product():
    """2^22-bit operands, L = 2^19: bit-exact vs GMP and the CPU oracle."""
    a = O.random_operand(1, 1 << 22)
    b = O.random_operand(2, 1 << 22)
    assert E.log_len(len(a), len(b)) == 19
    got = E.mul(a, b)
    assert np.array_equal(got, O.gmp_mul(a, b))
    assert np.array_equal(got, O.mul_ntt3(a, b, threads=3))


def test_device_api_and_async():
    a = O.random_operand(5, 1 << 20)
    b = O.random_operand(6, 1 << 19)
    want = O.gmp_mul(a, b)
    with E.DeviceBuffer.from_array(a) as da, E.DeviceBuffer.from_array(b) as db, \
            E.DeviceBuffer(len(a) + len(b)) as dc:
        E.mul_device(dc, da, len(a), db, len(b))
        assert np.array_equal(dc.download(), want)
        for _ in range(3):
            E.mul_device(dc, da, len(a), db, len(b), sync=False)
        E.engine.check(E.lib().ntt_engine_sync())
        assert np.array_equal(dc.download(), want)
        E.sqr_device(dc, da, len(a) // 2)
        h = a[: len(a) // 2]
        assert np.array_equal(dc.download(len(a)), O.gmp_mul(h, h))


def test_too_large_is_an_error():
    E.init()
    with pytest.raises(E.NttError) as ei:
        E.mul_device(E.DeviceBuffer(1), E.DeviceBuffer(1), 1 << 28, E.DeviceBuffer(1), 1 << 28)
    assert ei.value.code == 2


@pytest.mark.parametrize("na,nb", [(1 << 18, (1 << 18) + 5), (3 << 17, 1 << 16)])
def test_three_pass_transform(na, nb):
    """Transform lengths 2^21 and above run three passes (columns, columns, rows)."""
    a = O.splitmix(31 + na, na)
    b = O.splitmix(37 + nb, nb)
    assert E.log_len(na, nb) >= 21
    assert np.array_equal(E.mul(a, b), O.gmp_mul(a, b))
    assert np.array_equal(E.sqr(a), O.gmp_mul(a, a))


@pytest.mark.parametrize("na,nb,split", [(1 << 18, (1 << 18) + 5, None), (3 << 17, 1 << 16, None), (5000, 3000, None),
                                         (1 << 22, 1 << 22, None), (1 << 17, 1 << 17, (6, 8, 6)),
                                         (1 << 19, (1 << 19) - 7, (8, 8, 6))])
def test_cross_twiddle_tables_match_chains(na, nb, split):
    """Cross twiddles from per-block tables (forward column passes at store,
    inverse passes at load) and chains with their ratios from per-column
    tables (the first pass, the last inverse pass of blocks above 2^20
    residues) against the in-pass Montgomery chains with every ratio computed:
    the same products, both equal to GMP.  Lengths 2^21 (three passes: the
    middle block has tables), 2^15 (two passes: the inverse of the first pass,
    a 2^15 block, has them), 2^25 (three passes; its inverse table pass
    takes 6 tiles per block over 8192 tiles, the last blocks fewer), 2^20 split
    6/8/6 (a 64-point first pass with a ratio table) and 2^22 split 8/8/6 (the
    last inverse pass, a 2^22 block, by chain with a ratio table at 256
    points)."""
    E.init()
    a = O.splitmix(301 + na, na)
    b = O.splitmix(303 + nb, nb)
    want_ab, want_aa = O.gmp_mul(a, b), O.gmp_mul(a, a)
    log_len = E.log_len(na, nb)
    if split:
        assert sum(split) == log_len
        EN.set_pass_split(log_len, split)
    old = E.engine.set_xtab(False)
    try:
        chain_ab, chain_aa = E.mul(a, b), E.sqr(a)
        E.engine.set_xtab(True)
        tab_ab, tab_aa = E.mul(a, b), E.sqr(a)
    finally:
        E.engine.set_xtab(old)
        if split:
            EN.set_pass_split(log_len, ())
    assert np.array_equal(chain_ab, want_ab) and np.array_equal(chain_aa, want_aa)
    assert np.array_equal(tab_ab, want_ab) and np.array_equal(tab_aa, want_aa)


@pytest.mark.parametrize("tiles", [1, 2, 3, 4, 8])
def test_tile_batch_parity(tiles):
    """Forward / inverse passes with several tiles per block (the next tile's
    residues loaded while the current one is transformed): a 2^27-bit
    product (L = 2^24, 4096 tiles per row, ragged for 3) and its square."""
    E.init()
    a = O.random_operand(71, 1 << 27)
    b = O.random_operand(72, (1 << 27) - 4096)
    old = E.engine.set_tile_batch(tiles)
    try:
        assert np.array_equal(E.mul(a, b), O.gmp_mul(a, b))
        assert np.array_equal(E.sqr(b), O.gmp_mul(b, b))
    finally:
        E.engine.set_tile_batch(old)


def test_host_api_staged_transfers():
    """Host-buffer products above the staging threshold go through pinned
    64 MiB pieces (several per operand here): the result equals the
    device-resident product bit for bit, and the product modulo a prime."""
    n = 9 << 20                                      # 72 MiB per operand, 144 MiB product
    a = O.splitmix(91, n)
    b = O.splitmix(92, n)
    got = E.mul(a, b)
    da, db, dc = E.DeviceBuffer.from_array(a), E.DeviceBuffer.from_array(b), E.DeviceBuffer(2 * n)
    E.mul_device(dc, da, n, db, n)
    assert np.array_equal(got, dc.download())
    q = (1 << 61) - 1
    A, B, C = (int.from_bytes(x.tobytes(), "little") for x in (a, b, got))
    assert C % q == (A % q) * (B % q) % q
    sq = E.sqr(a)
    assert int.from_bytes(sq.tobytes(), "little") % q == (A % q) * (A % q) % q
