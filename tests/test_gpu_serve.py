"""GPU tests of the continuous-batching service: every result equals the same
product computed alone (and GMP), whatever mix it was served in, including
under an HBM budget small enough to make requests wait for admission."""
import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu


def make_requests(pool, n, seed, lo_bits=10, hi_bits=20):
    rng = np.random.default_rng(seed)
    items, host = [], []
    for i in range(n):
        ba = int(2 ** (lo_bits + (hi_bits - lo_bits) * rng.random()))
        bb = int(2 ** (lo_bits + (hi_bits - lo_bits) * rng.random()))
        a = O.random_operand(seed * 100000 + 2 * i, ba)
        b = O.random_operand(seed * 100000 + 2 * i + 1, bb)
        oa, ob, oc = pool.take(len(a)), pool.take(len(b)), pool.take(len(a) + len(b))
        pool.upload(oa, a)
        pool.upload(ob, b)
        items.append((pool.ptr(oa), len(a), pool.ptr(ob), len(b), pool.ptr(oc), len(a) + len(b), None, 0))
        host.append((a, b, oc))
    return EN.make_descs(items), host


@pytest.mark.parametrize("budget_mb,slots", [(4096, 3), (48, 2)])
def test_served_results_equal_alone(budget_mb, slots):
    pool = EN.DevicePool(40_000_000)
    try:
        descs, host = make_requests(pool, 200, seed=budget_mb)
        rng = np.random.default_rng(7)
        arrival = np.cumsum(rng.exponential(1.0 / 20000.0, size=len(descs)))
        t_done, st = EN.serve_replay(descs, arrival, budget_mb << 20, slots=slots)
        assert st["batches"] >= 1
        assert np.all(t_done >= arrival - 1e-6)
        for (a, b, oc) in host:
            got = pool.download(oc, len(a) + len(b))
            assert np.array_equal(got, O.gmp_mul(a, b))
            assert np.array_equal(got, E.mul(a, b))        # the same product alone
    finally:
        pool.close()


def test_request_larger_than_budget_is_rejected():
    pool = EN.DevicePool(3_000_000)
    try:
        descs, _ = make_requests(pool, 1, seed=3, lo_bits=22, hi_bits=22.1)
        with pytest.raises(E.NttError) as ei:
            EN.serve_replay(descs, np.zeros(1), 1 << 20, slots=1)
        assert ei.value.code == 2
    finally:
        pool.close()


def test_tree_modes_agree():
    P1, B1 = EN.split_tree(1, 20001, use_graph=True)
    t_graph = EN.split_tree_time(1, 20001, 1)
    t_each = EN.split_tree_time(1, 20001, 2)
    assert t_graph > 0 and t_each > 0
    Pg, Bg = O.binary_split(1, 20001, use_gmp=True)
    assert np.array_equal(P1, Pg) and np.array_equal(B1, Bg)


def test_autotune_keeps_results_exact(tmp_path):
    thr, caps = EN.autotune(11, 21, save=str(tmp_path / "tune.json"))
    assert 0 <= thr <= 128 and all(8 <= c <= 10 for c in caps.values())
    for n in (3, 50, 200, 5000, 40000):
        a = O.splitmix(n, n)
        b = O.splitmix(n + 1, n + 3)
        assert np.array_equal(E.mul(a, b), O.gmp_mul(a, b))
    thr2, caps2 = EN.autotune(11, 21, load=str(tmp_path / "tune.json"))
    assert thr2 == thr and caps2 == caps


def test_autotune_tile_batch(tmp_path):
    """Lengths of 2^24 and up also tune the tile batch; save/load keep it, and
    products stay exact under the chosen value."""
    saved = EN.get_tile_batch()
    try:
        EN.autotune(24, 24, save=str(tmp_path / "tune.json"))
        tb = EN.get_tile_batch()
        assert tb in (1, 2, 4)
        a = O.splitmix(1 << 21, 7)
        b = O.splitmix(1 << 21, 8)
        assert np.array_equal(E.mul(a, b), O.gmp_mul(a, b))
        EN.set_tile_batch(3)
        EN.autotune(24, 24, load=str(tmp_path / "tune.json"))
        assert EN.get_tile_batch() == tb
    finally:
        EN.set_tile_batch(saved)
