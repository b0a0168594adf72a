"""GPU parity of the batched multiply and its hipGraph replay (through the C ABI).

Each product of a batch must equal the same product computed alone, and GMP.
"""
import ctypes
import random

import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu


def build_batch(pool, specs, seed):
    """specs: list of (na, nb, nadd).  Returns descriptors and host operands."""
    items, host = [], []
    for i, (na, nb, nadd) in enumerate(specs):
        a = O.splitmix(seed * 1000 + 2 * i, na)
        b = O.splitmix(seed * 1000 + 2 * i + 1, nb)
        add = O.splitmix(seed * 7777 + i, nadd) if nadd else np.zeros(0, np.uint64)
        n_out = max(na + nb, nadd) + (1 if nadd else 0)
        oa, ob, oc = pool.take(na), pool.take(nb), pool.take(n_out)
        od = pool.take(nadd) if nadd else None
        pool.upload(oa, a)
        pool.upload(ob, b)
        if nadd:
            pool.upload(od, add)
        items.append((pool.ptr(oa), na, pool.ptr(ob), nb, pool.ptr(oc), n_out,
                      pool.ptr(od) if nadd else None, nadd))
        host.append((a, b, add, oc, n_out))
    return EN.make_descs(items), host


def expected(a, b, add, n_out):
    v = O.limbs_to_int(a) * O.limbs_to_int(b) + (O.limbs_to_int(add) if len(add) else 0)
    return O.int_to_limbs(v, n_out)


MIXED = [(1, 1, 0), (3, 5, 0), (0, 9, 0), (64, 64, 0), (64, 64, 70), (65, 3, 0), (100, 100, 0),
         (100, 100, 100), (500, 700, 0), (513, 511, 300), (1000, 1000, 0), (1000, 999, 1000),
         (4096, 4000, 0), (9000, 9000, 9000), (20000, 1, 0), (17, 30000, 0), (5, 5, 12)]


def test_batch_mixed_sizes_with_addend():
    pool = EN.DevicePool(400000)
    try:
        descs, host = build_batch(pool, MIXED, seed=3)
        EN.mul_batch_device(descs)
        for (a, b, add, oc, n_out) in host:
            assert np.array_equal(pool.download(oc, n_out), expected(a, b, add, n_out))
    finally:
        pool.close()


def test_batch_equals_alone_and_gmp():
    rng = random.Random(4)
    specs = [(rng.randrange(1, 3000), rng.randrange(1, 3000), 0) for _ in range(40)]
    pool = EN.DevicePool(600000)
    try:
        descs, host = build_batch(pool, specs, seed=4)
        EN.mul_batch_device(descs)
        for (a, b, add, oc, n_out) in host:
            got = pool.download(oc, n_out)
            assert np.array_equal(got, O.gmp_mul(a, b))
            assert np.array_equal(got, E.mul(a, b))          # the same product alone
    finally:
        pool.close()


def test_batch_graph_replays_match_eager():
    specs = [(7, 9, 0), (64, 64, 64), (300, 200, 0), (2000, 2100, 0), (2000, 2100, 50), (40, 5000, 0)]
    pool = EN.DevicePool(200000)
    try:
        descs, host = build_batch(pool, specs, seed=5)
        EN.mul_batch_device(descs)
        eager = [pool.download(oc, n_out) for (_, _, _, oc, n_out) in host]
        for (a, b, add, oc, n_out), e in zip(host, eager):
            assert np.array_equal(e, expected(a, b, add, n_out))
        g = EN.BatchGraph(descs)
        try:
            for rep in range(4):
                # clobber the outputs between replays with eager work
                for (_, _, _, oc, n_out) in host:
                    pool.upload(oc, np.full(n_out, 0xDEADBEEF, np.uint64))
                x = O.splitmix(rep, 100)
                assert np.array_equal(E.mul(x, x), O.gmp_mul(x, x))
                g.launch()
                EN.sync()
                for (_, _, _, oc, n_out), e in zip(host, eager):
                    assert np.array_equal(pool.download(oc, n_out), e), f"replay {rep}"
        finally:
            g.close()
    finally:
        pool.close()


def test_batch_overflow_is_reported():
    # n_out too small for the addend carry: integrity error at sync
    pool = EN.DevicePool(1000)
    try:
        a = np.full(4, 0xFFFFFFFFFFFFFFFF, np.uint64)
        oa, oc, od = pool.take(4), pool.take(8), pool.take(8)
        pool.upload(oa, a)
        pool.upload(od, np.full(8, 0xFFFFFFFFFFFFFFFF, np.uint64))
        descs = EN.make_descs([(pool.ptr(oa), 4, pool.ptr(oa), 4, pool.ptr(oc), 8, pool.ptr(od), 8)])
        with pytest.raises(E.NttError) as ei:
            EN.mul_batch_device(descs)
        assert ei.value.code == 5
        # the engine keeps working afterwards
        x = O.splitmix(9, 50)
        assert np.array_equal(E.mul(x, x), O.gmp_mul(x, x))
    finally:
        pool.close()


@pytest.mark.parametrize("nb,na_list", [
    (1500, [1400, 1500, 1600, 1550, 1450]),          # L = 2^14: b transformed once, shared
    (40000, [30000, 35000, 39000]),                  # L = 2^19
    (200, [150, 180, 210, 250]),                     # L = 2^11: single tiles hold several products, no sharing
])
def test_batch_shared_operand(nb, na_list):
    """Products of one size class that all multiply by the same b transform b
    once (the fused pass reads product 0's spectrum for all of them); each
    result equals GMP, also with a trailing product of another b that keeps
    its class unshared."""
    pool = EN.DevicePool(8 * (nb + max(na_list)) * (len(na_list) + 2) + 4096)
    try:
        b = O.splitmix(5150, nb)
        ob = pool.take(nb)
        pool.upload(ob, b)
        items, host = [], []
        for i, na in enumerate(na_list):
            a = O.splitmix(5151 + i, na)
            oa, oc = pool.take(na), pool.take(na + nb)
            pool.upload(oa, a)
            items.append((pool.ptr(oa), na, pool.ptr(ob), nb, pool.ptr(oc), na + nb, None, 0))
            host.append((a, b, oc))
        EN.mul_batch_device(EN.make_descs(items))
        for a, bb, oc in host:
            assert np.array_equal(pool.download(oc, len(a) + len(bb)), O.gmp_mul(a, bb))
        # the same class with one product of a different b: no sharing
        b2 = O.splitmix(5160, nb)
        ob2 = pool.take(nb)
        pool.upload(ob2, b2)
        a = host[0][0]
        oa, oc = pool.take(len(a)), pool.take(len(a) + nb)
        pool.upload(oa, a)
        items.append((pool.ptr(oa), len(a), pool.ptr(ob2), nb, pool.ptr(oc), len(a) + nb, None, 0))
        EN.mul_batch_device(EN.make_descs(items))
        assert np.array_equal(pool.download(oc, len(a) + nb), O.gmp_mul(a, b2))
        for a_, bb, oc_ in host:
            assert np.array_equal(pool.download(oc_, len(a_) + len(bb)), O.gmp_mul(a_, bb))
    finally:
        pool.close()


@pytest.mark.parametrize("count,na,nb", [(16, 3000, 2000), (8, 1 << 18, (1 << 18) - 77)])
def test_batch_table_passes_many_products(count, na, nb):
    """A size class of `count` products (a multiple of 8 groups): the passes
    that read cross-twiddle tables take their tiles column slice by column
    slice across 8 groups, so tiles of different products interleave.
    L = 2^15 (two passes: the inverse of the first from its table) and
    L = 2^21 (three passes: the inner pass and its inverse).  Eager and graph
    runs, tables on and off, all equal to GMP."""
    pool = EN.DevicePool(count * (2 * (na + nb)) + 4096)
    old = EN.set_xtab(True)
    try:
        descs, host = build_batch(pool, [(na, nb, 0)] * count, seed=61 + count)
        want = [O.gmp_mul(a, b) for (a, b, _, _, _) in host]
        for on in (True, False):
            EN.set_xtab(on)
            for (_, _, _, oc, n_out) in host:
                pool.upload(oc, np.zeros(n_out, np.uint64))
            EN.mul_batch_device(descs)
            for (_, _, _, oc, n_out), w in zip(host, want):
                assert np.array_equal(pool.download(oc, n_out), w), f"eager, tables {on}"
        EN.set_xtab(True)
        g = EN.BatchGraph(descs)
        try:
            for rep in range(2):
                for (_, _, _, oc, n_out) in host:
                    pool.upload(oc, np.full(n_out, 0xDEADBEEF, np.uint64))
                g.launch()
                EN.sync()
                for (_, _, _, oc, n_out), w in zip(host, want):
                    assert np.array_equal(pool.download(oc, n_out), w), f"graph replay {rep}"
        finally:
            g.close()
    finally:
        EN.set_xtab(old)
        pool.close()


def test_handles_refuse_to_run_after_reinit():
    """A batch graph or product tree captured before ntt_engine_teardown holds
    the old engine's tables: after a re-initialisation it must return
    NTT_ERR_NOT_INIT instead of replaying (its kernels are never launched),
    while handles made after it run normally."""
    E.init()
    L = E.lib()
    pool = EN.DevicePool(100000)
    try:
        descs, host = build_batch(pool, [(300, 200, 0), (2000, 2100, 0)], seed=9)
        old = EN.BatchGraph(descs)
        tree_h = ctypes.c_void_p()
        L.ntt_split_tree_create.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_void_p)]
        L.ntt_split_tree_run.argtypes = [ctypes.c_void_p]
        L.ntt_split_tree_destroy.argtypes = [ctypes.c_void_p]
        EN.check(L.ntt_split_tree_create(1, 3001, 1, ctypes.byref(tree_h)))
        L.ntt_engine_teardown()
        E.init()
        with pytest.raises(E.NttError) as ei:
            old.launch()
        assert ei.value.code == 1
        assert L.ntt_split_tree_run(tree_h) == 1
        old.close()
        L.ntt_split_tree_destroy(tree_h)
        new = EN.BatchGraph(descs)
        new.launch()
        EN.sync()
        for (a, b, add, oc, n_out) in host:
            assert np.array_equal(pool.download(oc, n_out), expected(a, b, add, n_out))
        new.close()
    finally:
        pool.close()
