"""GPU tests of the continuous-batching service (include/ntt_serve.h): every
result equals the same product computed alone (and GMP), whatever mix it was
served in, under budgets that make requests wait for admission, with
submitters on several threads, and with a pinned operand whose resident
spectrum replaces its transform."""
import json
import threading

import numpy as np
import pytest

import _oracle as O
import _workloads as W
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu


def make_requests(pool, n, seed, lo_bits=10, hi_bits=20, bits=None):
    rng = np.random.default_rng(seed)
    items, host = [], []
    for i in range(n):
        if bits is None:
            ba = int(2 ** (lo_bits + (hi_bits - lo_bits) * rng.random()))
            bb = int(2 ** (lo_bits + (hi_bits - lo_bits) * rng.random()))
        else:
            ba, bb = int(bits[i, 0]), int(bits[i, 1])
        a = O.random_operand(seed * 100000 + 2 * i, ba)
        b = O.random_operand(seed * 100000 + 2 * i + 1, bb)
        oa, ob, oc = pool.take(len(a)), pool.take(len(b)), pool.take(len(a) + len(b))
        pool.upload(oa, a)
        pool.upload(ob, b)
        items.append((pool.ptr(oa), len(a), pool.ptr(ob), len(b), pool.ptr(oc), len(a) + len(b), None, 0))
        host.append((a, b, oc))
    return EN.make_descs(items), host


def alone(pool, a, b):
    """The same product through the plain device multiply (ntt_mul_device)."""
    with E.DeviceBuffer.from_array(a) as da, E.DeviceBuffer.from_array(b) as db, \
            E.DeviceBuffer(len(a) + len(b)) as dc:
        E.mul_device(dc, da, len(a), db, len(b))
        return dc.download()


@pytest.mark.parametrize("budget_mb,slots", [(4096, 3), (48, 2)])
def test_served_results_equal_alone(budget_mb, slots):
    pool = EN.DevicePool(40_000_000)
    try:
        descs, host = make_requests(pool, 200, seed=budget_mb)
        rng = np.random.default_rng(7)
        arrival = np.cumsum(rng.exponential(1.0 / 20000.0, size=len(descs)))
        t_done, st = EN.serve_replay(descs, arrival, budget_mb << 20, slots=slots)
        assert st["batches"] >= 1 and st["completed"] == len(descs)
        assert np.all(t_done >= arrival - 1e-6)
        for (a, b, oc) in host:
            got = pool.download(oc, len(a) + len(b))
            assert np.array_equal(got, O.gmp_mul(a, b))
            assert np.array_equal(got, E.mul(a, b))        # the same product alone
    finally:
        pool.close()


@pytest.mark.timeout(600)
def test_config5_requests_up_to_2p28_bits_under_64gb():
    """The serving configuration's sizes: operands log-uniform in [2^10, 2^28]
    bits, Poisson arrivals, a 64 GB budget, three slots.  Each result equals
    GMP's product and the same product computed alone."""
    n = 24
    bits = W.config5_bits(n, seed=5)
    words = int(((bits + 63) // 64).sum() * 2) + 64
    pool = EN.DevicePool(words)
    try:
        descs, host = make_requests(pool, n, seed=55, bits=bits)
        arrival = np.cumsum(np.random.default_rng(8).exponential(1.0 / 400.0, size=n))
        t_done, st = EN.serve_replay(descs, arrival, 64 << 30, slots=3)
        assert st["completed"] == n and np.all(t_done >= arrival - 1e-6)
        print("batches", st["batches"], "max in flight", st["max_inflight"], flush=True)
        for i, (a, b, oc) in enumerate(host):
            got = pool.download(oc, len(a) + len(b))
            assert np.array_equal(got, O.gmp_mul(a, b)), i
            assert np.array_equal(got, alone(pool, a, b)), i
    finally:
        pool.close()


def test_submitters_on_several_threads():
    """Four threads submit 40 requests each while earlier ones run; every
    ticket completes with NTT_OK and every product equals GMP's."""
    pool = EN.DevicePool(30_000_000)
    try:
        descs, host = make_requests(pool, 160, seed=21, lo_bits=10, hi_bits=21)
        with EN.Server(1 << 30, max_inflight=4) as srv:
            tickets = [None] * len(descs)
            errors = []

            def worker(k):
                try:
                    for i in range(k, len(descs), 4):
                        tickets[i] = srv.submit(descs[i])
                except Exception as e:        # surfaced in the main thread
                    errors.append(e)

            th = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            assert not errors
            assert len(set(tickets)) == len(descs)
            for t in tickets:
                status, t_done = srv.wait(t)
                assert status == 0 and t_done > 0
            st = srv.stats()
            assert st["completed"] == len(descs) and st["batches"] >= 1
        for (a, b, oc) in host:
            assert np.array_equal(pool.download(oc, len(a) + len(b)), O.gmp_mul(a, b))
    finally:
        pool.close()


def test_pinned_operand_reuses_its_spectrum():
    """A pinned 2^20-bit operand multiplied by 30 operands of three transform
    lengths, on either side of the product and between unpinned requests:
    its forward transform is made once per length and reused, and every
    product equals GMP's and the unpinned path's."""
    pool = EN.DevicePool(20_000_000)
    try:
        p = O.random_operand(99, 1 << 20)
        op = pool.take(len(p))
        pool.upload(op, p)
        rng = np.random.default_rng(3)
        items, host = [], []
        for i in range(40):
            pinned = i % 4 != 3
            bits = int(2 ** (14 + 8 * rng.random()))
            a = O.random_operand(500 + i, bits)
            oa = pool.take(len(a))
            pool.upload(oa, a)
            if pinned:
                b, ob = p, op
            else:
                b = O.random_operand(900 + i, bits // 3 + 64)
                ob = pool.take(len(b))
                pool.upload(ob, b)
            oc = pool.take(len(a) + len(b))
            if pinned and i % 2:     # the pinned operand as a
                items.append((pool.ptr(ob), len(b), pool.ptr(oa), len(a), pool.ptr(oc), len(a) + len(b), None, 0))
            else:
                items.append((pool.ptr(oa), len(a), pool.ptr(ob), len(b), pool.ptr(oc), len(a) + len(b), None, 0))
            host.append((a, b, oc, pinned))
        descs = EN.make_descs(items)
        n_pinned = sum(1 for h in host if h[3])
        lens = {E.log_len(len(a), len(b)) for (a, b, _, pinned) in host if pinned}
        with EN.Server(2 << 30, max_inflight=3) as srv:
            srv.pin(pool.ptr(op), len(p))
            tickets = [srv.submit(d) for d in descs]
            for t in tickets:
                assert srv.wait(t)[0] == 0
            st = srv.stats()
            assert st["spectrum_hits"] == n_pinned
            assert st["spectra_made"] == len(lens)
            srv.unpin(pool.ptr(op), len(p))
            with pytest.raises(E.NttError):
                srv.unpin(pool.ptr(op), len(p))       # no longer pinned
        for (a, b, oc, pinned) in host:
            got = pool.download(oc, len(a) + len(b))
            assert np.array_equal(got, O.gmp_mul(a, b))
    finally:
        pool.close()


def test_pinned_spectrum_follows_the_pass_plan():
    """A resident spectrum is the pinned operand's rows after all forward passes
    but the last, so it depends on the length's pass plan.  After the plan of
    that length changes (a new split), the server makes a new spectrum instead
    of reading the old one under the new plan: every product stays exact."""
    pool = EN.DevicePool(8_000_000)
    log_len = 21
    before = EN.get_plan(log_len)
    try:
        p = O.random_operand(7, 1 << 24)        # 2^24-bit operands: L = 2^21
        op = pool.take(len(p))
        pool.upload(op, p)
        host = []
        for i in range(6):
            a = O.random_operand(700 + i, (1 << 24) - 1000 * i)
            oa = pool.take(len(a))
            pool.upload(oa, a)
            oc = pool.take(len(a) + len(p))
            host.append((a, oa, oc))
            assert E.log_len(len(a), len(p)) == log_len
        with EN.Server(2 << 30, max_inflight=2) as srv:
            srv.pin(pool.ptr(op), len(p))
            for k, split in enumerate(((7, 7, 7), (9, 6, 6), (6, 6, 9))):
                EN.set_pass_split(log_len, split)
                for (a, oa, oc) in host[2 * k: 2 * k + 2]:
                    d = EN.make_descs([(pool.ptr(oa), len(a), pool.ptr(op), len(p), pool.ptr(oc),
                                        len(a) + len(p), None, 0)])[0]
                    assert srv.wait(srv.submit(d))[0] == 0
                assert srv.stats()["spectra_made"] == k + 1
            srv.unpin(pool.ptr(op), len(p))
        for (a, oa, oc) in host:
            assert np.array_equal(pool.download(oc, len(a) + len(p)), O.gmp_mul(a, p))
    finally:
        EN.set_pass_split(log_len, ())
        if EN.get_plan(log_len) != before:
            EN.set_pass_split(log_len, before)
        pool.close()


def test_close_finishes_queued_requests():
    pool = EN.DevicePool(10_000_000)
    try:
        descs, host = make_requests(pool, 60, seed=31, lo_bits=12, hi_bits=20)
        srv = EN.Server(1 << 30, max_inflight=2)
        for d in descs:
            srv.submit(d)
        srv.close()                  # returns once every queued request is done
        for (a, b, oc) in host:
            assert np.array_equal(pool.download(oc, len(a) + len(b)), O.gmp_mul(a, b))
    finally:
        pool.close()


def test_request_larger_than_budget_is_rejected():
    pool = EN.DevicePool(3_000_000)
    try:
        descs, _ = make_requests(pool, 1, seed=3, lo_bits=22, hi_bits=22.1)
        with pytest.raises(E.NttError) as ei:
            EN.serve_replay(descs, np.zeros(1), 1 << 20, slots=1)
        assert ei.value.code == 2
        with EN.Server(1 << 20, max_inflight=1) as srv:
            with pytest.raises(E.NttError) as ei:
                srv.submit(descs[0])
            assert ei.value.code == 2
            with pytest.raises(E.NttError):
                srv.poll(12345)                       # unknown ticket
    finally:
        pool.close()


def test_tree_modes_agree():
    P1, B1 = EN.split_tree(1, 20001, use_graph=True)
    t_graph = EN.split_tree_time(1, 20001, 1)
    t_each = EN.split_tree_time(1, 20001, 2)
    assert t_graph > 0 and t_each > 0
    Pg, Bg = O.binary_split(1, 20001, use_gmp=True)
    assert np.array_equal(P1, Pg) and np.array_equal(B1, Bg)


def _clear_splits(lo, hi):
    for l in range(lo, hi + 1):
        EN.set_pass_split(l, ())


def test_autotune_keeps_results_exact(tmp_path):
    """The tuner's crossover and pass splits (every split it tried for 2^11 ..
    2^21 is a valid plan: parts of 2^3..2^10 points summing to the length);
    products stay exact under them; save/load give the same settings."""
    try:
        thr, plans = EN.autotune(11, 21, save=str(tmp_path / "tune.json"))
        assert 0 <= thr <= 128
        for l, p in plans.items():
            assert sum(p) == l and all(3 <= s <= 10 for s in p), (l, p)
        for n in (3, 50, 200, 5000, 40000):
            a = O.splitmix(n, n)
            b = O.splitmix(n + 1, n + 3)
            assert np.array_equal(E.mul(a, b), O.gmp_mul(a, b))
        doc = json.loads((tmp_path / "tune.json").read_text())
        assert {int(k): tuple(v) for k, v in doc["pass_split"].items()} == plans
        assert set(doc["split_us"]) == {str(l) for l in plans}
        _clear_splits(11, 21)
        thr2, plans2 = EN.autotune(11, 21, load=str(tmp_path / "tune.json"))
        assert thr2 == thr and plans2 == plans
    finally:
        _clear_splits(11, 21)


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
        _clear_splits(24, 24)
