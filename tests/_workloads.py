"""The synthetic workloads of the benchmark configurations, built once for both
bench.py and the GPU parity tests, so the tests check exactly what is timed.

  config4_batch   4096 independent products, each operand 2^(12 + 10u) bits
                  (u uniform, drawn independently: unequal pairs), uniform
                  random bits with the top bit set, seed 4096
  config5_requests  serving requests, operands 2^(10 + 18u) bits (up to 2^28)
"""
import numpy as np

import _oracle as O


def config4_sizes(n=4096, seed=4096):
    """Operand bit lengths of the 4096-product batch (n x 2 int64)."""
    rng = np.random.default_rng(seed)
    return np.floor(2.0 ** (12 + 10 * rng.random((n, 2)))).astype(np.int64)


def config4_batch(EN, n=4096, seed=4096):
    """Device pool, batch descriptors and host operands of the 4096-product batch:
    host[i] = (a, b, product offset in the pool)."""
    bits = config4_sizes(n, seed)
    la, lb = (bits[:, 0] + 63) // 64, (bits[:, 1] + 63) // 64
    pool = EN.DevicePool(int((la + lb).sum() * 2) + 16 * n)
    items, host = [], []
    for i in range(n):
        a = O.random_operand(10_000 + 2 * i, int(bits[i, 0]))
        b = O.random_operand(10_000 + 2 * i + 1, int(bits[i, 1]))
        oa, ob, oc = pool.take(len(a)), pool.take(len(b)), pool.take(len(a) + len(b))
        pool.upload(oa, a)
        pool.upload(ob, b)
        items.append((pool.ptr(oa), len(a), pool.ptr(ob), len(b), pool.ptr(oc), len(a) + len(b), None, 0))
        host.append((a, b, oc))
    return pool, items, host


def config5_bits(n, seed, lo=10, hi=28):
    """Operand bit lengths of n serving requests, log-uniform in [2^lo, 2^hi]."""
    rng = np.random.default_rng(seed)
    return np.floor(2.0 ** (lo + (hi - lo) * rng.random((n, 2)))).astype(np.int64)
