"""GPU parity on the benchmark workloads themselves (through the C ABI).

  config 2  the 2^30-bit x 2^30-bit product and the square of one operand,
            every limb against GMP mpz_mul (about 12 s of GMP each)
  config 4  the 4096-product batch of bench.py (tests/_workloads.py): every
            product against GMP, and the hipGraph replay against the eager batch
Each product is checked limb for limb, as the reference's own GMP oracle test
does (ntt/app/compute_e/test_gmp_oracle.c:63-76).
"""
import numpy as np
import pytest

import _oracle as O
import _workloads as W
import cryptography_amd as E
from cryptography_amd import engine as EN

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(400)
def test_config2_2p30_bit_product_and_square_vs_gmp():
    a = O.random_operand(1, 1 << 30)
    b = O.random_operand(2, 1 << 30)
    assert E.log_len(len(a), len(b)) == 27
    with E.DeviceBuffer.from_array(a) as da, E.DeviceBuffer.from_array(b) as db, \
            E.DeviceBuffer(len(a) + len(b)) as dc:
        E.mul_device(dc, da, len(a), db, len(b))
        got = dc.download()
        assert np.array_equal(got, O.gmp_mul(a, b))
        E.sqr_device(dc, da, len(a))
        got = dc.download()
    assert np.array_equal(got, O.gmp_mul(a, a))


@pytest.mark.timeout(300)
def test_config4_4096_product_graph_vs_gmp_and_eager():
    pool, items, host = W.config4_batch(EN)
    try:
        descs = EN.make_descs(items)
        EN.mul_batch_device(descs)
        eager = [pool.download(oc, len(a) + len(b)) for (a, b, oc) in host]
        bad = [i for i, ((a, b, _), got) in enumerate(zip(host, eager)) if not np.array_equal(got, O.gmp_mul(a, b))]
        assert not bad, f"{len(bad)} products differ from GMP (first {bad[:5]})"
        g = EN.BatchGraph(descs)
        try:
            for rep in range(3):
                # clear the outputs so each replay has to write every limb again
                for (a, b, oc) in host:
                    pool.upload(oc, np.zeros(len(a) + len(b), np.uint64))
                g.launch()
                EN.sync()
                for i, (a, b, oc) in enumerate(host):
                    assert np.array_equal(pool.download(oc, len(a) + len(b)), eager[i]), (rep, i)
        finally:
            g.close()
    finally:
        pool.close()
