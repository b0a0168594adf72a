"""The no-allocation multiply (ntt_mul_device_noalloc) against the allocating
device multiply (ntt_mul_device) through the C ABI: bit-identical products on
consecutive calls, each later one smaller than the one before (the static
scratch keeps the earlier call's residues), plus the size limit.
"""
import ctypes

import numpy as np
import pytest

import _oracle as O
import cryptography_amd as E

pytestmark = pytest.mark.gpu


def _noalloc(dc, da, na, db, nb):
    L = E.lib()
    L.ntt_mul_device_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_void_p, ctypes.c_size_t]
    E.engine.check(L.ntt_mul_device_noalloc(dc.ptr, da.ptr, na, db.ptr, nb))


# (bits of a, bits of b): the largest accepted size (two 2^22-bit operands,
# transform 2^19), then smaller ones, down to the schoolbook path
CALLS = [((1 << 22), (1 << 22)), ((1 << 20) + 77, (1 << 18) + 5), (5000 * 64, 3000 * 64), (64 * 64, 60 * 64)]


def test_noalloc_matches_device_multiply():
    E.init()
    for i, (ba, bb) in enumerate(CALLS):
        a = O.random_operand(700 + 2 * i, ba)
        b = O.random_operand(701 + 2 * i, bb)
        na, nb = len(a), len(b)
        print(f"call {i}: {na} x {nb} limbs, log_len {E.log_len(na, nb)}", flush=True)
        with E.DeviceBuffer.from_array(a) as da, E.DeviceBuffer.from_array(b) as db, \
                E.DeviceBuffer(na + nb) as d1, E.DeviceBuffer(na + nb) as d2:
            E.mul_device(d1, da, na, db, nb)
            _noalloc(d2, da, na, db, nb)
            want = d1.download()
            assert np.array_equal(d2.download(), want), f"call {i}"
            if i < 2:
                assert np.array_equal(want, O.gmp_mul(a, b))


def test_noalloc_rejects_products_beyond_its_scratch():
    E.init()
    with E.DeviceBuffer(1) as x:
        with pytest.raises(E.NttError) as ei:
            _noalloc(x, x, 1 << 16, x, (1 << 16) + 1)      # needs a 2^20 transform
        assert ei.value.code == 2
        assert "no-allocation" in str(ei.value)
