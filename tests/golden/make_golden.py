"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run in the build container only (it needs oracle/_ref/, built by
`make -C oracle ref` from the reference's own sources):

    python tests/golden/make_golden.py

Fixtures written (data only: inputs and the reference's outputs):
  ntt_spectra.json   reference ntt/ref/ntt_cpu.c ntt_forward on the build's
                     three primes and roots, seeded inputs, n = 8, 64, 1024
  ref_arith.json     reference binary_split cells (the cells of
                     ntt/app/compute_e/test_binary_split.c:220-257) and
                     reference bigint_mul (schoolbook) products on seeded operands
  ref_e.json         reference compute_e_host digits for d = 50..10^4 in full,
                     d = 10^5 as SHA-256
"""
import ctypes
import hashlib
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_u64p = ctypes.POINTER(ctypes.c_uint64)


class NttParams(ctypes.Structure):     # ntt/ref/ntt.h:30-38
    _fields_ = [("n", ctypes.c_uint64), ("q", ctypes.c_uint64), ("omega", ctypes.c_uint64),
                ("omega_inv", ctypes.c_uint64), ("n_inv", ctypes.c_uint64),
                ("log2_n", ctypes.c_uint32), ("reduce", ctypes.c_void_p)]


class BigInt(ctypes.Structure):        # ntt/lib/arith/bigint.h:69-74 (LIMB_BITS 64)
    _fields_ = [("limbs", _u64p), ("n_limbs", ctypes.c_int), ("cap", ctypes.c_int),
                ("managed", ctypes.c_int)]


class SplitResult(ctypes.Structure):   # ntt/app/compute_e/binary_split.h:28-31
    _fields_ = [("P", BigInt), ("B", BigInt)]


def bigint_hex(b):
    if b.n_limbs == 0:
        return "0x0"
    v = 0
    for i in range(b.n_limbs):
        v |= int(b.limbs[i]) << (64 * i)
    return hex(v)


def ntt_spectra():
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref_ntt.so"))
    L.ntt_params_init.argtypes = [ctypes.POINTER(NttParams)]
    L.ntt_alloc_twiddles.restype = ctypes.c_void_p
    L.ntt_alloc_twiddles.argtypes = [ctypes.POINTER(NttParams)]
    L.ntt_forward.argtypes = [_u64p, ctypes.c_void_p, ctypes.POINTER(NttParams)]
    libc = ctypes.CDLL(None)
    cases = []
    for pi, q in enumerate(O.primes()):
        for log_n in (3, 6, 10):
            n = 1 << log_n
            prm = NttParams(n=n, q=q, omega=O.orc().orc_root(pi, log_n))
            assert L.ntt_params_init(ctypes.byref(prm)) == 0
            seed = 1000 + 10 * pi + log_n
            x = O.splitmix(seed, n) % np.uint64(q)
            y = x.copy()
            tw = L.ntt_alloc_twiddles(ctypes.byref(prm))
            L.ntt_forward(y.ctypes.data_as(_u64p), tw, ctypes.byref(prm))
            libc.free(ctypes.c_void_p(tw))
            c = {"prime": pi, "q": hex(q), "log_n": log_n, "omega": hex(prm.omega),
                 "input_seed": seed, "input": "splitmix64(seed, n) mod q"}
            if n <= 64:
                c["spectrum"] = [hex(int(v)) for v in y]
            c["spectrum_sha256"] = hashlib.sha256(y.astype("<u8").tobytes()).hexdigest()
            cases.append(c)
    return cases


def ref_arith():
    L = ctypes.CDLL(os.path.join(REF_DIR, "libref_arith.so"))
    L.binary_split.argtypes = [ctypes.POINTER(SplitResult), ctypes.c_long, ctypes.c_long, ctypes.c_void_p]
    L.split_free.argtypes = [ctypes.POINTER(SplitResult)]
    L.bigint_alloc.restype = BigInt
    L.bigint_alloc.argtypes = [ctypes.c_int]
    L.bigint_free.argtypes = [ctypes.POINTER(BigInt)]
    L.bigint_mul.argtypes = [ctypes.POINTER(BigInt)] * 3
    L.split_terms_needed.restype = ctypes.c_long
    L.split_terms_needed.argtypes = [ctypes.c_long]
    cells = [(1, 1), (5, 5), (1, 2), (2, 3), (17, 18), (1024, 1025), (1, 3), (7, 9), (1, 5),
             (1, 9), (1, 17), (1, 65), (1, 257), (1, 4), (1, 8), (1, 13), (1, 100), (1, 200),
             (5, 50), (100, 200), (1000000, 1000001), (100000000, 100000001), (1, 500),
             (1, 1000), (1, 3500), (1, 5000)]
    split = []
    for a, b in cells:
        sr = SplitResult()
        L.binary_split(ctypes.byref(sr), a, b, None)
        split.append({"a": a, "b": b, "P": bigint_hex(sr.P), "B": bigint_hex(sr.B)})
        L.split_free(ctypes.byref(sr))
    prods = []
    for seed, na, nb in [(11, 1, 1), (12, 1, 7), (13, 5, 3), (14, 31, 64), (15, 64, 64),
                         (16, 100, 3), (17, 257, 129)]:
        a = O.splitmix(seed, na)
        b = O.splitmix(seed + 1000, nb)
        A, B, C = L.bigint_alloc(na), L.bigint_alloc(nb), L.bigint_alloc(1)
        for i in range(na):
            A.limbs[i] = int(a[i])
        for i in range(nb):
            B.limbs[i] = int(b[i])
        A.n_limbs, B.n_limbs = na, nb
        L.bigint_mul(ctypes.byref(C), ctypes.byref(A), ctypes.byref(B))
        prods.append({"seed_a": seed, "seed_b": seed + 1000, "na": na, "nb": nb,
                      "operands": "splitmix64 limbs", "product": bigint_hex(C)})
        for x in (A, B, C):
            L.bigint_free(ctypes.byref(x))
    terms = {str(d): int(L.split_terms_needed(d)) for d in (20, 70, 1020, 10020, 100020,
                                                             1000020, 100000020)}
    return {"binary_split": split, "bigint_mul": prods, "split_terms_needed": terms}


def ref_e():
    out = {}
    for d in (50, 100, 200, 1000, 10000, 100000):
        r = subprocess.run([os.path.join(REF_DIR, "compute_e_host"), "-d", str(d)],
                           capture_output=True, text=True, check=True)
        m = re.search(r"^e = (2\.[0-9]+)$", r.stdout, re.M)
        digits = m.group(1)
        ent = {"sha256": hashlib.sha256(digits.encode()).hexdigest(), "len": len(digits)}
        if d <= 10000:
            ent["e"] = digits
        out[str(d)] = ent
    return out


def main():
    with open(os.path.join(HERE, "ntt_spectra.json"), "w") as f:
        json.dump(ntt_spectra(), f, indent=1)
    with open(os.path.join(HERE, "ref_arith.json"), "w") as f:
        json.dump(ref_arith(), f, indent=1)
    with open(os.path.join(HERE, "ref_e.json"), "w") as f:
        json.dump(ref_e(), f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
