"""CPU tests that pin the oracle (oracle/oracle.c) before anything is checked
against it.  Each check is independent of the oracle's own formulas: Python
integers, closed forms, brute force on tiny sizes, GMP, and the reference's own
outputs (tests/golden/, written from the compiled reference)."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import _oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OEIS_E = ("2.71828182845904523536028747135266249775724709369995957496696762772407663035354759"
          "45713821785251664274274663919320030599218174135966290435729003342952605956307381"
          "3232862794349076323382988075319525101901")  # ntt/scripts/test_compute_e.sh:25


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):   # deterministic below 3.3e24
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def factor_small(n):
    f, d = {}, 2
    while d * d <= n:
        while n % d == 0:
            f[d] = f.get(d, 0) + 1
            n //= d
        d += 1
    if n > 1:
        f[n] = f.get(n, 0) + 1
    return f


# ---- primes and roots -------------------------------------------------------

def test_primes_form_and_generators():
    L = O.orc()
    P = O.primes()
    assert len(set(P)) == 3 and P[0] == max(P) and 2 ** 60 - P[0] < 2 ** 45
    for i, p in enumerate(P):
        k = L.orc_prime_k(i)
        # just below 2^60 and 1 mod 2^32: 2^60 - p < 2^60 / 17 (one-step reduction of
        # any 64-bit value) and < 2^45 for p0 (products of 2^29 coefficients below p0)
        assert 17 * (2 ** 60 - p) < 2 ** 60 and p % 2 ** 32 == 1 and is_prime(p)
        c = (p - 1) >> k
        assert c * 2 ** k + 1 == p and c % 2 == 1 and k >= 32
        # generator: g^((p-1)/q) != 1 for every prime q | p-1 (p-1 = c * 2^k, c < 2^28)
        g = L.orc_prime_gen(i)
        qs = set(factor_small(c)) | {2}
        assert all(is_prime(q) for q in qs)
        assert all(pow(g, (p - 1) // q, p) != 1 for q in qs)
        for log_n in (1, 10, 27, k):
            w = L.orc_root(i, log_n)
            assert pow(w, 2 ** log_n, p) == 1 and pow(w, 2 ** (log_n - 1), p) == p - 1


# ---- modular primitives -----------------------------------------------------

def test_modops_random_and_corners():
    L = O.orc()
    rng = random.Random(7)
    for p in O.primes():
        vals = [0, 1, 2, p - 1, p - 2, p // 2, (p + 1) // 2] + [rng.randrange(p) for _ in range(300)]
        for a in vals:
            b = rng.randrange(p)
            assert L.orc_addmod(a, b, p) == (a + b) % p
            assert L.orc_submod(a, b, p) == (a - b) % p
            assert L.orc_mulmod(a, b, p) == a * b % p
            wp = L.orc_shoup_pre(b, p)
            assert wp == (b << 64) // p
            # Shoup accepts any a < 2^64 (lazy inputs up to 16p)
            for x in (a, a + p, a + 3 * p, 2 ** 64 - 1):
                if x >= 2 ** 64:
                    continue
                lazy = L.orc_shoup_mul_lazy(x, b, wp, p)
                assert lazy < 2 * p and lazy % p == x * b % p
                assert L.orc_shoup_mul(x, b, wp, p) == x * b % p
        assert L.orc_invmod(3, p) * 3 % p == 1


# ---- NTT ---------------------------------------------------------------------

def dft_py(x, w, p):
    n = len(x)
    return [sum(int(x[j]) * pow(w, j * k, p) for j in range(n)) % p for k in range(n)]


@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 6])
def test_ntt_vs_naive_dft(log_n):
    for pi, p in enumerate(O.primes()):
        x = O.splitmix(50 + log_n + pi, 1 << log_n) % np.uint64(p)
        got = O.ntt_forward(x, log_n, pi)
        w = O.orc().orc_root(pi, log_n)
        assert [int(v) for v in got] == dft_py(x, w, p)
        assert np.array_equal(O.ntt_inverse(got, log_n, pi), x)


def test_ntt_known_answers():
    # KA1..KA5 of ntt/app/compute_e/test_ntt_kernel.hip:14-27 on the build's primes
    log_n, n = 10, 1024
    for pi, p in enumerate(O.primes()):
        w = O.orc().orc_root(pi, log_n)
        delta = np.zeros(n, np.uint64); delta[0] = 1
        assert np.all(O.ntt_forward(delta, log_n, pi) == 1)                       # delta -> ones
        e1 = np.zeros(n, np.uint64); e1[1] = 1
        spec = O.ntt_forward(e1, log_n, pi)
        assert [int(v) for v in spec[:8]] == [pow(w, k, p) for k in range(8)]    # e1 -> w^k
        assert np.array_equal(O.ntt_inverse(spec, log_n, pi), e1)                 # round trip
        prod = O.ntt_inverse((O.ntt_forward(delta, log_n, pi).astype(object) * spec.astype(object)
                              % p).astype(np.uint64), log_n, pi)
        assert np.array_equal(prod, e1)                                           # delta * e1 = e1


def test_ntt_matches_reference_spectra():
    cases = json.load(open(os.path.join(GOLD, "ntt_spectra.json")))
    assert len(cases) == 9
    for c in cases:
        pi, log_n, q = c["prime"], c["log_n"], int(c["q"], 16)
        assert q == O.primes()[pi]
        assert int(c["omega"], 16) == O.orc().orc_root(pi, log_n)
        x = O.splitmix(c["input_seed"], 1 << log_n) % np.uint64(q)
        y = O.ntt_forward(x, log_n, pi)
        if "spectrum" in c:
            assert [hex(int(v)) for v in y] == c["spectrum"]
        assert hashlib.sha256(y.astype("<u8").tobytes()).hexdigest() == c["spectrum_sha256"]


def test_cyclic_convolution_vs_schoolbook():
    log_n, n = 7, 128
    for pi, p in enumerate(O.primes()):
        a = O.splitmix(3, n) % np.uint64(p)
        b = O.splitmix(4, n) % np.uint64(p)
        fa, fb = O.ntt_forward(a, log_n, pi), O.ntt_forward(b, log_n, pi)
        c = O.ntt_inverse(((fa.astype(object) * fb.astype(object)) % p).astype(np.uint64), log_n, pi)
        want = [0] * n
        for i in range(n):
            for j in range(n):
                want[(i + j) % n] = (want[(i + j) % n] + int(a[i]) * int(b[j])) % p
        assert [int(v) for v in c] == want


# ---- Garner and carry ----------------------------------------------------------

def test_garner3_crt():
    P = O.primes()
    M = P[0] * P[1] * P[2]
    rng = random.Random(11)
    xs = [0, 1, M - 1, P[0], P[0] * P[1], 2 ** 64 - 1, 2 ** 128 + 5] + [rng.randrange(M) for _ in range(500)]
    for x in xs:
        r = [x % p for p in P]
        assert O.garner3(*r) == x
    # edge residue tuples of ntt/lib/test_garner.cpp:117-124: all zero, all p-1, single lane
    assert O.garner3(0, 0, 0) == 0
    assert O.garner3(P[0] - 1, P[1] - 1, P[2] - 1) == M - 1
    for lane in range(3):
        r = [0, 0, 0]; r[lane] = 1
        x = O.garner3(*r)
        assert [x % p for p in P] == r and x < M


def _pack_py(X, n_out):
    v = sum(int(x) << (16 * k) for k, x in enumerate(X))
    return O.int_to_limbs(v % (1 << (64 * n_out)), n_out), v >> (64 * n_out)


def test_pack16_and_cla():
    import ctypes
    L = O.orc()
    rng = random.Random(5)
    for trial in range(40):
        n = rng.randrange(1, 300)
        kind = trial % 4
        if kind == 0:
            X = [rng.randrange(2 ** 59) for _ in range(n)]
        elif kind == 1:   # long carry chains: all-ones limbs everywhere
            X = [0xFFFF] * n
        elif kind == 2:
            X = [2 ** 59 - 1] * n
        else:
            X = [rng.choice([0, 0xFFFF, 2 ** 48 - 1, rng.randrange(2 ** 62)]) for _ in range(n)]
        n_out = (n + 3) // 4 + 2
        Xa = np.zeros(3 * n, np.uint64)
        Xa[0::3] = np.array(X, dtype=np.uint64)
        want, hi = _pack_py(X, n_out)
        out = np.zeros(n_out, np.uint64)
        spill = L.orc_pack16(O._ptr(Xa), n, O._ptr(out), n_out)
        assert spill == 0 and hi == 0 and np.array_equal(out, want)
        out2 = np.zeros(n_out, np.uint64)
        assert L.orc_pack16_cla(O._ptr(Xa), n, O._ptr(out2), n_out) == 0
        assert np.array_equal(out2, want)
    # carry chain across many all-ones limbs: 2^(64m) - 1 + 1
    m = 50
    X = [0xFFFF] * (4 * m) + [0]
    X[0] = 0x10000      # adds one more unit -> ripples through every limb
    want, _ = _pack_py(X, m + 2)
    Xa = np.zeros(3 * len(X), np.uint64); Xa[0::3] = X
    out = np.zeros(m + 2, np.uint64)
    assert L.orc_pack16_cla(O._ptr(Xa), len(X), O._ptr(out), m + 2) == 0
    assert np.array_equal(out, want)
    del ctypes


# ---- whole products --------------------------------------------------------------

@pytest.mark.parametrize("na,nb", [(1, 1), (1, 5), (3, 2), (17, 64), (64, 64), (100, 37), (1000, 999),
                                   (4096, 4096), (8192, 3)])
def test_mul_ntt3_vs_python_and_gmp(na, nb):
    a = O.splitmix(na * 7 + 1, na)
    b = O.splitmix(nb * 13 + 2, nb)
    want = O.int_to_limbs(O.limbs_to_int(a) * O.limbs_to_int(b), na + nb)
    assert np.array_equal(O.mul_ntt3(a, b), want)
    assert np.array_equal(O.gmp_mul(a, b), want)
    if na * nb <= 1 << 16:
        assert np.array_equal(O.mul_schoolbook(a, b), want)


def test_mul_all_ones_operands():
    # all-max operands (test_e2e_oracle.c "all-max" cells): largest coefficients
    for n in (1, 2, 63, 64, 65, 1024):
        a = np.full(n, 0xFFFFFFFFFFFFFFFF, np.uint64)
        want = O.int_to_limbs(O.limbs_to_int(a) ** 2, 2 * n)
        assert np.array_equal(O.mul_ntt3(a, a), want)
        assert np.array_equal(O.mul_schoolbook(a, a), want)


def test_mul_matches_reference_schoolbook_products():
    d = json.load(open(os.path.join(GOLD, "ref_arith.json")))
    for c in d["bigint_mul"]:
        a = O.splitmix(c["seed_a"], c["na"])
        b = O.splitmix(c["seed_b"], c["nb"])
        got = O.limbs_to_int(O.mul_ntt3(a, b))
        assert hex(got) == c["product"]
        assert hex(O.limbs_to_int(O.mul_schoolbook(a, b))) == c["product"]


def test_mul_ntt3_multithreaded_2_22_bits():
    a = O.random_operand(1, 1 << 22)
    b = O.random_operand(2, 1 << 22)
    got = O.mul_ntt3(a, b, threads=3)
    assert np.array_equal(got, O.gmp_mul(a, b))
    assert O.orc().orc_log_len(len(a), len(b)) == 19


@pytest.mark.parametrize("na,nb,threads", [(1, 1, 4), (3, 5, 2), (1000, 999, 8), (40000, 3, 4),
                                           (1 << 16, (1 << 16) - 7, 8), (1 << 13, 1 << 13, 1)])
def test_mul_ntt3_all_core_restatement(na, nb, threads):
    """The all-core form of the restatement (the CPU baseline at full size:
    cache-blocked small stages, threaded large ones, one twiddle table per
    prime and direction) gives GMP's products and the sequential form's."""
    a = O.splitmix(na * 31 + 1, na)
    b = O.splitmix(nb * 37 + 2, nb)
    got = O.mul_ntt3_par(a, b, threads)
    assert np.array_equal(got, O.gmp_mul(a, b))
    if na + nb <= 4096:
        assert np.array_equal(got, O.mul_ntt3(a, b))


def test_mul_ntt3_all_core_all_ones():
    a = np.full(20000, 0xFFFFFFFFFFFFFFFF, np.uint64)
    assert np.array_equal(O.mul_ntt3_par(a, a, 8), O.gmp_mul(a, a))


# ---- product tree and e ------------------------------------------------------------

def test_binary_split_matches_reference_cells():
    d = json.load(open(os.path.join(GOLD, "ref_arith.json")))
    for c in d["binary_split"]:
        P, B = O.binary_split(c["a"], c["b"])
        assert hex(O.limbs_to_int(P)) == c["P"], (c["a"], c["b"])
        assert hex(O.limbs_to_int(B)) == c["B"], (c["a"], c["b"])
        if c["b"] > c["a"]:
            Pg, Bg = O.binary_split(c["a"], c["b"], use_gmp=True)
            assert O.limbs_to_int(Pg) == O.limbs_to_int(P) and O.limbs_to_int(Bg) == O.limbs_to_int(B)
    for k, v in d["split_terms_needed"].items():
        assert O.orc().orc_split_terms_needed(int(k)) == v


def test_binary_split_closed_form():
    # B = (b-1)!/(a-1)!,  P = sum_{k=a}^{b-1} (b-1)!/k!
    from math import factorial
    for a, b in [(1, 30), (3, 41), (10, 77)]:
        P, B = O.binary_split(a, b)
        assert O.limbs_to_int(B) == factorial(b - 1) // factorial(a - 1)
        assert O.limbs_to_int(P) == sum(factorial(b - 1) // factorial(k) for k in range(a, b))


def test_e_digits_gmp_vs_oeis_and_reference():
    assert O.gmp_e_digits(200) == OEIS_E.replace(".", "")[:200]
    d = json.load(open(os.path.join(GOLD, "ref_e.json")))
    for k, ent in d.items():
        n = int(k)
        if n > 10000:
            continue
        digits = O.gmp_e_digits(n)
        s = digits[0] + "." + digits[1:]
        assert s == ent["e"]
        assert hashlib.sha256(s.encode()).hexdigest() == ent["sha256"]
