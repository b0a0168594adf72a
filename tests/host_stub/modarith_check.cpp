/*
 * Host check of the modular primitives in cryptography_amd/csrc/modarith.h
 * against 128-bit integer arithmetic, for the three NTT primes: the value and
 * the documented output range of the Montgomery product, the Shoup product
 * (its small-operand form and its truncated-quotient form), the one-step reduction of any 64-bit value, the
 * compile-time roots, the lazy butterfly groups of the transform rounds, and
 * the CRT tests that three residues are those of one
 * integer below p0, or below 2 p0 (against Garner's digits).
 * Prints "bad N" and exits non-zero on any mismatch.
 */
#include "../../cryptography_amd/csrc/lazy_group.h"
#include "../../cryptography_amd/csrc/modarith.h"

#include <stdio.h>

using namespace nttc;
typedef unsigned __int128 u128;

static uint64_t rs = 0x9e3779b97f4a7c15ull;
static uint64_t rnd()
{
    rs ^= rs << 13;
    rs ^= rs >> 7;
    rs ^= rs << 17;
    return rs;
}

static int bad = 0;
static void expect(bool ok, const char *what, int pi, uint64_t a, uint64_t b)
{
    if (!ok && bad++ < 10)
        printf("%s p%d a=%016llx b=%016llx\n", what, pi, (unsigned long long)a, (unsigned long long)b);
}

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t p)
{
    uint64_t r = 1;
    for (; e; e >>= 1, b = (uint64_t)((u128)b * b % p))
        if (e & 1) r = (uint64_t)((u128)r * b % p);
    return r;
}

template <int PI>
static void check(long iters)
{
    const uint64_t p = prime(PI);
    const uint64_t R = (uint64_t)(((u128)1 << 64) % p);   /* 2^64 mod p */
    const uint64_t edges[] = {0, 1, 2, p - 1, p, p + 1, 2 * p - 1, 2 * p, 3 * p - 1, 3 * p, 4 * p - 1,
                              0xffffffffull, 0x100000000ull, p >> 1, p >> 32};
    auto mont = [&](uint64_t a, uint64_t b) {
        const uint64_t r = mont_mul_lazy<PI>(a, b);
        expect((u128)(r % p) * R % p == (u128)a * b % p, "mont value", PI, a, b);
        expect(r > 0 && r < 3 * p, "mont range (any a, b<2p)", PI, a, b);
        if (a < 8 * p && b < 2 * p) expect(r < 2 * p, "mont range (a<8p, b<2p)", PI, a, b);
        if (b < p) expect(r < 2 * p, "mont range (any a, b<p)", PI, a, b);
    };
    auto shoup = [&](uint64_t a, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        const uint64_t r = shoup_lazy<PI>(a, w, wp);
        expect(r < 2 * p && r % p == (uint64_t)((u128)a * w % p), "shoup", PI, a, w);
    };
    /* Shoup with the truncated quotient: any a, result below 4p */
    auto shoup4 = [&](uint64_t a, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        const uint64_t r = shoup_lazy4<PI>(a, w, wp);
        expect(r < 4 * p && r % p == (uint64_t)((u128)a * w % p), "shoup4", PI, a, w);
    };
    /* small-operand Shoup: the same quotient as the full one, so the same value */
    auto shoup_s = [&](uint32_t a, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        const uint64_t r = shoup_small<PI>(a, w, wp);
        expect(r < 2 * p && r % p == (uint64_t)((u128)a * w % p), "shoup small", PI, a, w);
        expect(r == shoup_lazy<PI>(a, w, wp), "shoup small quotient", PI, a, w);
    };
    /* any 64-bit value to [0, 2p) in one step */
    auto red = [&](uint64_t x) {
        const uint64_t r = reduce_any<PI>(x);
        expect(r < 2 * p && r % p == x % p, "reduce_any", PI, x, r);
    };
    for (uint64_t x : edges) red(x);
    for (uint64_t q = 0; q < 16; q++)
        for (uint64_t x : {q << 60, (q << 60) - 1, (q << 60) + 1, q * p, q * p - 1, q * p + 2 * p - 1})
            red(x);
    red(~0ull);
    red(16 * p - 1);
    for (long i = 0; i < iters; i++) red(rnd());
    for (uint64_t a : edges)
        for (uint64_t b : edges) {
            if (b < 2 * p) mont(a, b);
            if (b < p) shoup(a, b);
            if (b < p) shoup4(a, b);
            if (b < p && a <= 0xffffffffull) shoup_s((uint32_t)a, b);
        }
    for (long i = 0; i < iters; i++) {
        mont(rnd(), rnd() % (2 * p));
        mont(rnd() % (8 * p), rnd() % (2 * p));
        mont(~0ull - (rnd() & 0xffff), 2 * p - 1 - (rnd() & 0xffff));
        mont(rnd() % (4 * p), rnd() % (2 * p));
        mont(rnd() % (2 * p), rnd() % (2 * p));
        mont(rnd() % (4 * p), rnd() % p);
        shoup(rnd(), rnd() % p);
        shoup4(rnd(), rnd() % p);
        shoup4(~0ull - (rnd() & 0xffffff), p - 1 - (rnd() & 0xffff));
        shoup4(rnd() | 0xffffffff00000000ull, rnd() % p);
        shoup_s((uint32_t)rnd(), rnd() % p);
        shoup_s((uint32_t)(rnd() & 0xffff), rnd() % p);
    }
    /* compile-time roots: w of order exactly 2^LOGORD, and w^E */
    using T = ConstTw<PI, 12, 1>;
    expect(powmod(T::w, 1u << 12, p) == 1 && powmod(T::w, 1u << 11, p) == p - 1, "root order", PI, T::w, 0);
    using T5 = ConstTw<PI, 12, 5>;
    expect(T5::w == powmod(T::w, 5, p) && T5::wp == (uint64_t)(((u128)T5::w << 64) / p), "root power", PI, T5::w, 0);
}

/* Lazy butterfly groups (lazy_group.h) against the same butterfly network in
 * exact arithmetic mod p: every plan the kernels use and more, inputs at and
 * just below their bounds as well as random, random table twiddles (also the
 * negated-twiddle inverse form) and the compile-time roots.  Each output must
 * be congruent to the exact one and below its planned bound. */
struct RandTw {
    static constexpr bool NEG = false;
    TwPair t[3][8];
    template <int S, int I>
    TwPair get() const { return t[S][I]; }
};
struct RandTwNeg : RandTw {
    static constexpr bool NEG = true;
};

template <class TW, int RL, int S = 0, int I = 0>
static void tw_table(const TW &tw, uint64_t (&w)[3][8])
{
    if constexpr (S < RL) {
        if constexpr (I < (1 << RL)) {
            w[S][I] = tw.template get<S, I>().w;
            tw_table<TW, RL, S, I + 1>(tw, w);
        } else {
            tw_table<TW, RL, S + 1, 0>(tw, w);
        }
    }
}

template <int PI, int RL, bool DIT, bool TW1, int BIN, int BOUT, int J0, class TW>
static void check_lazy_one(const TW &tw, long iters)
{
    const uint64_t p = prime(PI);
    constexpr int R = 1 << RL;
    constexpr LzPlan PL = LzRound<RL, DIT, TW1, BIN, BOUT, J0>::plan;
    uint64_t w[3][8];
    tw_table<TW, RL>(tw, w);
    for (long it = 0; it < iters; it++) {
        uint64_t x[8], e[8];
        for (int i = 0; i < R; i++) {
            const uint64_t top = (uint64_t)BIN * p;
            const int mode = (int)(rnd() % 4);
            x[i] = mode == 0 ? top - 1 - (rnd() & 0xfff) : mode == 1 ? top - 1 : rnd() % top;
            e[i] = x[i] % p;
        }
        lz_group<PI, RL, DIT, TW1, BIN, BOUT, J0>(x, tw);
        for (int j = J0; j < RL; j++) {
            const int S = DIT ? RL - 1 - j : j, half = R >> (S + 1);
            for (int i = 0; i < R; i++) {
                if (i & half) continue;
                const bool one = TW1 && (i & (half - 1)) == 0;
                const uint64_t wv = one ? 1 : w[S][i], a = e[i], b = e[i + half];
                if (!DIT) {
                    e[i] = (a + b) % p;
                    e[i + half] = (uint64_t)((u128)((a + p - b) % p) * wv % p);
                } else {
                    const uint64_t t = (uint64_t)((u128)b * wv % p);
                    const bool neg = TW::NEG && !one;
                    e[i] = neg ? (a + p - t) % p : (a + t) % p;
                    e[i + half] = neg ? (a + t) % p : (a + p - t) % p;
                }
            }
        }
        for (int i = 0; i < R; i++) {
            expect(x[i] % p == e[i], "lazy group value", PI, (uint64_t)(RL * 1000 + DIT * 100 + TW1 * 10 + i), x[i]);
            expect(PL.out[i] <= BOUT && (u128)x[i] < (u128)PL.out[i] * p, "lazy group bound", PI,
                   (uint64_t)(RL * 1000 + DIT * 100 + TW1 * 10 + i), x[i]);
        }
    }
}

template <int PI, int RL>
static void check_lazy_rl(long iters)
{
    const uint64_t p = prime(PI);
    RandTw tw;
    RandTwNeg twn;
    for (int s = 0; s < 3; s++)
        for (int i = 0; i < 8; i++) {
            const uint64_t v = (s + i) % 3 == 0 ? p - 1 : rnd() % p;
            tw.t[s][i] = twn.t[s][i] = TwPair{v, (uint64_t)(((u128)v << 64) / p)};
        }
    check_lazy_one<PI, RL, false, false, 2, 2, 0>(tw, iters);
    check_lazy_one<PI, RL, false, true, 2, 2, 0>(ConstTwF<PI, RL, false>{}, iters);
    check_lazy_one<PI, RL, false, true, 2, 8, 0>(ConstTwF<PI, RL, false>{}, iters);
    check_lazy_one<PI, RL, false, true, 2, LZ_MAX, 0>(ConstTwF<PI, RL, false>{}, iters);
    check_lazy_one<PI, RL, true, false, 4, 4, 0>(tw, iters);
    check_lazy_one<PI, RL, true, false, 4, LZ_MAX, 0>(tw, iters);
    check_lazy_one<PI, RL, true, false, 4, 4, 0>(twn, iters);
    check_lazy_one<PI, RL, true, false, 4, LZ_MAX, 0>(twn, iters);
    check_lazy_one<PI, RL, true, true, 2, 4, 0>(ConstTwF<PI, RL, true>{}, iters);
    check_lazy_one<PI, RL, true, true, 4, 4, 0>(ConstTwF<PI, RL, true>{}, iters);
    check_lazy_one<PI, RL, true, true, 4, LZ_MAX, 0>(ConstTwF<PI, RL, true>{}, iters);
    check_lazy_one<PI, RL, true, false, LZ_MAX, 4, 0>(tw, iters);
    /* inverse rounds chained unreduced: inputs up to 16p */
    check_lazy_one<PI, RL, true, false, 6, LZ_MAX, 0>(tw, iters);
    check_lazy_one<PI, RL, true, false, 8, LZ_MAX, 0>(twn, iters);
    check_lazy_one<PI, RL, true, false, 10, LZ_MAX, 0>(tw, iters);
    check_lazy_one<PI, RL, true, false, 12, LZ_MAX, 0>(twn, iters);
    check_lazy_one<PI, RL, true, false, 14, LZ_MAX, 0>(tw, iters);
    check_lazy_one<PI, RL, true, false, LZ_MAX, LZ_MAX, 0>(twn, iters);
    check_lazy_one<PI, RL, true, true, 4, LZ_MAX, 0>(ConstTwF<PI, RL, true>{}, iters);
    if constexpr (RL == 3) check_lazy_one<PI, RL, false, false, 2, 2, 1>(tw, iters);
}

template <int PI>
static void check_lazy(long iters)
{
    check_lazy_rl<PI, 1>(iters);
    check_lazy_rl<PI, 2>(iters);
    check_lazy_rl<PI, 3>(iters);
}

/* crt_below_p0 against the textbook Garner digits (128-bit arithmetic):
 * residues of X are those of an integer below p0 iff a1 = a2 = 0 in
 * X = x0 + a1 p0 + a2 p0 p1.  Integers below p0 (also in [p2, p0), where
 * x0 must be reduced before the comparison), integers from p0 up, and
 * random tuples. */
static void check_crt(long iters)
{
    const uint64_t p0 = prime(0), p1 = prime(1), p2 = prime(2);
    const uint64_t i01 = powmod(p0 % p1, p1 - 2, p1);
    const uint64_t i012 = (uint64_t)((u128)powmod(p0 % p2, p2 - 2, p2) * powmod(p1 % p2, p2 - 2, p2) % p2);
    auto digits_zero = [&](uint64_t x0, uint64_t x1, uint64_t x2) {
        const uint64_t a1 = (uint64_t)((u128)((x1 + p1 - x0 % p1) % p1) * i01 % p1);
        const uint64_t b = (uint64_t)(((u128)x0 + (u128)a1 * p0) % p2);
        const uint64_t a2 = (uint64_t)((u128)((x2 + p2 - b) % p2) * i012 % p2);
        return a1 == 0 && a2 == 0;
    };
    /* residues of X */
    auto res = [&](u128 X, uint64_t &x0, uint64_t &x1, uint64_t &x2) {
        x0 = (uint64_t)(X % p0);
        x1 = (uint64_t)(X % p1);
        x2 = (uint64_t)(X % p2);
    };
    /* a1 in {0, 1} and a2 = 0: the integer below 2 p0 */
    auto digits_small = [&](uint64_t x0, uint64_t x1, uint64_t x2) {
        const uint64_t a1 = (uint64_t)((u128)((x1 + p1 - x0 % p1) % p1) * i01 % p1);
        const uint64_t b = (uint64_t)(((u128)x0 + (u128)a1 * p0) % p2);
        const uint64_t a2 = (uint64_t)((u128)((x2 + p2 - b) % p2) * i012 % p2);
        return a1 <= 1 && a2 == 0;
    };
    auto one = [&](u128 X, bool below) {
        uint64_t x0, x1, x2;
        res(X, x0, x1, x2);
        const bool got = crt_below_p0(x0, x1, x2);
        expect(got == digits_zero(x0, x1, x2), "crt vs Garner digits", 0, (uint64_t)(X >> 64), (uint64_t)X);
        expect(got == below, "crt below p0", 0, (uint64_t)(X >> 64), (uint64_t)X);
        bool ok;
        const uint64_t v = crt_below_2p0(x0, x1, x2, ok);
        const bool below2 = X < 2 * (u128)p0;
        expect(ok == digits_small(x0, x1, x2), "crt2 vs Garner digits", 0, (uint64_t)(X >> 64), (uint64_t)X);
        expect(ok == below2 && (!ok || v == (uint64_t)X), "crt below 2 p0", 0, (uint64_t)(X >> 64), (uint64_t)X);
    };
    const u128 edges[] = {0, 1, p2 - 1, p2, p2 + 1, p1 - 1, p1, p1 + 1, p0 - 1, p0, p0 + 1,
                          p0 + p2, p0 + p1, 2 * (u128)p0 - 1, 2 * (u128)p0, 2 * (u128)p0 + 1, (u128)p0 * p1,
                          (u128)p0 * p1 - 1, (u128)p0 * p1 + p0 - 1, ~(u128)0};
    for (u128 X : edges) one(X, X < p0);
    for (long i = 0; i < iters; i++) {
        one(rnd() % p0, true);
        one(p2 + rnd() % (p0 - p2), true);
        one(p0 + rnd() % p0, false);
        one(p0 + (((u128)rnd() << 64) | rnd()) % ((~(u128)0) - p0), false);
        const uint64_t x0 = rnd() % p0, x1 = rnd() % p1, x2 = rnd() % p2;
        expect(crt_below_p0(x0, x1, x2) == digits_zero(x0, x1, x2), "crt random tuple", 0, x0, x1);
        bool ok;
        (void)crt_below_2p0(x0, x1, x2, ok);
        expect(ok == digits_small(x0, x1, x2), "crt2 random tuple", 0, x0, x1);
    }
}

int main(int argc, char **argv)
{
    long iters = argc > 1 ? atol(argv[1]) : 2000000;
    check<0>(iters);
    check<1>(iters);
    check<2>(iters);
    check_lazy<0>(iters / 20);
    check_lazy<2>(iters / 20);
    check_crt(iters / 8);
    printf("bad %d\n", bad);
    return bad != 0;
}
