/*
 * Host check of the modular primitives in cryptography_amd/csrc/modarith.h
 * against 128-bit integer arithmetic, for the three NTT primes: the value and
 * the documented output range of the Montgomery product, the Shoup product
 * (and its small-operand form), both butterflies (the inverse one also with a
 * negated twiddle), and the compile-time roots.  Prints "bad N" and
 * exits non-zero on any mismatch.
 */
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
        expect(r > 0 && r < 3 * p, "mont range (a<4p, b<2p)", PI, a, b);
        if (a < 2 * p && b < 2 * p) expect(r < 2 * p, "mont range (a,b<2p)", PI, a, b);
        if (b < p) expect(r < 2 * p, "mont range (a<4p, b<p)", PI, a, b);
    };
    auto shoup = [&](uint64_t a, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        const uint64_t r = shoup_lazy<PI>(a, w, wp);
        expect(r < 2 * p && r % p == (uint64_t)((u128)a * w % p), "shoup", PI, a, w);
    };
    /* small-operand Shoup: the same quotient as the full one, so the same value */
    auto shoup_s = [&](uint32_t a, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        const uint64_t r = shoup_small<PI>(a, w, wp);
        expect(r < 2 * p && r % p == (uint64_t)((u128)a * w % p), "shoup small", PI, a, w);
        expect(r == shoup_lazy<PI>(a, w, wp), "shoup small quotient", PI, a, w);
    };
    auto bfly = [&](uint64_t x, uint64_t y, uint64_t w) {
        const uint64_t wp = (uint64_t)(((u128)w << 64) / p);
        uint64_t a = x, b = y;
        dif_bfly<PI>(a, b, w, wp);            /* inputs [0, 2p) */
        expect(a < 2 * p && a % p == (x + y) % p, "dif sum", PI, x, y);
        expect(b < 2 * p && b % p == (uint64_t)((u128)((x + 2 * p - y) % p) * w % p), "dif diff", PI, x, y);
        const uint64_t X = x * 2 % (4 * p), Y = y * 2 % (4 * p);   /* inverse inputs range over [0, 4p) */
        a = X;
        b = Y;
        dit_bfly<PI>(a, b, w, wp);
        const uint64_t t = (uint64_t)((u128)(Y % p) * w % p);
        expect(a < 4 * p && a % p == (X % p + t) % p, "dit sum", PI, X, Y);
        expect(b < 4 * p && b % p == (X % p + p - t) % p, "dit diff", PI, X, Y);
        /* negated-twiddle form: W = -w gives the butterfly of twiddle w */
        const uint64_t W = (p - w) % p, Wp = (uint64_t)(((u128)W << 64) / p);
        a = X;
        b = Y;
        dit_bfly_neg<PI>(a, b, W, Wp);
        expect(a < 4 * p && a % p == (X % p + t) % p, "dit neg sum", PI, X, Y);
        expect(b < 4 * p && b % p == (X % p + p - t) % p, "dit neg diff", PI, X, Y);
    };
    for (uint64_t a : edges)
        for (uint64_t b : edges) {
            if (a < 4 * p && b < 2 * p) mont(a, b);
            if (b < p) shoup(a, b);
            if (b < p && a <= 0xffffffffull) shoup_s((uint32_t)a, b);
            if (a < 2 * p && b < 2 * p) bfly(a, b, (b % p) ? b % p : 1);
        }
    for (long i = 0; i < iters; i++) {
        mont(rnd() % (4 * p), rnd() % (2 * p));
        mont(rnd() % (2 * p), rnd() % (2 * p));
        mont(rnd() % (4 * p), rnd() % p);
        shoup(rnd(), rnd() % p);
        shoup_s((uint32_t)rnd(), rnd() % p);
        shoup_s((uint32_t)(rnd() & 0xffff), rnd() % p);
        bfly(rnd() % (2 * p), rnd() % (2 * p), rnd() % p);
    }
    /* compile-time roots: w of order exactly 2^LOGORD, and w^E */
    using T = ConstTw<PI, 12, 1>;
    expect(powmod(T::w, 1u << 12, p) == 1 && powmod(T::w, 1u << 11, p) == p - 1, "root order", PI, T::w, 0);
    using T5 = ConstTw<PI, 12, 5>;
    expect(T5::w == powmod(T::w, 5, p) && T5::wp == (uint64_t)(((u128)T5::w << 64) / p), "root power", PI, T5::w, 0);
}

int main(int argc, char **argv)
{
    long iters = argc > 1 ? atol(argv[1]) : 2000000;
    check<0>(iters);
    check<1>(iters);
    check<2>(iters);
    printf("bad %d\n", bad);
    return bad != 0;
}
