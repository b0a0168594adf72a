/*
 * bigint_caller.c — a C caller of the drop-in multiply API, compiled against
 * include/ only and linked to libntt_amd.so (and GMP as the checker).
 *
 * It makes the calls a caller of the reference's ntt/lib/arith/multiply.h
 * makes, with BigInt structs:
 *
 *   bigint_caller gmp <log_n> [seed]
 *       The reference's own differential check
 *       (ntt/app/compute_e/test_gmp_oracle.c:55-76): two xorshift64 operands
 *       of N/2 - 16 limbs (N = 2^log_n, top bit set), ntt_bigint_mul, every
 *       limb against mpz_mul; then bigint_mul and bigint_sqr of the same
 *       operands.  Seed 0x123456789abcdef by default, as there.
 *   bigint_caller api
 *       init / teardown / re-init, the threshold get/set round trip, the
 *       dispatch count, mul_limb, c aliasing a or b, a c shorter than the
 *       product (resized), zero operands, products around the schoolbook
 *       crossover on both paths — each against mpz_mul.
 *
 * Exit status 0 when every check passes; one line per check on stdout.
 */
#include <gmp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bigint.h"
#include "multiply.h"

static int n_fail = 0;

static void report(const char *what, int ok)
{
    printf("%s %s\n", ok ? "PASS" : "FAIL", what);
    fflush(stdout);
    if (!ok) n_fail++;
}

/* Marsaglia xorshift64 (shifts 13, 7, 17), the generator of the reference's
 * GMP oracle test. */
static uint64_t xs_state;
static uint64_t xs_next(void)
{
    uint64_t x = xs_state;
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return xs_state = x;
}

/* BigInt of n random limbs with the top bit set. */
static BigInt random_bigint(int n)
{
    BigInt a = bigint_alloc(n + 8);
    for (int i = 0; i < n; i++) a.limbs[i] = xs_next();
    if (n) a.limbs[n - 1] |= 1ull << 63;
    a.n_limbs = n;
    return a;
}

/* Whether c (trimmed BigInt) equals x*y computed by GMP, limb for limb. */
static int equals_gmp_product(const BigInt *c, const BigInt *x, const BigInt *y)
{
    mpz_t gx, gy, gz;
    mpz_inits(gx, gy, gz, NULL);
    mpz_import(gx, (size_t)x->n_limbs, -1, 8, 0, 0, x->limbs);
    mpz_import(gy, (size_t)y->n_limbs, -1, 8, 0, 0, y->limbs);
    mpz_mul(gz, gx, gy);
    size_t nz = mpz_sgn(gz) ? (mpz_sizeinbase(gz, 2) + 63) / 64 : 0;
    int ok = (size_t)c->n_limbs == nz;
    if (ok && nz) {
        uint64_t *ref = (uint64_t *)calloc(nz, 8);
        size_t cnt = 0;
        mpz_export(ref, &cnt, -1, 8, 0, 0, gz);
        ok = cnt == nz && memcmp(ref, c->limbs, nz * 8) == 0;
        if (!ok) {
            size_t i = 0;
            while (i < nz && ref[i] == c->limbs[i]) i++;
            printf("     first differing limb %zu of %zu\n", i, nz);
        }
        free(ref);
    }
    mpz_clears(gx, gy, gz, NULL);
    return ok;
}

static int run_gmp(int log_n, uint64_t seed)
{
    const int N = 1 << log_n;
    const int n = N / 2 - 16;
    xs_state = seed ? seed : 1;
    ntt_mul_init(log_n);
    BigInt a = random_bigint(n), b = random_bigint(n);
    BigInt c = bigint_alloc(2 * n + 16);
    char what[160];
    unsigned long d0 = ntt_get_gpu_dispatch_count();
    ntt_bigint_mul(&c, &a, &b);
    snprintf(what, sizeof what, "ntt_bigint_mul log_n=%d operand_limbs=%d (every limb vs mpz_mul)", log_n, n);
    report(what, equals_gmp_product(&c, &a, &b));
    report("ntt_get_gpu_dispatch_count advanced by one", ntt_get_gpu_dispatch_count() == d0 + 1);
    BigInt c2 = bigint_alloc(1);
    bigint_mul(&c2, &a, &b);
    report("bigint_mul into a one-limb c (resized) equals ntt_bigint_mul",
           c2.n_limbs == c.n_limbs && memcmp(c2.limbs, c.limbs, (size_t)c.n_limbs * 8) == 0);
    bigint_sqr(&c2, &a);
    snprintf(what, sizeof what, "bigint_sqr log_n=%d (every limb vs mpz_mul)", log_n);
    report(what, equals_gmp_product(&c2, &a, &a));
    bigint_free(&a);
    bigint_free(&b);
    bigint_free(&c);
    bigint_free(&c2);
    ntt_mul_teardown();
    return n_fail ? 1 : 0;
}

static int run_api(void)
{
    /* init, teardown, init again */
    ntt_mul_init(20);
    ntt_mul_teardown();
    ntt_mul_init_impl(24);

    /* threshold: default, set, restore */
    const int t0 = bigint_mul_get_threshold();
    report("default threshold is BIGINT_MUL_THRESHOLD", t0 == BIGINT_MUL_THRESHOLD);
    bigint_mul_set_threshold(7);
    report("bigint_mul_set_threshold(7) reads back", bigint_mul_get_threshold() == 7);
    bigint_mul_set_threshold(t0);

    /* mul_limb against the 128-bit product */
    xs_state = 0x243F6A8885A308D3ull;
    int ok = 1;
    for (int i = 0; i < 100000; i++) {
        const uint64_t x = xs_next(), y = i < 4 ? ~0ull : xs_next();
        limb_t lo, hi;
        mul_limb(x, y, &lo, &hi);
        const unsigned __int128 p = (unsigned __int128)x * y;
        ok &= lo == (uint64_t)p && hi == (uint64_t)(p >> 64);
    }
    report("mul_limb equals the 128-bit product (10^5 pairs)", ok);

    /* products around the crossover, on the threshold's path and forced NTT */
    static const int sizes[][2] = {{1, 1}, {2, 1}, {63, 64}, {64, 64}, {65, 64}, {64, 65}, {200, 3},
                                   {1000, 999}, {4096, 4096}, {30000, 17}};
    xs_state = 0x123456789abcdefull;
    for (size_t k = 0; k < sizeof sizes / sizeof sizes[0]; k++) {
        BigInt a = random_bigint(sizes[k][0]), b = random_bigint(sizes[k][1]);
        BigInt c = bigint_alloc(1);
        char what[128];
        unsigned long d0 = ntt_get_gpu_dispatch_count();
        bigint_mul(&c, &a, &b);
        snprintf(what, sizeof what, "bigint_mul %d x %d limbs", sizes[k][0], sizes[k][1]);
        report(what, equals_gmp_product(&c, &a, &b));
        ntt_bigint_mul(&c, &a, &b);
        snprintf(what, sizeof what, "ntt_bigint_mul %d x %d limbs", sizes[k][0], sizes[k][1]);
        report(what, equals_gmp_product(&c, &a, &b));
        report("  threshold unchanged by ntt_bigint_mul", bigint_mul_get_threshold() == t0);
        report("  two dispatches counted", ntt_get_gpu_dispatch_count() == d0 + 2);
        bigint_free(&a);
        bigint_free(&b);
        bigint_free(&c);
    }

    /* aliasing: c == a, c == b, c == a == b; each against the unaliased product */
    {
        BigInt a = random_bigint(3000), b = random_bigint(2500);
        BigInt ac = bigint_alloc(1), bc = bigint_alloc(1), want = bigint_alloc(1);
        bigint_copy(&ac, &a);
        bigint_copy(&bc, &b);
        bigint_mul(&want, &a, &b);
        bigint_mul(&ac, &ac, &b);
        report("bigint_mul(c, c, b) with c == a", bigint_cmp(&ac, &want) == 0 && equals_gmp_product(&ac, &a, &b));
        bigint_mul(&bc, &a, &bc);
        report("bigint_mul(c, a, c) with c == b", bigint_cmp(&bc, &want) == 0);
        bigint_copy(&ac, &a);
        bigint_mul(&ac, &ac, &ac);
        report("bigint_mul(c, c, c)", equals_gmp_product(&ac, &a, &a));
        bigint_copy(&ac, &a);
        bigint_sqr(&ac, &ac);
        report("bigint_sqr(c, c)", equals_gmp_product(&ac, &a, &a));
        bigint_copy(&ac, &a);
        ntt_bigint_mul(&ac, &ac, &b);
        report("ntt_bigint_mul(c, c, b) with c == a", bigint_cmp(&ac, &want) == 0);
        bigint_free(&a);
        bigint_free(&b);
        bigint_free(&ac);
        bigint_free(&bc);
        bigint_free(&want);
    }

    /* zero operands: the product is zero, trimmed to no limbs */
    {
        BigInt z = bigint_alloc(4), a = random_bigint(500), c = random_bigint(10);
        bigint_mul(&c, &z, &a);
        report("bigint_mul(0, a) is zero", bigint_is_zero(&c));
        c = (bigint_free(&c), random_bigint(10));
        bigint_mul(&c, &a, &z);
        report("bigint_mul(a, 0) is zero", bigint_is_zero(&c));
        c = (bigint_free(&c), random_bigint(10));
        ntt_bigint_mul(&c, &z, &z);
        report("ntt_bigint_mul(0, 0) is zero", bigint_is_zero(&c));
        bigint_free(&z);
        bigint_free(&a);
        bigint_free(&c);
    }

    /* a c larger than the product: limbs past the product are cleared */
    {
        BigInt a = random_bigint(100), b = random_bigint(100), c = bigint_alloc(400);
        for (int i = 0; i < 400; i++) c.limbs[i] = ~0ull;
        c.n_limbs = 400;
        bigint_mul(&c, &a, &b);
        int clean = 1;
        for (int i = c.n_limbs; i < c.cap; i++) clean &= c.limbs[i] == 0;
        report("bigint_mul into a longer c: product exact, high limbs cleared",
               equals_gmp_product(&c, &a, &b) && clean && c.cap == 400);
        bigint_free(&a);
        bigint_free(&b);
        bigint_free(&c);
    }

    /* a managed c pre-sized to the product (it is never reallocated) */
    {
        BigInt a = random_bigint(2000), b = random_bigint(1000), c = bigint_alloc_managed(3000);
        bigint_mul(&c, &a, &b);
        report("bigint_mul into a pre-sized managed c", c.managed == 1 && equals_gmp_product(&c, &a, &b));
        bigint_free(&a);
        bigint_free(&b);
        bigint_free(&c);
    }

    ntt_mul_teardown_impl();
    return n_fail ? 1 : 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && strcmp(argv[1], "gmp") == 0)
        return run_gmp(atoi(argv[2]), argc > 3 ? strtoull(argv[3], NULL, 0) : 0x123456789abcdefull);
    if (argc >= 2 && strcmp(argv[1], "api") == 0) return run_api();
    fprintf(stderr, "usage: %s gmp <log_n> [seed] | api\n", argv[0]);
    return 2;
}
