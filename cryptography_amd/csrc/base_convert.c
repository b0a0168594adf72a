/*
 * base_convert.c — decimal conversion for include/base_convert.h (the
 * reference's ntt/lib/arith/base_convert.h), every product through bigint_mul
 * on the GPU.
 *
 * Divide and conquer on the digit count: a value below 10^(2^(k+1)) splits at
 * 10^(2^k) into a quotient (the high digits) and a remainder (exactly 2^k low
 * digits, zero-padded), by bigint_div_newton with the cached reciprocal of
 * 10^(2^k).  Pieces below 10^2048 are converted on the host by repeated
 * division by 10^19 (the splits there would be GPU launches of a few limbs
 * each); values below 10^16 are printed directly.  The reference's
 * algorithm is the same divide and conquer (base_convert.c:135-259).
 * Above DECIMAL_DEVICE_LIMBS limbs the whole conversion runs on the device
 * instead (ntt_to_decimal: a fixed-point radix tree with one batched
 * multiply per level), with the same result.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../../include/base_convert.h"
#include "../../include/multiply.h"
#include "../../include/newton.h"
#include "../../include/ntt_engine.h"

#define MAX_LEVEL 40
#define DIRECT_LEVEL 3        /* below 10^(2^(3+1)) = 10^16: one u64 */
#define BASE_LEVEL 10         /* below 10^(2^11): converted on the host by repeated division by 10^19 */
#define DECIMAL_DEVICE_LIMBS 256   /* above this (about 4900 digits): ntt_to_decimal, all on the device */

static BigInt pow10_k[MAX_LEVEL + 1], recip_k[MAX_LEVEL + 1];
static int cache_level = -1;

static void die(const char *what)
{
    fprintf(stderr, "base_convert: %s\n", what);
    exit(1);
}

void base_convert_teardown(void)
{
    for (int k = 0; k <= cache_level; k++) {
        bigint_free(&pow10_k[k]);
        bigint_free(&recip_k[k]);
    }
    cache_level = -1;
}

void base_convert_init(int max_level)
{
    if (max_level < 0 || max_level > MAX_LEVEL) die("base_convert_init: max_level out of range [0, 40]");
    if (max_level <= cache_level) return;
    for (int k = cache_level + 1; k <= max_level; k++) {
        pow10_k[k] = bigint_alloc(4);
        recip_k[k] = bigint_alloc(4);
        if (k == 0) bigint_set_u64(&pow10_k[0], 10);
        else bigint_sqr(&pow10_k[k], &pow10_k[k - 1]);
        newton_reciprocal(&recip_k[k], &pow10_k[k]);
        cache_level = k;
    }
}

static void print_u64(char *out, uint64_t v, int width)   /* width digits, zero-padded */
{
    for (int i = width - 1; i >= 0; i--) {
        out[i] = (char)('0' + v % 10);
        v /= 10;
    }
}

static int u64_digits(uint64_t v)
{
    int n = 1;
    while (v >= 10) {
        v /= 10;
        n++;
    }
    return n;
}

/* Exactly `width` digits of N (N < 10^width), zero-padded, by repeated
 * division of the limbs by 10^19: the base case below 10^2048, where the
 * quadratic loop is a few microseconds and the GPU splits would be dominated
 * by their launches. */
static void emit_small(const BigInt *N, size_t width, char *out)
{
    int n = N->n_limbs;
    uint64_t *t = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
    if (!t) die("out of memory");
    memcpy(t, N->limbs, (size_t)n * sizeof(uint64_t));
    const uint64_t D = 10000000000000000000ull;   /* 10^19 */
    size_t pos = width;
    while (pos > 0) {
        unsigned __int128 r = 0;
        for (int i = n - 1; i >= 0; i--) {
            unsigned __int128 cur = (r << 64) | t[i];
            t[i] = (uint64_t)(cur / D);
            r = cur % D;
        }
        while (n > 0 && t[n - 1] == 0) n--;
        uint64_t d = (uint64_t)r;
        for (int k = 0; k < 19 && pos > 0; k++) {
            out[--pos] = (char)('0' + d % 10);
            d /= 10;
        }
    }
    free(t);
}

/* N < 10^(2^(k+1)): q = N / 10^(2^k), r = N mod 10^(2^k). */
static void split(const BigInt *N, int k, BigInt *q, BigInt *r)
{
    BigInt t = bigint_alloc(4);
    bigint_div_newton(q, N, &pow10_k[k], &recip_k[k]);
    bigint_mul(&t, q, &pow10_k[k]);
    bigint_sub(r, N, &t);
    bigint_free(&t);
}

/* Exactly 2^(k+1) digits of N < 10^(2^(k+1)), zero-padded. */
static void emit_padded(const BigInt *N, int k, char *out)
{
    if (k < DIRECT_LEVEL + 1) {
        print_u64(out, N->n_limbs ? N->limbs[0] : 0, 1 << (k + 1));
        return;
    }
    if (k <= BASE_LEVEL) {
        emit_small(N, (size_t)1 << (k + 1), out);
        return;
    }
    BigInt q = bigint_alloc(4), r = bigint_alloc(4);
    split(N, k, &q, &r);
    emit_padded(&q, k - 1, out);
    emit_padded(&r, k - 1, out + ((size_t)1 << k));
    bigint_free(&q);
    bigint_free(&r);
}

/* Digits of N < 10^(2^(k+1)) without leading zeros (N > 0); returns the count. */
static size_t emit_top(const BigInt *N, int k, char *out)
{
    if (k < DIRECT_LEVEL + 1) {
        uint64_t v = N->limbs[0];
        int n = u64_digits(v);
        print_u64(out, v, n);
        return (size_t)n;
    }
    if (k <= BASE_LEVEL) {
        /* N < 10^(2^(k+1)): all digits, then drop the leading zeros */
        size_t w = (size_t)1 << (k + 1), z = 0;
        emit_small(N, w, out);
        while (z + 1 < w && out[z] == '0') z++;
        memmove(out, out + z, w - z);
        return w - z;
    }
    if (bigint_cmp(N, &pow10_k[k]) < 0) return emit_top(N, k - 1, out);
    BigInt q = bigint_alloc(4), r = bigint_alloc(4);
    split(N, k, &q, &r);
    size_t n = emit_top(&q, k - 1, out);
    emit_padded(&r, k - 1, out + n);
    bigint_free(&q);
    bigint_free(&r);
    return n + ((size_t)1 << k);
}

char *bigint_to_decimal(const BigInt *N)
{
    if (bigint_is_zero(N)) {
        char *s = (char *)malloc(2);
        if (!s) die("out of memory");
        s[0] = '0';
        s[1] = 0;
        return s;
    }
    if (cache_level < 0) return NULL;
    /* smallest K with N < 10^(2^(K+1)): pow10_k[K+1], or past the cache the
     * square of its top entry */
    int K = -2;
    for (int k = 0; k <= cache_level; k++)
        if (bigint_cmp(N, &pow10_k[k]) < 0) {
            K = k - 1;
            break;
        }
    if (K == -2) {
        BigInt sq = bigint_alloc(4);
        bigint_sqr(&sq, &pow10_k[cache_level]);
        int fits = bigint_cmp(N, &sq) < 0;
        bigint_free(&sq);
        if (!fits) return NULL;
        K = cache_level;
    }
    if (N->n_limbs > DECIMAL_DEVICE_LIMBS) {
        /* one device-resident conversion instead of a division per node */
        if (!ntt_engine_is_init() && ntt_engine_init(0)) die("engine initialisation failed");
        size_t cap = (size_t)((double)bigint_bits(N) * 0.30103) + 2, len = 0;
        char *s = (char *)malloc(cap);
        if (!s) die("out of memory");
        int rc = ntt_to_decimal(N->limbs, (size_t)N->n_limbs, s, cap, &len);
        if (rc) {
            fprintf(stderr, "base_convert: %s\n", ntt_engine_last_error());
            die("device decimal conversion failed");
        }
        return s;
    }
    if (K < DIRECT_LEVEL) K = DIRECT_LEVEL;   /* the direct case covers everything below 10^16 */
    size_t cap = ((size_t)1 << (K + 1)) + 1;
    char *s = (char *)malloc(cap);
    if (!s) die("out of memory");
    size_t n = emit_top(N, K, s);
    s[n] = 0;
    return s;
}

long bigint_write_decimal(const BigInt *N, int fd)
{
    char *s = bigint_to_decimal(N);
    if (!s) return -1;
    size_t n = strlen(s), done = 0;
    while (done < n) {
        ssize_t w = write(fd, s + done, n - done);
        if (w <= 0) {
            free(s);
            return -1;
        }
        done += (size_t)w;
    }
    free(s);
    return (long)n;
}

void base_convert_pow10_exact(BigInt *out, long n)
{
    if (n < 0) die("base_convert_pow10_exact: negative exponent");
    bigint_set_u64(out, 1);
    for (int k = 0; n >> k; k++) {
        if (!((n >> k) & 1)) continue;
        if (k > cache_level) die("base_convert_pow10_exact: cache too small (call base_convert_init)");
        bigint_mul(out, out, &pow10_k[k]);
    }
}
