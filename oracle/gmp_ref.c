/*
 * gmp_ref.c — GMP-backed checks (TEST INFRASTRUCTURE ONLY).
 *
 * GMP (6.2.1 in this image, /opt/conda) is the canonical-integer oracle the
 * reference's own tests use (ntt/app/compute_e/test_gmp_oracle.c,
 * ntt/lib/arith/test_e2e_oracle.c).  Products here are mpz_mul; e digits are
 * the reference's formula (ntt/app/compute_e/main.c:167-169,267-345:
 * n_total = D + 20 guard digits, e = (B+P)*10^n_total / B truncated, D
 * significant digits) evaluated with mpz binary splitting.
 */
#include <gmp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "oracle.h"

static void imp(mpz_t z, const uint64_t *d, size_t n) { mpz_import(z, n, -1, 8, 0, 0, d); }

static size_t exp_limbs(const mpz_t z, uint64_t *out, size_t cap)
{
    size_t cnt = 0;
    size_t need = (mpz_sizeinbase(z, 2) + 63) / 64;
    if (mpz_sgn(z) == 0) need = 0;
    if (need > cap) return (size_t)-1;
    memset(out, 0, cap * sizeof(uint64_t));
    if (need) mpz_export(out, &cnt, -1, 8, 0, 0, z);
    return cnt;
}

/* out (na+nb limbs) = a*b via mpz_mul. Returns the trimmed limb count. */
size_t gmpref_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out)
{
    mpz_t x, y, z;
    mpz_inits(x, y, z, NULL);
    imp(x, a, na); imp(y, b, nb);
    mpz_mul(z, x, y);
    size_t r = exp_limbs(z, out, na + nb);
    mpz_clears(x, y, z, NULL);
    return r;
}

/* Time n_rep mpz_mul calls of a and b; returns seconds per multiply. */
double gmpref_time_mul(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int n_rep)
{
    mpz_t x, y, z;
    mpz_inits(x, y, z, NULL);
    imp(x, a, na); imp(y, b, nb);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < n_rep; i++) mpz_mul(z, x, y);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    mpz_clears(x, y, z, NULL);
    return ((t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec)) / n_rep;
}

static void bs(long a, long b, mpz_t P, mpz_t B)
{
    if (b - a == 1) { mpz_set_ui(P, 1); mpz_set_ui(B, (unsigned long)a); return; }
    long m = a + (b - a) / 2;
    mpz_t PL, BL, PR, BR;
    mpz_inits(PL, BL, PR, BR, NULL);
    bs(a, m, PL, BL);
    bs(m, b, PR, BR);
    mpz_mul(B, BL, BR);
    mpz_mul(P, PL, BR);
    mpz_add(P, P, PR);
    mpz_clears(PL, BL, PR, BR, NULL);
}

/* P, B for [a,b) as limb arrays (caller frees with free()). */
void gmpref_binary_split(long a, long b, uint64_t **P, size_t *np, uint64_t **B, size_t *nb)
{
    mpz_t p, q;
    mpz_inits(p, q, NULL);
    if (b <= a) { mpz_set_ui(p, 0); mpz_set_ui(q, 1); }
    else bs(a, b, p, q);
    size_t lp = (mpz_sizeinbase(p, 2) + 63) / 64 + 1, lq = (mpz_sizeinbase(q, 2) + 63) / 64 + 1;
    *P = (uint64_t *)calloc(lp, 8);
    *B = (uint64_t *)calloc(lq, 8);
    *np = mpz_sgn(p) ? exp_limbs(p, *P, lp) : 0;
    *nb = exp_limbs(q, *B, lq);
    mpz_clears(p, q, NULL);
}

/* First n_digits significant digits of e (no decimal point) into out
 * (n_digits+1 bytes). Returns 0 on success. */
int gmpref_e_digits(long n_digits, char *out)
{
    long n_total = n_digits + 20;
    long n_terms = orc_split_terms_needed(n_total);
    mpz_t P, B, num, t;
    mpz_inits(P, B, num, t, NULL);
    bs(1, n_terms + 1, P, B);
    mpz_add(num, B, P);
    mpz_ui_pow_ui(t, 10, (unsigned long)n_total);
    mpz_mul(num, num, t);
    mpz_fdiv_q(num, num, B);
    char *s = mpz_get_str(NULL, 10, num);
    int rc = 0;
    if ((long)strlen(s) < n_digits) rc = -1;
    else { memcpy(out, s, (size_t)n_digits); out[n_digits] = 0; }
    void (*freefunc)(void *, size_t);
    mp_get_memory_functions(NULL, NULL, &freefunc);
    freefunc(s, strlen(s) + 1);
    mpz_clears(P, B, num, t, NULL);
    return rc;
}

/* Decimal digits of the n-limb integer d (mpz_get_str), NUL-terminated, into
 * out (cap bytes); returns the digit count, or -1 when cap is too small. */
long gmpref_to_decimal(const uint64_t *d, size_t n, char *out, size_t cap)
{
    mpz_t z;
    mpz_init(z);
    imp(z, d, n);
    char *s = mpz_get_str(NULL, 10, z);
    size_t len = strlen(s);
    long rc = -1;
    if (len + 1 <= cap) {
        memcpy(out, s, len + 1);
        rc = (long)len;
    }
    void (*freefunc)(void *, size_t);
    mp_get_memory_functions(NULL, NULL, &freefunc);
    freefunc(s, len + 1);
    mpz_clear(z);
    return rc;
}
