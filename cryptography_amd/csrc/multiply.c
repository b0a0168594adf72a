/*
 * multiply.c — the reference's public multiply API (ntt/lib/arith/multiply.c)
 * over the GPU engine.
 *
 * Dispatch: both operands at or below the threshold -> GPU schoolbook kernel,
 * otherwise the CRT-NTT (multiply.c:78-90 dispatches the same way, with the
 * schoolbook on the host).  Errors are fatal, as in the reference
 * (ntt_mul.hip:101-109, :997-999, :1013-1022).
 */
#include "../../include/multiply.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ntt_engine.h"

static void fatal(const char *what, int rc)
{
    fprintf(stderr, "%s: %s (status %d)\n", what, ntt_engine_last_error(), rc);
    exit(1);
}

static void ensure_init(const char *what)
{
    if (!ntt_engine_is_init()) {
        int rc = ntt_engine_init(0);
        if (rc) fatal(what, rc);
    }
}

int bigint_mul_get_threshold(void) { return ntt_engine_get_threshold(); }
void bigint_mul_set_threshold(int t) { ntt_engine_set_threshold(t); }

void ntt_mul_init_impl(int max_log_n)
{
    if (max_log_n < 3 || max_log_n > 30) {
        fprintf(stderr, "ntt_mul_init: max_log_n=%d outside [3, 30]\n", max_log_n);
        exit(1);
    }
    int rc = ntt_engine_init(0);
    if (rc) fatal("ntt_mul_init", rc);
}

void ntt_mul_teardown_impl(void) { ntt_engine_teardown(); }
void ntt_mul_init(int max_log_n) { ntt_mul_init_impl(max_log_n); }
void ntt_mul_teardown(void) { ntt_mul_teardown_impl(); }
unsigned long ntt_get_gpu_dispatch_count(void) { return ntt_engine_dispatch_count(); }

/* Product into a temporary, then into c: alias-safe and c's storage is only
 * resized after the operands have been read.  A managed c that is too short
 * aborts in bigint_resize, as in the reference (bigint.c:46-60). */
static void mul_common(BigInt *c, const BigInt *a, const BigInt *b, int square, int force_ntt, const char *what)
{
    ensure_init(what);
    size_t na = (size_t)a->n_limbs, nb = square ? na : (size_t)b->n_limbs;
    size_t n = na + nb;
    if (na == 0 || nb == 0) {
        bigint_resize(c, 1);
        bigint_zero(c);
        return;
    }
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    if (!tmp) {
        fprintf(stderr, "%s: out of memory\n", what);
        exit(1);
    }
    /* the forced NTT path is a flag of the call, not a change of the shared
     * threshold, so concurrent bigint_mul callers never see it */
    int rc = square      ? ntt_sqr_limbs(tmp, a->limbs, na)
             : force_ntt ? ntt_mul_limbs_ntt(tmp, a->limbs, na, b->limbs, nb)
                         : ntt_mul_limbs(tmp, a->limbs, na, b->limbs, nb);
    if (rc) {
        free(tmp);
        fatal(what, rc);
    }
    if ((size_t)c->cap < n) bigint_resize(c, (int)n);
    memcpy(c->limbs, tmp, n * sizeof(uint64_t));
    if ((size_t)c->cap > n) memset(c->limbs + n, 0, ((size_t)c->cap - n) * sizeof(uint64_t));
    free(tmp);
    c->n_limbs = (int)n;
    while (c->n_limbs > 0 && c->limbs[c->n_limbs - 1] == 0) c->n_limbs--;
}

void bigint_mul(BigInt *c, const BigInt *a, const BigInt *b) { mul_common(c, a, b, 0, 0, "bigint_mul"); }
void bigint_sqr(BigInt *c, const BigInt *a) { mul_common(c, a, a, 1, 0, "bigint_sqr"); }
void ntt_bigint_mul(BigInt *c, const BigInt *a, const BigInt *b) { mul_common(c, a, b, 0, 1, "ntt_bigint_mul"); }
