/*
 * multiply_cpu.c — TEST HARNESS ONLY.  A CPU bigint_mul / bigint_sqr over the
 * oracle's 3-prime NTT multiply, so the host-side logic of newton.c and
 * base_convert.c (which only multiplies through bigint_mul) can be checked on
 * a machine without a GPU.  Never linked into libntt_amd.so.
 */
#include <stdlib.h>
#include <string.h>

#include "bigint.h"
#include "oracle.h"

limb_t *ntt_managed_alloc_limbs(int cap) { return (limb_t *)calloc((size_t)cap, sizeof(limb_t)); }
void ntt_managed_free_limbs(limb_t *p) { free(p); }

void bigint_mul(BigInt *c, const BigInt *a, const BigInt *b)
{
    int na = a->n_limbs, nb = b->n_limbs;
    if (!na || !nb) {
        bigint_resize(c, 1);
        bigint_zero(c);
        return;
    }
    uint64_t *t = (uint64_t *)calloc((size_t)na + nb, 8);
    if (na + nb <= 64) orc_mul_schoolbook(a->limbs, na, b->limbs, nb, t);
    else orc_mul_ntt3(a->limbs, na, b->limbs, nb, t, 1);
    bigint_resize(c, na + nb);
    bigint_zero(c);
    memcpy(c->limbs, t, ((size_t)na + nb) * 8);
    c->n_limbs = na + nb;
    while (c->n_limbs > 0 && c->limbs[c->n_limbs - 1] == 0) c->n_limbs--;
    free(t);
}

void bigint_sqr(BigInt *c, const BigInt *a) { bigint_mul(c, a, a); }

/* Engine entry points base_convert.c names for values above its device
 * threshold (256 limbs).  The CPU checks stay below it; reaching these is a
 * harness error. */
#include <stdio.h>
int ntt_engine_is_init(void) { return 1; }
int ntt_engine_init(int device) { (void)device; return 0; }
const char *ntt_engine_last_error(void) { return "no device in the CPU harness"; }
int ntt_to_decimal(const uint64_t *limbs, size_t n, char *out, size_t cap, size_t *len)
{
    (void)limbs; (void)n; (void)out; (void)cap; (void)len;
    fprintf(stderr, "multiply_cpu: ntt_to_decimal called in the CPU harness\n");
    return 1;
}
