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
