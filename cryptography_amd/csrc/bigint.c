/*
 * bigint.c — BigInt helpers (host, O(n) bookkeeping only).
 *
 * Behaviour matches the reference's ntt/lib/arith/bigint.c (LIMB_BITS 64):
 * resize grows with zero fill and aborts on a managed BigInt (bigint.c:46-60),
 * add/sub/shr are alias-safe, results are trimmed.  Every multiply runs on the
 * GPU (multiply.c); nothing here multiplies big operands.
 */
#include "../../include/bigint.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* Managed allocation lives in the HIP part of the library (ntt_engine.hip:
 * hipMallocManaged, cf. ntt/lib/arith/bigint_hip_alloc.hip). */
extern limb_t *ntt_managed_alloc_limbs(int cap);
extern void ntt_managed_free_limbs(limb_t *p);

static void die(const char *msg)
{
    fprintf(stderr, "%s\n", msg);
    exit(1);
}

static void trim(BigInt *a)
{
    while (a->n_limbs > 0 && a->limbs[a->n_limbs - 1] == 0) a->n_limbs--;
}

BigInt bigint_alloc(int cap)
{
    BigInt a;
    if (cap < 1) cap = 1;
    a.limbs = (limb_t *)calloc((size_t)cap, sizeof(limb_t));
    if (!a.limbs) die("bigint_alloc: out of memory");
    a.n_limbs = 0;
    a.cap = cap;
    a.managed = 0;
    return a;
}

BigInt bigint_alloc_managed(int cap)
{
    BigInt a;
    if (cap < 1) cap = 1;
    a.limbs = ntt_managed_alloc_limbs(cap);
    if (!a.limbs) die("bigint_alloc_managed: hipMallocManaged failed");
    a.n_limbs = 0;
    a.cap = cap;
    a.managed = 1;
    return a;
}

void bigint_free(BigInt *a)
{
    if (a->limbs) {
        if (a->managed) ntt_managed_free_limbs(a->limbs);
        else free(a->limbs);
    }
    a->limbs = NULL;
    a->n_limbs = 0;
    a->cap = 0;
    a->managed = 0;
}

void bigint_resize(BigInt *a, int new_cap)
{
    if (new_cap <= a->cap) return;
    if (a->managed) {
        fprintf(stderr, "bigint_resize: managed BigInt cap=%d cannot grow to %d; pre-allocate it\n", a->cap,
                new_cap);
        abort();
    }
    limb_t *p = (limb_t *)realloc(a->limbs, (size_t)new_cap * sizeof(limb_t));
    if (!p) die("bigint_resize: out of memory");
    memset(p + a->cap, 0, (size_t)(new_cap - a->cap) * sizeof(limb_t));
    a->limbs = p;
    a->cap = new_cap;
}

void bigint_zero(BigInt *a)
{
    if (a->cap) memset(a->limbs, 0, (size_t)a->cap * sizeof(limb_t));
    a->n_limbs = 0;
}

void bigint_set_u64(BigInt *a, uint64_t v)
{
    bigint_resize(a, 1);
    bigint_zero(a);
    if (v) {
        a->limbs[0] = v;
        a->n_limbs = 1;
    }
}

void bigint_copy(BigInt *dst, const BigInt *src)
{
    if (dst == src) return;
    int n = src->n_limbs;
    bigint_resize(dst, n > 0 ? n : 1);
    memmove(dst->limbs, src->limbs, (size_t)n * sizeof(limb_t));
    memset(dst->limbs + n, 0, (size_t)(dst->cap - n) * sizeof(limb_t));
    dst->n_limbs = n;
}

int bigint_is_zero(const BigInt *a) { return a->n_limbs == 0; }

int bigint_cmp(const BigInt *a, const BigInt *b)
{
    if (a->n_limbs != b->n_limbs) return a->n_limbs > b->n_limbs ? 1 : -1;
    for (int i = a->n_limbs - 1; i >= 0; i--)
        if (a->limbs[i] != b->limbs[i]) return a->limbs[i] > b->limbs[i] ? 1 : -1;
    return 0;
}

void bigint_add(BigInt *c, const BigInt *a, const BigInt *b)
{
    int na = a->n_limbs, nb = b->n_limbs;
    int n = (na > nb ? na : nb) + 1;
    /* c may alias a or b: keep the limb pointers valid across the resize */
    int ca = (c == a), cb = (c == b);
    bigint_resize(c, n);
    const limb_t *pa = ca ? c->limbs : a->limbs, *pb = cb ? c->limbs : b->limbs;
    uint64_t carry = 0;
    for (int i = 0; i < n - 1; i++) {
        u128 s = (u128)(i < na ? pa[i] : 0) + (i < nb ? pb[i] : 0) + carry;
        c->limbs[i] = (limb_t)s;
        carry = (uint64_t)(s >> 64);
    }
    c->limbs[n - 1] = carry;
    for (int i = n; i < c->cap && i < c->n_limbs; i++) c->limbs[i] = 0;
    c->n_limbs = n;
    trim(c);
}

void bigint_sub(BigInt *c, const BigInt *a, const BigInt *b)
{
    int na = a->n_limbs, nb = b->n_limbs;
    int ca = (c == a), cb = (c == b);
    bigint_resize(c, na > 0 ? na : 1);
    const limb_t *pa = ca ? c->limbs : a->limbs, *pb = cb ? c->limbs : b->limbs;
    uint64_t borrow = 0;
    for (int i = 0; i < na; i++) {
        uint64_t x = pa[i], y = i < nb ? pb[i] : 0;
        uint64_t d = x - y - borrow;
        borrow = ((u128)y + borrow > x) ? 1 : 0;
        c->limbs[i] = d;
    }
    for (int i = na; i < c->cap && i < c->n_limbs; i++) c->limbs[i] = 0;
    c->n_limbs = na;
    trim(c);
}

void bigint_add_u64(BigInt *a, uint64_t v)
{
    bigint_resize(a, a->n_limbs + 2);
    uint64_t carry = v;
    for (int i = 0; carry; i++) {
        u128 s = (u128)a->limbs[i] + carry;
        a->limbs[i] = (limb_t)s;
        carry = (uint64_t)(s >> 64);
        if (i + 1 > a->n_limbs) a->n_limbs = i + 1;
    }
    trim(a);
}

void bigint_sub_u64(BigInt *a, uint64_t v)
{
    uint64_t borrow = v;
    for (int i = 0; i < a->n_limbs && borrow; i++) {
        uint64_t x = a->limbs[i];
        a->limbs[i] = x - borrow;
        borrow = x < borrow ? 1 : 0;
    }
    trim(a);
}

void bigint_shl(BigInt *c, const BigInt *a, int bits)
{
    int ls = bits / 64, bs = bits % 64;
    int na = a->n_limbs;
    int n = na + ls + 1;
    limb_t *tmp = (limb_t *)calloc((size_t)n, sizeof(limb_t));
    if (!tmp) die("bigint_shl: out of memory");
    for (int i = 0; i < na; i++) {
        tmp[i + ls] |= bs ? a->limbs[i] << bs : a->limbs[i];
        if (bs) tmp[i + ls + 1] |= a->limbs[i] >> (64 - bs);
    }
    bigint_resize(c, n);
    bigint_zero(c);
    memcpy(c->limbs, tmp, (size_t)n * sizeof(limb_t));
    free(tmp);
    c->n_limbs = n;
    trim(c);
}

void bigint_shr(BigInt *c, const BigInt *a, int bits)
{
    int ls = bits / 64, bs = bits % 64;
    int na = a->n_limbs;
    if (ls >= na) {
        bigint_resize(c, 1);
        bigint_zero(c);
        return;
    }
    int n = na - ls;
    int ca = (c == a);
    bigint_resize(c, n);
    const limb_t *pa = ca ? c->limbs : a->limbs;
    for (int i = 0; i < n; i++) {   /* reads lead writes: alias-safe */
        uint64_t lo = pa[i + ls] >> bs;
        uint64_t hi = (bs && i + ls + 1 < na) ? pa[i + ls + 1] << (64 - bs) : 0;
        c->limbs[i] = lo | hi;
    }
    for (int i = n; i < c->cap && i < na; i++) c->limbs[i] = 0;
    c->n_limbs = n;
    trim(c);
}

void bigint_mul_u64(BigInt *c, const BigInt *a, uint64_t v)
{
    int na = a->n_limbs;
    int ca = (c == a);
    bigint_resize(c, na + 1);
    const limb_t *pa = ca ? c->limbs : a->limbs;
    uint64_t carry = 0;
    for (int i = 0; i < na; i++) {
        u128 pr = (u128)pa[i] * v + carry;
        c->limbs[i] = (limb_t)pr;
        carry = (uint64_t)(pr >> 64);
    }
    c->limbs[na] = carry;
    for (int i = na + 1; i < c->cap && i < c->n_limbs; i++) c->limbs[i] = 0;
    c->n_limbs = na + 1;
    trim(c);
}

int bigint_bits(const BigInt *a)
{
    if (a->n_limbs == 0) return 0;
    uint64_t w = a->limbs[a->n_limbs - 1];
    return (a->n_limbs - 1) * 64 + (64 - __builtin_clzll(w));
}

void bigint_print_hex(const BigInt *a, const char *label)
{
    if (label) fprintf(stderr, "%s: ", label);
    if (a->n_limbs == 0) {
        fprintf(stderr, "0\n");
        return;
    }
    for (int i = a->n_limbs - 1; i >= 0; i--) fprintf(stderr, "%016llx", (unsigned long long)a->limbs[i]);
    fprintf(stderr, "\n");
}
