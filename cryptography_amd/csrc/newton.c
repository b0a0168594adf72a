/*
 * newton.c — reciprocal and division for include/newton.h (the reference's
 * ntt/lib/arith/newton.h:17-30), with every product through bigint_mul, i.e.
 * on the GPU.
 *
 * Reciprocal by precision doubling on truncated operands: for the top k bits
 * Q_k = floor(Q / 2^(b-k)) of Q (b = bits(Q)), X_k = floor(2^(2k) / Q_k).
 *   - k <= 62: one 128-bit division.
 *   - otherwise, from X_h with h = floor(k/2) + 2: Y = X_h 2^(k-h) carries
 *     about h - 2 correct bits; one Newton update Y + Y (2^(2k) - Q_k Y) / 2^(2k)
 *     squares the error to a few units; the remainder 2^(2k) - Q_k X_k then
 *     moves X_k to the exact floor.
 * At k = b this is floor(2^(2b) / Q).  The reference instead seeds a full-size
 * iterate from the top 64 bits of Q (newton.c:55-130); both satisfy its
 * contract (the floor or one less), this one always returns the floor.
 *
 * Division: q0 = floor(N X / 2^(2b)) is the quotient or up to 3 below when
 * bits(N) <= 2b; a longer N uses floor(2^n / D) (n = bits(N)), which is the
 * reciprocal of D 2^(n-2b).  The remainder N - q0 D corrects q0.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../include/multiply.h"
#include "../../include/newton.h"

typedef unsigned __int128 u128;

#define MAX_CORRECTIONS 64   /* far above the analysis (a few); more means a corrupted product */

static void die(const char *what)
{
    fprintf(stderr, "newton: %s\n", what);
    exit(1);
}

static void set_pow2(BigInt *x, int e)
{
    int L = e / 64;
    bigint_resize(x, L + 1);
    bigint_zero(x);
    x->limbs[L] = (limb_t)1 << (e % 64);
    x->n_limbs = L + 1;
}

static void set_u128(BigInt *x, u128 v)
{
    bigint_resize(x, 2);
    bigint_zero(x);
    x->limbs[0] = (limb_t)v;
    x->limbs[1] = (limb_t)(v >> 64);
    x->n_limbs = x->limbs[1] ? 2 : (x->limbs[0] ? 1 : 0);
}

/* Make X the exact floor(T / Q), given X within MAX_CORRECTIONS of it. */
static void correct_quotient(BigInt *X, const BigInt *T, const BigInt *Q)
{
    BigInt P = bigint_alloc(4), r = bigint_alloc(4);
    bigint_mul(&P, Q, X);
    int steps = 0;
    while (bigint_cmp(&P, T) > 0) {           /* X too large */
        if (++steps > MAX_CORRECTIONS) die("reciprocal did not converge (corrupted product?)");
        bigint_sub_u64(X, 1);
        bigint_sub(&P, &P, Q);
    }
    bigint_sub(&r, T, &P);
    while (bigint_cmp(&r, Q) >= 0) {          /* X too small */
        if (++steps > MAX_CORRECTIONS) die("reciprocal did not converge (corrupted product?)");
        bigint_add_u64(X, 1);
        bigint_sub(&r, &r, Q);
    }
    bigint_free(&P);
    bigint_free(&r);
}

/* X = floor(2^(2k) / Q_k), Q_k the top k bits of Q (b = bits(Q) >= k). */
static void recip_top(BigInt *X, const BigInt *Q, int b, int k)
{
    BigInt Qk = bigint_alloc(4);
    bigint_shr(&Qk, Q, b - k);
    if (k <= 62) {
        u128 t = (u128)1 << (2 * k);
        set_u128(X, t / (u128)Qk.limbs[0]);
        bigint_free(&Qk);
        return;
    }
    const int h = k / 2 + 2;
    BigInt Y = bigint_alloc(4), P = bigint_alloc(4), T = bigint_alloc(4), e = bigint_alloc(4),
           Z = bigint_alloc(4);
    recip_top(&Y, Q, b, h);
    bigint_shl(&Y, &Y, k - h);                /* ~ 2^(2k) / Q_k */
    set_pow2(&T, 2 * k);
    bigint_mul(&P, &Qk, &Y);
    if (bigint_cmp(&P, &T) <= 0) {            /* Y below: add Y e / 2^(2k) */
        bigint_sub(&e, &T, &P);
        bigint_mul(&Z, &Y, &e);
        bigint_shr(&Z, &Z, 2 * k);
        bigint_add(X, &Y, &Z);
    } else {                                  /* Y above: subtract */
        bigint_sub(&e, &P, &T);
        bigint_mul(&Z, &Y, &e);
        bigint_shr(&Z, &Z, 2 * k);
        if (bigint_cmp(&Z, &Y) >= 0) die("reciprocal iterate underflow");
        bigint_sub(X, &Y, &Z);
    }
    correct_quotient(X, &T, &Qk);
    bigint_free(&Qk);
    bigint_free(&Y);
    bigint_free(&P);
    bigint_free(&T);
    bigint_free(&e);
    bigint_free(&Z);
}

void newton_reciprocal(BigInt *x, const BigInt *Q)
{
    if (bigint_is_zero(Q)) die("newton_reciprocal: division by zero");
    int b = bigint_bits(Q);
    recip_top(x, Q, b, b);
}

void bigint_div_newton(BigInt *q, const BigInt *N, const BigInt *D, const BigInt *recip)
{
    if (bigint_is_zero(D)) die("bigint_div_newton: division by zero");
    if (bigint_cmp(N, D) < 0) {
        bigint_resize(q, 1);
        bigint_zero(q);
        return;
    }
    const int b = bigint_bits(D), n = bigint_bits(N);
    BigInt X = bigint_alloc(4), q0 = bigint_alloc(4);
    int s;
    if (n <= 2 * b) {
        bigint_copy(&X, recip);
        s = 2 * b;
    } else {
        BigInt Ds = bigint_alloc(4);
        bigint_shl(&Ds, D, n - 2 * b);        /* reciprocal of D 2^(n-2b) is floor(2^n / D) */
        newton_reciprocal(&X, &Ds);
        bigint_free(&Ds);
        s = n;
    }
    bigint_mul(&q0, N, &X);
    bigint_shr(&q0, &q0, s);
    /* q0 <= floor(N / D) <= q0 + 3: correct from the remainder */
    BigInt P = bigint_alloc(4), r = bigint_alloc(4);
    bigint_mul(&P, &q0, D);
    int steps = 0;
    while (bigint_cmp(&P, N) > 0) {           /* only if recip was too large */
        if (++steps > MAX_CORRECTIONS) die("bigint_div_newton: no convergence (bad reciprocal?)");
        bigint_sub_u64(&q0, 1);
        bigint_sub(&P, &P, D);
    }
    bigint_sub(&r, N, &P);
    while (bigint_cmp(&r, D) >= 0) {
        if (++steps > MAX_CORRECTIONS) die("bigint_div_newton: no convergence (bad reciprocal?)");
        bigint_add_u64(&q0, 1);
        bigint_sub(&r, &r, D);
    }
    bigint_copy(q, &q0);
    bigint_free(&X);
    bigint_free(&q0);
    bigint_free(&P);
    bigint_free(&r);
}
