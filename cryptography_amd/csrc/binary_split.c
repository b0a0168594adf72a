/*
 * binary_split.c — product tree for sum 1/k! (reference API) and a BigInt
 * slab pool.
 *
 * binary_split follows ntt/app/compute_e/binary_split.c:74-133 (empty-range
 * identity, leaf P = 1, B = a, split m = a + (b-a)/2, B = B_L*B_R,
 * P = P_L*B_R + P_R) with every multiply on the GPU through bigint_mul.
 * binary_split_batched evaluates the same tree on the device level by level.
 */
#include "../../include/binary_split.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/multiply.h"
#include "../../include/ntt_engine.h"

/* ---- slab pool: free list of limb buffers, best fit (mem_pool.h:1-33) ---- */

#define POOL_SLOTS 256

struct MemPool {
    limb_t *buf[POOL_SLOTS];
    int cap[POOL_SLOTS];
    int n;
};

MemPool *mem_pool_create(void)
{
    MemPool *p = (MemPool *)calloc(1, sizeof(MemPool));
    if (!p) {
        fprintf(stderr, "mem_pool_create: out of memory\n");
        exit(1);
    }
    return p;
}

void mem_pool_destroy(MemPool *pool)
{
    if (!pool) return;
    for (int i = 0; i < pool->n; i++) free(pool->buf[i]);
    free(pool);
}

BigInt mem_pool_get(MemPool *pool, int min_cap)
{
    if (min_cap < 1) min_cap = 1;
    int best = -1;
    for (int i = 0; i < pool->n; i++)
        if (pool->cap[i] >= min_cap && (best < 0 || pool->cap[i] < pool->cap[best])) best = i;
    if (best < 0) return bigint_alloc(min_cap);
    BigInt b;
    b.limbs = pool->buf[best];
    b.cap = pool->cap[best];
    b.n_limbs = 0;
    b.managed = 0;
    memset(b.limbs, 0, (size_t)b.cap * sizeof(limb_t));
    pool->n--;
    pool->buf[best] = pool->buf[pool->n];
    pool->cap[best] = pool->cap[pool->n];
    return b;
}

void mem_pool_put(MemPool *pool, BigInt *b)
{
    if (!b->limbs) return;
    if (b->managed || pool->n == POOL_SLOTS) {
        bigint_free(b);
        return;
    }
    pool->buf[pool->n] = b->limbs;
    pool->cap[pool->n] = b->cap;
    pool->n++;
    b->limbs = NULL;
    b->n_limbs = 0;
    b->cap = 0;
}

/* ---- tree ------------------------------------------------------------------ */

static BigInt take(MemPool *pool, int cap) { return pool ? mem_pool_get(pool, cap) : bigint_alloc(cap); }
static void give(MemPool *pool, BigInt *b)
{
    if (pool) mem_pool_put(pool, b);
    else bigint_free(b);
}

void split_free(SplitResult *s)
{
    bigint_free(&s->P);
    bigint_free(&s->B);
}

void binary_split(SplitResult *out, long a, long b, MemPool *pool)
{
    if (b <= a) {
        out->P = take(pool, 1);
        out->B = take(pool, 1);
        bigint_set_u64(&out->P, 0);
        bigint_set_u64(&out->B, 1);
        return;
    }
    if (b - a == 1) {
        out->P = take(pool, 1);
        out->B = take(pool, 1);
        bigint_set_u64(&out->P, 1);
        bigint_set_u64(&out->B, (uint64_t)a);
        return;
    }
    long m = a + (b - a) / 2;
    SplitResult L, R;
    binary_split(&L, a, m, pool);
    binary_split(&R, m, b, pool);
    out->B = take(pool, L.B.n_limbs + R.B.n_limbs + 2);
    bigint_mul(&out->B, &L.B, &R.B);
    BigInt t = take(pool, L.P.n_limbs + R.B.n_limbs + 2);
    bigint_mul(&t, &L.P, &R.B);
    out->P = take(pool, (t.n_limbs > R.P.n_limbs ? t.n_limbs : R.P.n_limbs) + 2);
    bigint_add(&out->P, &t, &R.P);
    give(pool, &t);
    give(pool, &L.P);
    give(pool, &L.B);
    give(pool, &R.P);
    give(pool, &R.B);
}

/* binary_split.c:143-155: Stirling fixed point for N ln N - N > D ln 10,
 * then a 5 % margin and +64 terms. */
long split_terms_needed(long n_digits)
{
    double target = (double)n_digits * log(10.0);
    double N = target / log(target);
    for (int i = 0; i < 50; i++) {
        double lnN = log(N);
        if (lnN <= 1.0) {
            N = target;
            break;
        }
        N = target / (lnN - 1.0);
    }
    return (long)(N * 1.05) + 64;
}

static void tree_fatal(const char *what, int rc)
{
    fprintf(stderr, "binary_split_batched: %s: %s (status %d)\n", what, ntt_engine_last_error(), rc);
    exit(1);
}

void binary_split_batched(SplitResult *out, long a, long b)
{
    if (!ntt_engine_is_init()) {
        int rc = ntt_engine_init(0);
        if (rc) tree_fatal("init", rc);
    }
    ntt_split_tree *t = NULL;
    int rc = ntt_split_tree_create(a, b, 1, &t);
    if (rc) tree_fatal("create", rc);
    rc = ntt_split_tree_run(t);
    if (rc) tree_fatal("run", rc);
    size_t capP, capB, nP, nB;
    ntt_split_tree_root_caps(t, &capP, &capB);
    out->P = bigint_alloc((int)capP);
    out->B = bigint_alloc((int)capB);
    rc = ntt_split_tree_result(t, out->P.limbs, capP, &nP, out->B.limbs, capB, &nB);
    if (rc) tree_fatal("result", rc);
    out->P.n_limbs = (int)nP;
    out->B.n_limbs = (int)nB;
    ntt_split_tree_destroy(t);
}
