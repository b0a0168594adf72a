#pragma once
/*
 * binary_split.h — drop-in for the reference's
 * ntt/app/compute_e/binary_split.h and mem_pool.h, plus the level-batched
 * device tree.
 *
 * split(a, b) gives (P, B) for sum_{k=a}^{b-1} 1/k! with B = (b-1)!/(a-1)!
 * and P / (b-1)! equal to that partial sum; leaf P = 1, B = a; merge at
 * m = a + (b-a)/2: B = B_L*B_R, P = P_L*B_R + P_R.
 */
#include <stddef.h>
#include <stdint.h>

#include "bigint.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- mem_pool.h:19-33 ---------------------------------------------------- */
typedef struct MemPool MemPool;
MemPool *mem_pool_create(void);
void mem_pool_destroy(MemPool *pool);
BigInt mem_pool_get(MemPool *pool, int min_cap);   /* zeroed, n_limbs 0 */
void mem_pool_put(MemPool *pool, BigInt *b);       /* recycles, detaches b */

/* ---- binary_split.h:28-47 ------------------------------------------------- */
typedef struct {
    BigInt P;
    BigInt B;
} SplitResult;

/* Depth-first recursion with one bigint_mul per product (the reference's
 * plain path; every product runs on the GPU). pool may be NULL. */
void binary_split(SplitResult *out, long a, long b, MemPool *pool);
void split_free(SplitResult *s);
long split_terms_needed(long n_digits);

/* No counterpart in the reference: the same (P, B), evaluated level by level
 * on the device (leaf kernel + one batch per tree height, one hipGraph).
 * Fail-stop like binary_split. */
void binary_split_batched(SplitResult *out, long a, long b);

/* ---- status-returning device tree ------------------------------------------
 * create: plan the tree and allocate its device arena.  use_graph: 1 captures
 * the whole level-batched evaluation into one hipGraph, 0 launches the same
 * batches eagerly, 2 launches every product on its own (the unbatched device
 * path, for comparison).  run: evaluate and wait. */
typedef struct ntt_split_tree ntt_split_tree;
int ntt_split_tree_create(long a, long b, int use_graph, ntt_split_tree **out);
int ntt_split_tree_run(ntt_split_tree *t);
int ntt_split_tree_info(ntt_split_tree *t, size_t *n_products, int *n_levels, size_t *n_leaves,
                        size_t *arena_bytes);
int ntt_split_tree_root_caps(ntt_split_tree *t, size_t *capP, size_t *capB);
/* Root P and B copied to host buffers of at least the root capacities;
 * *nP / *nB receive the trimmed lengths. */
int ntt_split_tree_result(ntt_split_tree *t, uint64_t *P, size_t capP, size_t *nP, uint64_t *B, size_t capB,
                          size_t *nB);
/* Root P and B left on the device (pointers into the tree's arena, valid
 * until destroy; lengths are capacities, high limbs zero). */
int ntt_split_tree_device_result(ntt_split_tree *t, const uint64_t **dP, size_t *capP, const uint64_t **dB,
                                 size_t *capB);
void ntt_split_tree_destroy(ntt_split_tree *t);

#ifdef __cplusplus
}
#endif
