#pragma once
/*
 * multiply.h — drop-in for the reference's ntt/lib/arith/multiply.h.
 *
 * Same entry points and fail-stop behaviour (errors print and exit(1)); every
 * product is computed on the GPU: the schoolbook kernel at or below the
 * threshold, the 3-prime CRT-NTT above it.  Status-returning forms of the
 * same operations are in ntt_engine.h.
 */
#include "bigint.h"

#ifdef __cplusplus
extern "C" {
#endif

/* multiply.h:23-25: compile-time default for the crossover, in limbs. */
#ifndef BIGINT_MUL_THRESHOLD
#define BIGINT_MUL_THRESHOLD 64
#endif

/* multiply.h:30-31 */
int bigint_mul_get_threshold(void);
void bigint_mul_set_threshold(int t);

/* multiply.h:35-36.  max_log_n (3..30) is the largest transform the caller
 * plans to use; transforms are sized per product, so it only bounds the
 * range check the reference performs at init (ntt_mul.hip:396-403). */
void ntt_mul_init(int max_log_n);
void ntt_mul_teardown(void);

/* multiply.h:40 */
unsigned long ntt_get_gpu_dispatch_count(void);

/* multiply.h:46: the NTT path, whatever the operand sizes. */
void ntt_bigint_mul(BigInt *c, const BigInt *a, const BigInt *b);

/* multiply.h:50: c = a*b; c is resized as needed, the result trimmed.
 * c may alias a or b. */
void bigint_mul(BigInt *c, const BigInt *a, const BigInt *b);

/* No counterpart in the reference (it has no square entry point): c = a*a
 * with one forward transform instead of two. */
void bigint_sqr(BigInt *c, const BigInt *a);

/* multiply.c:94-98 test hooks and ntt_mul.hip:1247 signal-handler reset. */
void ntt_mul_init_impl(int max_log_n);
void ntt_mul_teardown_impl(void);
void ntt_gpu_emergency_reset(void);

#ifdef __cplusplus
}
#endif
