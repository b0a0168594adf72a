#pragma once
/*
 * newton.h — drop-in for the reference's ntt/lib/arith/newton.h:17-30:
 * reciprocal and division of big integers, every product on the GPU through
 * bigint_mul.  Fail-stop like the reference (errors print and exit(1)).
 */
#include "bigint.h"

#ifdef __cplusplus
extern "C" {
#endif

/* newton.h:23.  x = floor(2^(2b) / Q), b = bigint_bits(Q), Q > 0.  The
 * reference promises the floor or one less; this one returns the floor. */
void newton_reciprocal(BigInt *x, const BigInt *Q);

/* newton.h:28.  q = floor(N / D), exactly.  recip must be newton_reciprocal(D)
 * (the floor or one less); when N has more than 2*bits(D) bits a reciprocal
 * at a larger scale is formed internally. */
void bigint_div_newton(BigInt *q, const BigInt *N, const BigInt *D, const BigInt *recip);

#ifdef __cplusplus
}
#endif
