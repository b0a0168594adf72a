#pragma once
/*
 * ntt_poly.h — products of integer polynomials on the CRT-NTT engine, the
 * second client of the transform kernels.
 *
 * Counterpart of the reference's polynomial pipeline (cyclic
 * ntt/lib/main.hip:434-552, negacyclic :554-683 with the twist kernels of
 * ntt/lib/ntt_kernel.hip:517-564), which runs across four GPUs with four
 * primes and returns 256-bit CRT values; here one GPU, three primes, 192-bit
 * values.
 *
 * out[3k .. 3k+2] (little-endian words) = coefficient k of f * g mod (X^n - 1)
 * (negacyclic = 0) or mod (X^n + 1) (negacyclic = 1), as its residue modulo
 * P = p0 p1 p2 (P > 2^179): exact for every input (|c_k| < n 2^128 < P/2), a
 * negative negacyclic coefficient c reads as P + c.  f and g hold n
 * coefficients < 2^64 each; n is a power of two in [2, 2^NTT_POLY_MAX_LOG].
 * Returns an ntt_status (ntt_engine.h).
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NTT_POLY_MAX_LOG 27

/* Device pointers (out: 3n words). */
int ntt_poly_mul_device(uint64_t *d_out, const uint64_t *d_f, const uint64_t *d_g, size_t n, int negacyclic);
/* Host buffers, staged through the device. */
int ntt_poly_mul(uint64_t *out, const uint64_t *f, const uint64_t *g, size_t n, int negacyclic);

#ifdef __cplusplus
}
#endif
