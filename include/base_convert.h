#pragma once
/*
 * base_convert.h — drop-in for the reference's ntt/lib/arith/base_convert.h:
 * decimal conversion of big integers by divide and conquer over a cache of
 * 10^(2^k) and their reciprocals; every product on the GPU through
 * bigint_mul.  Fail-stop like the reference.
 */
#include <stddef.h>
#include <stdint.h>

#include "bigint.h"

#ifdef __cplusplus
extern "C" {
#endif

/* base_convert.h:19-20.  Cache pow10[k] = 10^(2^k) and floor(2^(2 bits)/pow10[k])
 * for k = 0..max_level (max_level <= 40). */
void base_convert_init(int max_level);
void base_convert_teardown(void);

/* base_convert.h:24.  Decimal digits of N, NUL-terminated, malloc'd (caller
 * frees); "0" for zero.  NULL when N >= 10^(2^(max_level+1)) (the cache is too
 * small) or before base_convert_init. */
char *bigint_to_decimal(const BigInt *N);

/* No single counterpart in the reference (its conversion is the divide and
 * conquer above, base_convert.c:135-259): the digits of the integer held in
 * n little-endian u64 limbs, without leading zeros and NUL-terminated, into
 * out (cap >= floor(bits(N) * 0.30103) + 2 bytes); *len = the digit count.
 * Device-resident: one upload, a Newton reciprocal of 2 * 10^D and one
 * product for the fraction (N + 1/2) / 10^D, a level-batched radix tree,
 * one download.  bigint_to_decimal uses it above DECIMAL_DEVICE_LIMBS limbs.
 * Returns an NTT_ERR_* code (integrity errors included: see ntt_engine.h). */
int ntt_to_decimal(const uint64_t *limbs, size_t n, char *out, size_t cap, size_t *len);

/* base_convert.h:28.  Writes the digits of N to fd (no terminator); returns
 * the digit count, or -1 on error. */
long bigint_write_decimal(const BigInt *N, int fd);

/* base_convert.h:32.  out = 10^n from the cached 10^(2^k) (one product per set
 * bit of n); needs max_level >= floor(log2(n)). */
void base_convert_pow10_exact(BigInt *out, long n);

#ifdef __cplusplus
}
#endif
