#pragma once
/*
 * base_convert.h — drop-in for the reference's ntt/lib/arith/base_convert.h:
 * decimal conversion of big integers by divide and conquer over a cache of
 * 10^(2^k) and their reciprocals; every product on the GPU through
 * bigint_mul.  Fail-stop like the reference.
 */
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

/* base_convert.h:28.  Writes the digits of N to fd (no terminator); returns
 * the digit count, or -1 on error. */
long bigint_write_decimal(const BigInt *N, int fd);

/* base_convert.h:32.  out = 10^n from the cached 10^(2^k) (one product per set
 * bit of n); needs max_level >= floor(log2(n)). */
void base_convert_pow10_exact(BigInt *out, long n);

#ifdef __cplusplus
}
#endif
