#pragma once
/*
 * bigint.h — arbitrary-precision unsigned integer, drop-in for the reference's
 * ntt/lib/arith/bigint.h (LIMB_BITS = 64 build).
 *
 * value = sum_i limbs[i] * 2^(64 i), little-endian, n_limbs trimmed.
 * Same struct layout and the same helper set as the reference
 * (bigint.h:69-150) so callers link unchanged.  The reference's host-side NTT
 * helpers bigint_scatter / bigint_gather (bigint.h:113-138) have no
 * counterpart: the split and the carry gather run on the GPU inside the
 * multiply.
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIMB_BITS 64
typedef uint64_t limb_t;
#define LIMB_MASK (~(uint64_t)0)
#define LIMB_BYTES ((int)sizeof(limb_t))

/* bigint.h:52-66: the full 128-bit product of two limbs, *hi:*lo = a*b.  At
 * 64-bit limbs the compiler's 128-bit multiply gives it directly. */
static inline void mul_limb(limb_t a, limb_t b, limb_t *lo, limb_t *hi)
{
    const unsigned __int128 p = (unsigned __int128)a * b;
    *lo = (limb_t)p;
    *hi = (limb_t)(p >> 64);
}

/* bigint.h:69-74 */
typedef struct {
    limb_t *limbs;
    int n_limbs;
    int cap;
    int managed;   /* 1: limbs from hipMallocManaged (never realloc'd) */
} BigInt;

/* bigint.h:77-84 */
BigInt bigint_alloc(int cap);
BigInt bigint_alloc_managed(int cap);
void bigint_free(BigInt *a);
void bigint_resize(BigInt *a, int new_cap);

/* bigint.h:87-89 */
void bigint_zero(BigInt *a);
void bigint_set_u64(BigInt *a, uint64_t v);
void bigint_copy(BigInt *dst, const BigInt *src);

/* bigint.h:92-93 */
int bigint_is_zero(const BigInt *a);
int bigint_cmp(const BigInt *a, const BigInt *b);

/* bigint.h:96-112 */
void bigint_add(BigInt *c, const BigInt *a, const BigInt *b);
void bigint_sub(BigInt *c, const BigInt *a, const BigInt *b);
void bigint_add_u64(BigInt *a, uint64_t v);
void bigint_sub_u64(BigInt *a, uint64_t v);
void bigint_shl(BigInt *c, const BigInt *a, int bits);
void bigint_shr(BigInt *c, const BigInt *a, int bits);
void bigint_mul_u64(BigInt *c, const BigInt *a, uint64_t v);

/* bigint.h:141-150 */
int bigint_bits(const BigInt *a);
static inline int bigint_limbs_for_digits(int n_digits)
{
    return (int)((uint64_t)n_digits * 332194 / 100000 / 64) + 4;
}
void bigint_print_hex(const BigInt *a, const char *label);

#ifdef __cplusplus
}
#endif
