/*
 * modarith.h — exact 64-bit modular arithmetic for the three NTT primes.
 *
 * Host and device.  The primes are p = 2^60 - c*2^32 + 1 (c < 2^8), so
 * p = 1 (mod 2^32), every power-of-two transform length up to 2^35 has
 * roots, lazily reduced values below 16p still fit one u64 word, and one
 * subtraction of floor(x / 2^60) * p brings any 64-bit x below 2p
 * (reduce_any).  Arithmetic follows the reference's primitives
 * (ntt/lib/primes.h:193-279, ntt/lib/shoup.h:45-112) with the build's primes
 * and lazy ranges.
 */
#ifndef CRYPTOGRAPHY_AMD_MODARITH_H
#define CRYPTOGRAPHY_AMD_MODARITH_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#define NTT_NPRIMES 3
/* Twiddles are powers of one root of order 2^NTT_ROOT_LOG per prime. */
#define NTT_ROOT_LOG 30
#define NTT_TW_LOG 15                      /* two-level table: 2^15 entries each */
#define NTT_TW_SIZE (1u << NTT_TW_LOG)
#define NTT_MAX_LOG_LEN NTT_ROOT_LOG       /* longest transform: 2^30 coefficients */
#define NTT_NOALLOC_MAX_LOG 19             /* longest transform of the no-allocation multiply */

namespace nttc {

/* prime i: value (2^60 - c*2^32 + 1 for c = 96, 120, 168), exponent k of
 * the largest power of two dividing p - 1 (p = c'*2^k + 1, c' odd),
 * generator of Z_p^* */
__host__ __device__ constexpr uint64_t prime(int i)
{
    return i == 0 ? 0x0fffffa000000001ull : i == 1 ? 0x0fffff8800000001ull : 0x0fffff5800000001ull;
}
__host__ __device__ constexpr uint64_t prime_gen(int i) { return i == 0 ? 3 : i == 1 ? 3 : 6; }
__host__ __device__ constexpr int prime_k(int i) { return i == 0 ? 37 : i == 1 ? 35 : 35; }

/* -p^-1 mod 2^64 by Newton iteration (each step doubles the correct bits). */
__host__ __device__ constexpr uint64_t neg_inv64(uint64_t p)
{
    uint64_t x = p;                 /* p*p == 1 mod 8 for odd p: 3 bits */
    x *= 2 - p * x; x *= 2 - p * x; x *= 2 - p * x; x *= 2 - p * x; x *= 2 - p * x;
    return (uint64_t)0 - x;
}

#if defined(__HIP_DEVICE_COMPILE__)
/* gfx950 keeps 64-bit operands in even-aligned register pairs, so a 32-bit
 * value used as a zero-extended 64-bit addend of v_mad_u64_u32 costs register
 * moves when it sits in the odd half of a pair (the high word of a product).
 * The two partial products of weight 2^32 of a 64 x 64-bit product are
 * therefore summed with their full 64-bit partner in one v_mad_u64_u32 whose
 * carry out (a lane mask) goes into the top word afterwards.
 * mad_carry: y = a*b + x (mod 2^64) and its carry out. */
__device__ __forceinline__ uint64_t mad_carry(uint32_t a, uint32_t b, uint64_t x, uint64_t &carry)
{
    uint64_t y;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(y), "=s"(carry) : "v"(a), "v"(b), "v"(x));
    return y;
}
/* z + carry * 2^32 (the carry out of mad_carry). */
__device__ __forceinline__ uint64_t add_carry_hi(uint64_t z, uint64_t carry)
{
    uint32_t hi;
    uint64_t c2;
    asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(hi), "=s"(c2) : "v"((uint32_t)(z >> 32)), "s"(carry));
    return ((uint64_t)hi << 32) | (uint32_t)z;
}
#endif

__host__ __device__ __forceinline__ uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* hi64(a*b) = a1*b1 + floor((a1*b0 + X) / 2^32), X = a0*b1 + hi32(a0*b0) < 2^64 */
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t X = (uint64_t)a0 * b1 + __umulhi(a0, b0);
    uint64_t c;
    const uint64_t Y = mad_carry(a1, b0, X, c);
    return add_carry_hi((uint64_t)a1 * b1 + (Y >> 32), c);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

template <int PI>
struct Prime {
    static constexpr uint64_t P = prime(PI);
    static constexpr uint64_t NINV = neg_inv64(prime(PI));   /* -p^-1 mod 2^64 */
};

/* x in [0, 2p) -> [0, p) */
template <int PI>
__host__ __device__ __forceinline__ uint64_t reduce1(uint64_t x)
{
    return x >= Prime<PI>::P ? x - Prime<PI>::P : x;
}
/* any x < 2^64 -> [0, 2p), congruent: x - q*p for q = floor(x / 2^60) < 16.
 * With x = q*2^60 + s (s < 2^60) and d = 2^60 - p = c*2^32 - 1, the result is
 * s + q*d < 2^60 + 15d < 2p (as 17d < 2^60), and it is >= 0 as q*p <= x.
 * p = ph*2^32 + 1, so q*p = q + (q*ph)*2^32 with q*ph < 2^32: a shift, one
 * 32-bit multiply and a 64-bit subtraction. */
template <int PI>
__host__ __device__ __forceinline__ uint64_t reduce_any(uint64_t x)
{
    static_assert(prime(PI) < (1ull << 60) && 17 * ((1ull << 60) - prime(PI)) < (1ull << 60),
                  "p just below 2^60");
    constexpr uint32_t PH = (uint32_t)(prime(PI) >> 32);
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t q = xh >> 28;
    uint32_t t = q * PH;
#if defined(__HIP_DEVICE_COMPILE__)
    /* a plain multiply: merged into xh - q*PH as one 64-bit multiply-add, xh
     * would first need a move into an even register pair */
    asm("" : "+v"(t));
#endif
    const uint32_t lo = xl - q;
    const uint32_t hi = xh - t - (uint32_t)(xl < q);
    return ((uint64_t)hi << 32) | lo;
}

/* Shoup product a*w mod p for any a < 2^64, w < p, wp = floor(w*2^64/p);
 * result in [0, 2p).  q = hi(a*wp) is floor(a*w/p) or one less. */
template <int PI>
__host__ __device__ __forceinline__ uint64_t shoup_lazy(uint64_t a, uint64_t w, uint64_t wp)
{
    uint64_t q = mulhi64(a, wp);
    return a * w - q * Prime<PI>::P;
}

/* Shoup product a*w mod p with a truncated quotient, for any a < 2^64, w < p,
 * wp = floor(w*2^64/p); result in [0, 4p).  With a = a1*2^32 + a0 and
 * wp = wp1*2^32 + wp0, floor(a*wp / 2^64) = a1*wp1 + floor((A + B + C) / 2^32)
 * for A = a1*wp0, B = a0*wp1, C = hi32(a0*wp0).  q = a1*wp1 + hi32(A) + hi32(B)
 * drops (A mod 2^32) + (B mod 2^32) + C < 3*2^32, so it is at most two below
 * floor(a*wp / 2^64), itself floor(a*w/p) or one less: a*w - q*p < 4p.  Eight
 * 32-bit multiplies and no carry (shoup_lazy: nine and a carry). */
template <int PI>
__host__ __device__ __forceinline__ uint64_t shoup_lazy4(uint64_t a, uint64_t w, uint64_t wp)
{
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), w0 = (uint32_t)wp, w1 = (uint32_t)(wp >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t h = (uint64_t)__umulhi(a1, w0) + __umulhi(a0, w1);
#else
    const uint64_t h = (((uint64_t)a1 * w0) >> 32) + (((uint64_t)a0 * w1) >> 32);
#endif
    const uint64_t q = (uint64_t)a1 * w1 + h;
    return a * w - q * Prime<PI>::P;
}

/* Shoup product a*w mod p for a < 2^32; result in [0, 2p).  q = hi64(a*wp)
 * exactly: a*wp = a*wp1*2^32 + a*wp0, and with t = a*wp1 + hi32(a*wp0) (an
 * integer below 2^64) floor(a*wp / 2^64) = floor(t / 2^32), since the
 * dropped part is below 1.  Five 32-bit multiplies instead of nine. */
template <int PI>
__host__ __device__ __forceinline__ uint64_t shoup_small(uint32_t a, uint64_t w, uint64_t wp)
{
    const uint64_t t = (uint64_t)a * (uint32_t)(wp >> 32) + (uint32_t)(((uint64_t)a * (uint32_t)wp) >> 32);
    const uint32_t q = (uint32_t)(t >> 32);
    return (uint64_t)a * w - (uint64_t)q * Prime<PI>::P;
}

/* Whether residues x_i in [0, p_i) of the three primes are those of one
 * integer below p0, i.e. whether the Garner digits a1, a2 of
 * x0 + a1*p0 + a2*p0*p1 are both 0, without computing them:
 * a1 = (x1 - x0) * p0^-1 mod p1 is 0 iff x1 = x0 mod p1, and then
 * a2 = (x2 - x0) * (p0 p1)^-1 mod p2 is 0 iff x2 = x0 mod p2.  As p0 < 2 p1
 * and p0 < 2 p2, one conditional subtraction reduces x0 modulo each. */
__host__ __device__ __forceinline__ bool crt_below_p0(uint64_t x0, uint64_t x1, uint64_t x2)
{
    static_assert(prime(0) < 2 * prime(1) && prime(0) < 2 * prime(2), "p0 < 2 p1 and p0 < 2 p2");
    const uint64_t m1 = x0 >= prime(1) ? x0 - prime(1) : x0;
    const uint64_t m2 = x0 >= prime(2) ? x0 - prime(2) : x0;
    return (m1 == x1) & (m2 == x2);
}

/* The integer X < 2 p0 with residues x_i in [0, p_i), if there is one: X is
 * x0 (Garner digits a1 = a2 = 0, crt_below_p0) or x0 + p0 (a1 = 1, a2 = 0:
 * x0 + p0 = x1 mod p1 and = x2 mod p2).  p0 > p1, p2 and p0 - p_i < p_i, so
 * (x0 + p0) mod p_i = (x0 mod p_i) + (p0 - p_i), less p_i when it reaches
 * p_i.  Sets ok to false when neither holds. */
__host__ __device__ __forceinline__ uint64_t crt_below_2p0(uint64_t x0, uint64_t x1, uint64_t x2, bool &ok)
{
    static_assert(prime(0) > prime(1) && prime(0) > prime(2) && prime(0) < 2 * prime(1) &&
                      prime(0) < 2 * prime(2),
                  "p1, p2 < p0 < 2 p1, 2 p2");
    constexpr uint64_t D1 = prime(0) - prime(1), D2 = prime(0) - prime(2);
    const uint64_t m1 = x0 >= prime(1) ? x0 - prime(1) : x0;
    const uint64_t m2 = x0 >= prime(2) ? x0 - prime(2) : x0;
    uint64_t n1 = m1 + D1, n2 = m2 + D2;
    n1 = n1 >= prime(1) ? n1 - prime(1) : n1;
    n2 = n2 >= prime(2) ? n2 - prime(2) : n2;
    const bool lo = (m1 == x1) & (m2 == x2), hi = (n1 == x1) & (n2 == x2);
    ok = lo | hi;
    return hi && !lo ? x0 + prime(0) : x0;
}

/* Montgomery product a*b*2^-64 mod p for a*b < 2^64 * p; result in [0, p). */
template <int PI>
__host__ __device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b)
{
    uint64_t tlo = a * b, thi = mulhi64(a, b);
    uint64_t m = tlo * Prime<PI>::NINV;
    uint64_t mphi = mulhi64(m, Prime<PI>::P);
    /* t + m*p is divisible by 2^64; its low word sum carries iff tlo != 0 */
    uint64_t u = thi + mphi + (tlo != 0);
    return u >= Prime<PI>::P ? u - Prime<PI>::P : u;
}

/* ---- compile-time roots --------------------------------------------------
 * Evaluated by the compiler only (static constexpr members), for the few
 * twiddles of the smallest stages, which are then immediates in the code. */
__host__ __device__ constexpr uint64_t c_mulmod(uint64_t a, uint64_t b, uint64_t p)
{
    return (uint64_t)((unsigned __int128)a * b % p);
}
__host__ __device__ constexpr uint64_t c_powmod(uint64_t b, uint64_t e, uint64_t p)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = c_mulmod(r, b, p);
        b = c_mulmod(b, b, p);
        e >>= 1;
    }
    return r;
}
/* w^E for the root w = g^((p-1)/2^LOGORD) of order 2^LOGORD (the same root as
 * the device tables' w^(2^(30-LOGORD))), with its Shoup partner. */
template <int PI, int LOGORD, int E>
struct ConstTw {
    static constexpr uint64_t P = prime(PI);
    static constexpr uint64_t root = c_powmod(prime_gen(PI), (P - 1) >> LOGORD, P);
    static constexpr uint64_t w = c_powmod(root, (uint64_t)E, P);
    static constexpr uint64_t wp = (uint64_t)(((unsigned __int128)w << 64) / P);
};

/* Montgomery product a*b*2^-64 mod p for any a < 2^64 and b < 2p; result in
 * (T/2^64, T/2^64 + p] for T = a*b (derivation below): below 3p always,
 * below 2p when T < 2^64 p (b < p, or a < 2^63), below p + 1 when T < 2^64.
 * Every prime here is 1 mod 2^32 (p = ph*2^32 + 1), so p^-1 = 1 (mod 2^32)
 * and the reduction is two subtractive 32-bit steps with free multipliers:
 * T = a*b = hi*2^64 + l1*2^32 + l0; T - l0*p = 2^32 (hi*2^32 + D) with
 * D = l1 - l0*ph (signed); with s0 = D mod 2^32, (hi*2^32 + D) - s0*p =
 * 2^32 (hi + floor(D / 2^32) - s0*ph).  That last value is
 * (T - m p) / 2^64 for m = l0 + s0*2^32 < 2^64, so it lies in
 * (T/2^64 - p, T/2^64]; adding p gives (T/2^64, T/2^64 + p], and
 * T/2^64 < 2p for T < 2^65 p.  floor(D / 2^32) lies in [-ph, 0], an int32
 * as ph < 2^28.  The 128-bit product takes four
 * 32x32->64 multiplies and each step one more: no carries to compare. */
template <int PI>
__host__ __device__ __forceinline__ uint64_t mont_mul_lazy(uint64_t a, uint64_t b)
{
    static_assert((prime(PI) & 0xffffffffull) == 1, "p must be 1 mod 2^32");
    constexpr uint64_t P = prime(PI);
    constexpr uint32_t NPH = 0u - (uint32_t)(P >> 32);             /* 2^32 - ph */
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t cy;   /* u = a1*b0 + t as a 65-bit sum: low 64 bits and carry */
    const uint64_t u = mad_carry(a1, b0, t, cy);
    const uint64_t hi = add_carry_hi((uint64_t)a1 * b1 + (u >> 32), cy);
#else
    const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;
    const uint64_t hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
#endif
    const uint32_t l0 = (uint32_t)p00, l1 = (uint32_t)u;
    const uint64_t dd = (uint64_t)l0 * NPH + l1;                   /* D + l0*2^32 */
    const uint32_t s0 = (uint32_t)dd;                              /* D mod 2^32 */
    uint32_t ddh = (uint32_t)(dd >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    /* the two high-word subtractions stay 32-bit (the compiler otherwise
     * widens each into a 64-bit subtraction with a borrow: 19 VALU
     * instructions for the product instead of 15) */
    asm("" : "+v"(ddh));
#endif
    const uint32_t dhi = ddh - l0;                                 /* floor(D / 2^32), two's complement */
    const uint64_t x = hi + (uint64_t)(int64_t)(int32_t)dhi + P;
    const uint64_t y = (uint64_t)s0 * NPH + x;
    uint32_t yh = (uint32_t)(y >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+v"(yh));
#endif
    return ((uint64_t)(yh - s0) << 32) | (uint32_t)y;               /* x - s0*ph */
}

}  // namespace nttc

#endif
