/*
 * ntt_kernels.h — kernel argument blocks and declarations shared by
 * ntt_kernels.hip and the host engine.
 */
#ifndef CRYPTOGRAPHY_AMD_NTT_KERNELS_H
#define CRYPTOGRAPHY_AMD_NTT_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

/* Shoup pair: twiddle w and floor(w * 2^64 / p), one 16-byte load. */
struct __attribute__((aligned(16))) TwPair {
    uint64_t w, wp;
};

#define CARRY_E 8   /* output limbs per thread in the carry kernels */

struct PassArgs {
    uint64_t *res;              /* res[op][prime][len] */
    const TwPair *tw;           /* twiddle tables, 3 primes x 2 x 2^15 */
    const uint16_t *src16[2];   /* first forward pass: 16-bit coefficients per operand */
    size_t n16[2];              /* coefficient counts (zero beyond) */
    size_t len;                 /* transform length L */
    uint32_t logS, logT, logC, logG;
    int first;
};

struct PointwiseArgs {
    uint64_t *res;
    size_t len;
    uint64_t scale[3], scale_p[3];   /* 2^64 / L mod p and Shoup partners */
    int square;
};

struct GarnerConsts {
    uint64_t inv01, inv01p;      /* p0^-1 mod p1 */
    uint64_t p0m2, p0m2p;        /* p0 mod p2 */
    uint64_t inv012, inv012p;    /* (p0 p1)^-1 mod p2 */
};

struct GarnerArgs {
    const uint64_t *res;        /* res[prime][len], operand a's slot */
    size_t len;
    size_t n_coef;              /* coefficients that may be nonzero */
    size_t n_out;               /* output limbs */
    uint64_t *out;
    uint8_t *gen;               /* one byte of generate bits per thread */
    uint8_t *blk;               /* block aggregates, then block carry-ins */
    int *flag;                  /* integrity flag: set on a corrupted product */
    GarnerConsts gc;
};

__global__ void k_twiddle_gen(TwPair *tw, uint64_t r2_0, uint64_t r2_1, uint64_t r2_2);
__global__ void k_fwd_pass(PassArgs pa);
__global__ void k_inv_pass(PassArgs pa);
__global__ void k_pointwise(PointwiseArgs pw);
__global__ void k_garner_pack(GarnerArgs ga);
__global__ void k_carry_scan(uint8_t *blk, size_t nblk);
__global__ void k_carry_apply(GarnerArgs ga);
__global__ void k_schoolbook(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out);

#endif
