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

#define CW_ROUNDS 4                     /* carry kernels: rounds of 64 limbs per wavefront */
#define CW_LIMBS (64 * CW_ROUNDS)       /* limbs per wavefront */
#define CB_LIMBS (4 * CW_LIMBS)         /* limbs per 256-thread block */

/* Residue rows are padded: after every 2^pad_log residues come NTT_PAD_WORDS
 * unused words, so the rows of the first (most strided) pass of a long
 * transform are not a power-of-two number of bytes apart.  pad_log =
 * NTT_PAD_NONE disables it (every index is below 2^40).  A 1 KB pad measured
 * best for the 2 MB rows of the longest products: with 512 B the first pass
 * ran about 8% slower in a third of the processes. */
#define NTT_PAD_WORDS 128
#define NTT_PAD_NONE 40
__host__ __device__ __forceinline__ size_t res_at(size_t a, uint32_t pad_log)
{
    return a + ((a >> pad_log) * NTT_PAD_WORDS);
}

/* One product of a batch: c[0..n_out) = a*b (+ add).  Device pointers. */
struct MulDesc {
    const uint64_t *a, *b;      /* operand limbs */
    uint64_t *c;                /* product limbs */
    const uint64_t *add;        /* optional addend limbs (nullptr: none) */
    uint32_t na, nb, nadd, n_out;
};

/* A size class: nprod products sharing one transform length L = 2^logL.
 * Residues of product i occupy [i*L, (i+1)*L) of each (op, prime) row. */
struct PassArgs {
    uint64_t *res;              /* res[op][prime][len] */
    const TwPair *tw;           /* twiddle tables, 3 primes x 2 x 2^15 */
    const TwPair *small_tw;     /* w_4096^e, e < 2048, per prime */
    const MulDesc *desc;        /* nprod descriptors (device), or nullptr: use one */
    MulDesc one;
    size_t len;                 /* row stride in words (padded) */
    uint32_t pad_log;           /* residue index a lives at res_at(a, pad_log) */
    uint32_t nprod;
    uint32_t logL;
    uint32_t logT, logC, logG;  /* this pass: column stride, tile columns, tile blocks */
    uint32_t ntiles;            /* tiles per (operand, prime) row */
    uint32_t tpb;               /* forward / inverse passes: tiles per block */
    int first;                  /* read the operands' 16-bit coefficients */
    int square;                 /* b == a: one forward transform */
    uint32_t op0;               /* forward passes: operand of grid z = 0 (1: b only) */
    uint32_t b_tiles;           /* fused pass: b spectra repeat, tile t's is at
                                   t mod b_tiles (0: every product has its own) */
    uint64_t scale[3], scale_p[3];   /* 2^64 / L mod p and Shoup partners (middle pass) */
    uint64_t scale_m[3];        /* the same times 2^64 mod p (Montgomery form) */
    int scale_late;             /* middle pass: leave the 1/L scale to the last inverse pass */
    int half_zero;              /* first pass: every operand has at most L/2 coefficients, so
                                   rows S/2.. of each column are zero */
    int iscale;                 /* kinds 5 and 9: multiply by 1/L (the chain's first twiddle
                                   carries it); kind 7 reads a table that includes it */
    const uint64_t *xtab;       /* kinds 6 and 7: this pass's cross twiddles, one
                                   entry per residue of an S x T block, per prime */
    const TwPair *btab;         /* kinds 8 and 9: the cross-twiddle chain's ratio B
                                   per column with its Shoup partner, T entries per
                                   prime (one direction of a btab_for table) */
    const uint64_t *bres;       /* fused pass: b's rows (after all forward passes
                                   but the last) kept outside the class, one
                                   product long, prime i at bres + i*blen; every
                                   product reads them (b_tiles = L/4096); nullptr:
                                   b is the class's second operand */
    size_t blen;
};

#define FT_TILE 4096            /* residues per transform tile */
#define FT_MAX_LOGS 10          /* largest sub-transform of one pass */
hipError_t launch_fast_pass(int logS, int kind, const PassArgs &pa, size_t ntiles, int nops, hipStream_t st);
hipError_t launch_small_twiddles(const TwPair *tw, TwPair *small, hipStream_t st);
/* Cross-twiddle tables of a pass with S = 2^logS rows and T = 2^logT columns
 * (N = S*T <= 2^NTT_XTAB_MAX_LOG): out[dir][prime][m*T + r] =
 * w_N^(+-r*bitrev_S(m)) * 2^64 mod p (Montgomery form, in [0, p)); dir 0
 * forward, 1 inverse, 2 inverse times scale[prime] (2^64 / N mod p: the
 * 1/N of a whole length-N transform folded into its last inverse pass). */
#define NTT_XTAB_MAX_LOG 20
#define NTT_XTAB_DIRS 3
struct XtabScale {
    uint64_t s[3], sp[3];       /* 2^64 / N mod p and Shoup partners */
};
hipError_t launch_xtab_gen(const TwPair *tw, uint64_t *out, int logS, int logT, const XtabScale &sc, hipStream_t st);
/* Chain ratios of a pass with S = 2^logS rows and T = 2^logT columns (N = S*T):
 * out[dir][prime][r] = (B, floor(B * 2^64 / p)) for column r < T, B in [0, p):
 * dir 0 B = w_N^(r * S/8) (the first forward pass's 8-point groups of
 * consecutive rows), dir 1 B = w_N^-r (the inverse pass that applies its
 * cross twiddles at the load, rows m0 + k*S/8). */
#define NTT_BTAB_DIRS 2
hipError_t launch_btab_gen(const TwPair *tw, TwPair *out, int logS, int logT, hipStream_t st);

struct GarnerConsts {
    uint64_t inv01, inv01p;      /* p0^-1 mod p1 */
    uint64_t p0m2, p0m2p;        /* p0 mod p2 */
    uint64_t inv012, inv012p;    /* (p0 p1)^-1 mod p2 */
};

/* Garner + carry over a size class: grid.y = product. */
struct GarnerArgs {
    const uint64_t *res;        /* res[prime][len], operand a's slot */
    size_t len;                 /* row stride in words (padded) */
    uint32_t pad_log;           /* residue index a lives at res_at(a, pad_log) */
    uint32_t logL;
    const MulDesc *desc;        /* or nullptr: use one */
    MulDesc one;
    uint32_t blk_per_prod;      /* carry blocks per product */
    uint8_t *gen;               /* u64 generate mask per 64-limb round (8-byte aligned) */
    uint8_t *blk;               /* block aggregates, then block carry-ins */
    int *flag;                  /* integrity flag: set on a corrupted product */
};

__global__ void k_twiddle_gen(TwPair *tw, uint64_t r2_0, uint64_t r2_1, uint64_t r2_2);
__global__ void k_garner_pack(GarnerArgs ga);
/* Block carry-ins from block (G,P) cells; cin0 is the carry into block 0. */
__global__ void k_carry_scan(uint8_t *blk, uint32_t nblk, uint8_t cin0);
__device__ __forceinline__ uint8_t cla_compose(uint8_t lo, uint8_t hi)
{
    uint8_t g = (hi & 1) | ((hi >> 1) & lo & 1);
    uint8_t p = (uint8_t)((hi >> 1) & (lo >> 1) & 1);
    return (uint8_t)(g | (p << 1));
}
/* Round cell (generate, propagate) of 64 limbs from their ballot masks. */
__device__ __forceinline__ uint8_t round_cell(uint64_t G, uint64_t P)
{
    uint64_t X = G | P, Y = G;
    uint64_t s = X + Y;
    uint8_t gen = s < X;
    uint8_t prop = (s == ~0ull) && !gen;
    return (uint8_t)(gen | (prop << 1));
}
__global__ void k_carry_apply(GarnerArgs ga);
/* Batched schoolbook: one wavefront per product, 4 products per block. */
#define SB_MAX_OUT 256
__global__ void k_schoolbook(const MulDesc *desc, MulDesc one, uint32_t nprod, int *flag);

#endif
