/*
 * ntt_fast.hip — the transform passes of the CRT-NTT multiply, register-blocked.
 *
 * A four-step (here three-factor) decomposition without transposes: pass i
 * runs S-point DIF sub-transforms down the columns of an S x T view of the
 * current block and multiplies by the cross twiddle w_N^(r*bitrev(m)); the
 * spectrum stays in that scrambled order, and the inverse runs the mirrored
 * DIT passes.  What matters on gfx950:
 *   - the S-point sub-transform runs in radix-8 rounds: each thread takes 8
 *     elements from LDS, does three butterfly stages in registers and writes
 *     them back, so a 512-point transform costs 3 LDS round trips and 3
 *     barriers instead of 9;
 *   - the S/2 stage twiddles (and their inverses) are staged once per block in
 *     LDS as Shoup pairs; the three smallest stages use compile-time twiddles;
 *   - the cross twiddles are applied inside the last forward (first inverse)
 *     radix round as a Montgomery chain, four table look-ups per 8 points;
 *     passes whose S x T block has at most 2^20 residues (the inner pass of
 *     a three-pass transform, the inverse of a two-pass one) instead read one
 *     precomputed Montgomery-form twiddle per residue (kinds 6 and 7);
 *   - the last forward pass of both operands, the pointwise product and the
 *     first inverse pass are one kernel (kind 2): both spectra of a tile are
 *     formed, multiplied and transformed back without leaving the chip (the
 *     1/L scale rides on the last inverse pass's cross twiddles when that
 *     pass applies them at the load, else on the pointwise product);
 *   - a first pass whose operands fill at most half the transform (every
 *     balanced product) skips the zero upper half of its columns: the top
 *     stage is a bare product of 16-bit inputs by their twiddles;
 *   - the forward column pass, the in-round inverse and the fused pass take
 *     their tiles into LDS by DMA (global_load_lds_dwordx4, no registers),
 *     writing the stage twiddles and looking up the cross-twiddle bases while
 *     the transfer is in flight (the first pass expands operand limbs, the
 *     load-chain inverse twiddles in registers: both load through VGPRs);
 *     block barriers wait for LDS work only, so stores stay in flight.
 * A tile is 4096 residues (32 KB of LDS) worked on by 512 threads: with 8 KB
 * of stage twiddles, three blocks share a CU (registers bound the fourth);
 * the fused middle pass keeps one twiddle table for both directions and, at
 * 64 VGPRs, four.
 * Rows of a size class are padded to a multiple of 4096 residues, and residue
 * index a lives at res_at(a, pad_log) (a pad after every first-pass row).
 *
 * Reference counterparts: ntt/lib/ntt_kernel.hip:625-675 (stockham_kernel),
 * :739-778 (stockham_kernel_had: Hadamard fused into the inverse's first pass),
 * :880-918 (xtranspose_fwd/inv: cross twiddle).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lazy_group.h"
#include "modarith.h"
#include "ntt_kernels.h"
#include "tile_geom.h"

using namespace nttc;

/* Shoup pairs w_4096^e, e < 2048, per prime: the stage twiddles of every
 * sub-transform of up to 4096 points (w_S^x = w_4096^(x*4096/S)). */
__global__ void __launch_bounds__(256) k_small_twiddles(const TwPair *tw, TwPair *small)
{
    /* w_4096^e = W^(e * 2^18) for the 2^30 root W: entry e*2^3 of the hi table */
    uint32_t e = blockIdx.x * 256 + threadIdx.x;
    uint32_t pi = blockIdx.y;
    if (e >= FT_SIZE / 2) return;
    small[pi * (FT_SIZE / 2) + e] = tw[(size_t)pi * 2 * NTT_TW_SIZE + NTT_TW_SIZE + (e << 3)];
}

/* LDS layout: the tile unpadded.  A pad word after every 8 residues removes
 * most bank conflicts of the strided rounds, but its 4 KB per block and its
 * address arithmetic cost occupancy (the fused middle pass drops from three
 * blocks per CU to two); measured, the unpadded layout is faster. */
#define LP(i) (i)
#define FT_LDS (FT_SIZE)

/* Block barrier for LDS traffic only: waits for this wave's LDS operations
 * (lgkmcnt) but not for its global ones, so global stores, and LDS DMA
 * transfers that a later explicit vmcnt wait covers, stay in flight across
 * it (a release fence would wait for those too: a DMA transfer is an LDS
 * write).  Threads of a pass block share data through LDS alone.  The
 * memory clobber keeps the compiler from moving memory accesses across. */
__device__ __forceinline__ void lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

/* Ordering for LDS data that only the calling wave wrote and reads: its LDS
 * operations complete in issue order, so waiting for them (and keeping the
 * compiler from moving memory accesses across) is enough; no block barrier. */
__device__ __forceinline__ void wave_lds_sync()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}
template <bool WL>
__device__ __forceinline__ void round_sync()
{
    if constexpr (WL) wave_lds_sync();
    else lds_sync();
}

/* XCD-aware tile order: consecutive blocks go to different XCDs (round robin
 * over 8), so tile numbers are dealt out in 8 contiguous ranges, one per XCD;
 * neighbouring column tiles then share their 128-byte lines in one L2. */
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n)
{
    uint32_t x = b & 7, i = b >> 3, q = n >> 3, rem = n & 7;
    return x * q + (x < rem ? x : rem) + i;
}

/* Tile order of the table-twiddle passes: tile t of the launch covers column
 * tile ct of group (gb*8 + gi) for t = (gb*colTiles + ct)*8 + gi, so the
 * tiles one XCD works on at a time share their twiddle columns across 8
 * groups (the table's slice stays in L2), and column neighbours are 8 tiles
 * apart (they still share their 128-byte lines).  Needs a multiple of 8
 * groups; otherwise the order is unchanged. */
__device__ __forceinline__ uint32_t xtab_tile(uint32_t t, uint32_t logCT, uint32_t ntiles)
{
    if (((ntiles >> logCT) & 7) != 0 || (ntiles & ((1u << logCT) - 1)) != 0) return t;
    const uint32_t gi = t & 7, rest = t >> 3;
    const uint32_t ct = rest & ((1u << logCT) - 1), gb = rest >> logCT;
    return (((gb << 3) | gi) << logCT) | ct;
}

__device__ __forceinline__ uint32_t brev_n(uint32_t x, int bits)
{
    return bits ? (__brev(x) >> (32 - bits)) : 0;
}

/* Residue k of thread tid's share of tile t (q = tid + k * FTPB) sits at
 * res_at(ft_addr(q)) = base + k * step: linear in k for every pass geometry,
 * pads included (the pad index a >> pad_log is constant across a tile or
 * grows with the row by one per row).  Two address computations per tile
 * instead of one per residue. */
template <int LOGS>
__device__ __forceinline__ void tile_span(const PassArgs &pa, const FGeom &g, uint32_t t, uint32_t tid, size_t &base,
                                          size_t &step)
{
    uint32_t m, r;
    base = res_at(ft_addr(LOGS, g, t, tid, m, r), pa.pad_log);
    step = res_at(ft_addr(LOGS, g, t, tid + FTPB, m, r), pa.pad_log) - base;
}

/* Group store x[gb + i*gs] = v[i], i < R, by a running pointer: one 64-bit
 * add per store.  gs is a per-lane value to the compiler, which otherwise
 * forms each i*gs with two 64-bit multiplies and two moves. */
template <int R>
__device__ __forceinline__ void group_store(uint64_t *x, size_t gb, size_t gs, const uint64_t *v)
{
    uint64_t *xp = x + gb;
#pragma unroll
    for (int i = 0; i < R; i++) {
        *xp = v[i];
        xp += gs;
        asm("" : "+v"(xp));     /* keeps the adds from being folded back into i*gs */
    }
}

/* Forward group of a compile-time round, inputs below 2p. */
template <int PI, int RL, int BOUT>
__device__ __forceinline__ void const_dif(uint64_t *x)
{
    lz_group<PI, RL, false, true, 2, BOUT>(x, ConstTwF<PI, RL, false>{});
}
/* Inverse group of a compile-time round. */
template <int PI, int RL, int BIN, int BOUT>
__device__ __forceinline__ void const_dit(uint64_t *x)
{
    lz_group<PI, RL, true, true, BIN, BOUT>(x, ConstTwF<PI, RL, true>{});
}

/* ---- cross twiddles in the last forward / first inverse round -------------
 * The four-step cross twiddle of row m, column r of a pass is
 * w_N^(r*bitrev_S(m)).  The smallest-stage round holds, per thread, the 2^RL
 * consecutive rows m0 + i (m0 a multiple of 2^RL) of one column, and
 * bitrev_S(m0 + i) = bitrev_S(m0) + bitrev_RL(i) * 2^(LOGS-RL), so its
 * twiddles are A * B^bitrev_RL(i) with A = w_N^(r*bitrev_S(m0)) and
 * B = w_N^(r*2^(LOGS-RL)): two table look-ups (lo and hi part each) and a
 * chain of Montgomery products instead of 2^(RL+1) random look-ups. */
struct XtwCtx {
    const TwPair *tw;     /* this prime's two-level table */
    uint32_t col0;        /* first column of the tile */
    uint32_t shift;       /* NTT_ROOT_LOG - log2(N): exponents of the 2^30 root */
    uint64_t ta, tb;      /* A, B of this thread's group (Montgomery form) when
                             the caller looked them up ahead (xt_ahead); with a
                             ratio table (kind 8), tb is B itself, in [0, p) */
    uint64_t tbp;         /* kind 8: B's Shoup partner */
    const TwPair *btab;   /* kind 8: this prime's ratios, B of column r at btab[r] */
};

/* When the cross-twiddle round has 8-point groups and one group per thread
 * (S = 2^LOGS with LOGS a multiple of 3, so its groups are FT_SIZE/8 =
 * FTPB), a thread's A and B are known at tile start: looking them up there
 * overlaps the table loads with the tile transfer instead of stalling the
 * round on them. */
template <int LOGS>
constexpr bool xt_ahead() { return LOGS % 3 == 0 && (FT_SIZE >> 3) == FTPB; }

/* w^E * 2^64 mod p (Montgomery form), in [0, 2p), for E < 2^30. */
template <int PI>
__device__ __forceinline__ uint64_t tw_mont(const TwPair *tw, uint32_t E)
{
    TwPair lo = tw[E & (NTT_TW_SIZE - 1)];
    TwPair hi = tw[NTT_TW_SIZE + (E >> NTT_TW_LOG)];
    /* lo.w * 2^64 mod p = -lo.wp * p mod 2^64 (the Shoup remainder) */
    uint64_t lo_m = (uint64_t)0 - lo.wp * Prime<PI>::P;
    return shoup_lazy<PI>(lo_m, hi.w, hi.wp);
}

/* Cross-twiddle table entry e = m*T + r of an S x T pass (N = S*T):
 * w_N^(r*bitrev_S(m)) in Montgomery form, w_N^-(r*bitrev_S(m)) for INV,
 * reduced to [0, p). */
template <int PI>
__device__ __forceinline__ uint64_t xtab_entry(const TwPair *tw, uint32_t logS, uint32_t logT, uint32_t e, bool inv)
{
    const uint32_t m = e >> logT, r = e & ((1u << logT) - 1);
    uint32_t E = (r * brev_n(m, logS)) << (NTT_ROOT_LOG - logS - logT);
    if (inv) E = 0u - E;
    return reduce1<PI>(tw_mont<PI>(tw + (size_t)PI * 2 * NTT_TW_SIZE, E & ((1u << NTT_ROOT_LOG) - 1)));
}

/* grid (N / 256, primes, NTT_XTAB_DIRS) */
__global__ void __launch_bounds__(256) k_xtab_gen(const TwPair *tw, uint64_t *out, uint32_t logS, uint32_t logT,
                                                  XtabScale sc)
{
    const uint32_t N = 1u << (logS + logT);
    const uint32_t e = blockIdx.x * 256 + threadIdx.x, pi = blockIdx.y;
    const bool inv = blockIdx.z != 0;
    if (e >= N) return;
    uint64_t v = pi == 0   ? xtab_entry<0>(tw, logS, logT, e, inv)
                 : pi == 1 ? xtab_entry<1>(tw, logS, logT, e, inv)
                           : xtab_entry<2>(tw, logS, logT, e, inv);
    if (blockIdx.z == 2)
        v = pi == 0   ? reduce1<0>(shoup_lazy<0>(v, sc.s[0], sc.sp[0]))
            : pi == 1 ? reduce1<1>(shoup_lazy<1>(v, sc.s[1], sc.sp[1]))
                      : reduce1<2>(shoup_lazy<2>(v, sc.s[2], sc.sp[2]));
    out[((size_t)blockIdx.z * NTT_NPRIMES + pi) * N + e] = v;
}

/* x[i] *= w^(EA + bitrev_RL(i) * EB) for i < 2^RL (exponents mod 2^30);
 * the Montgomery products leave x below 8p in (0, 2p) and any x in (0, 3p). */
template <int PI, int RL>
__device__ __forceinline__ void chain_apply(uint64_t *x, uint64_t t, const uint64_t b)
{
    constexpr int R = 1 << RL;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int i = RL == 3 ? ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1) : RL == 2 ? ((k & 1) << 1) | (k >> 1) : k;
        x[i] = mont_mul_lazy<PI>(x[i], t);
        if (k + 1 < R) t = mont_mul_lazy<PI>(t, b);
    }
}
/* The same with the ratio B in plain form with its Shoup partner (a ratio
 * table): the chain's steps are Shoup products (14 VALU instructions against
 * the Montgomery product's 24), which keep t in Montgomery form and below
 * 2p, as the products by t need. */
template <int PI, int RL>
__device__ __forceinline__ void chain_apply_tab(uint64_t *x, uint64_t t, const uint64_t b, const uint64_t bp)
{
    constexpr int R = 1 << RL;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int i = RL == 3 ? ((k & 1) << 2) | (k & 2) | ((k >> 2) & 1) : RL == 2 ? ((k & 1) << 1) | (k >> 1) : k;
        x[i] = mont_mul_lazy<PI>(x[i], t);
        if (k + 1 < R) t = shoup_lazy<PI>(t, b, bp);
    }
}
template <int PI, int RL>
__device__ __forceinline__ void chain_twiddle(uint64_t *x, const TwPair *tw, uint32_t EA, uint32_t EB)
{
    chain_apply<PI, RL>(x, tw_mont<PI>(tw, EA & ((1u << NTT_ROOT_LOG) - 1)),
                        tw_mont<PI>(tw, EB & ((1u << NTT_ROOT_LOG) - 1)));
}

/* A and B of this thread's group in the 8-point cross-twiddle round of tile
 * column col0 (xt_ahead): the group is thread tid's, rows m0 .. m0+7 with
 * m0 = ((tid >> logC) mod 2^(LOGS-3)) * 8, column r = col0 + (tid mod C);
 * INV gives the inverse twiddles. */
template <int PI, int LOGS, bool INV, bool BTAB = false>
__device__ __forceinline__ void xt_lookup(XtwCtx &xc, uint32_t logC)
{
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t r = xc.col0 + (tid & ((1u << logC) - 1));
    const uint32_t m0 = ((tid >> logC) & ((1u << (LOGS - 3)) - 1)) << 3;
    uint32_t ea = (r * brev_n(m0, LOGS)) << xc.shift, eb = (r << (LOGS - 3)) << xc.shift;
    if (INV) {
        ea = 0u - ea;
        eb = 0u - eb;
    }
    xc.ta = tw_mont<PI>(xc.tw, ea & ((1u << NTT_ROOT_LOG) - 1));
    if constexpr (BTAB) {
        static_assert(!INV, "ratio tables serve the first forward pass");
        const TwPair b = xc.btab[r];
        xc.tb = b.w;
        xc.tbp = b.wp;
    } else {
        xc.tb = tw_mont<PI>(xc.tw, eb & ((1u << NTT_ROOT_LOG) - 1));
    }
}

/* ---- stage twiddles of the fused passes: one table, both directions --------
 * The fused passes (kinds 2 and 4) run forward and inverse rounds over one
 * table.  Stage l (half-size hs = 2^l) of the rounds that read LDS (l >= RLL;
 * the smallest RLL stages use compile-time twiddles) keeps w_(2hs)^k for
 * k = 0 .. hs at [mid_off(l), mid_off(l) + hs]: the forward twiddle w^e at
 * +e, and the inverse twiddle w^-e = -w^(hs-e) as the entry at +(hs - e)
 * with the butterfly's outputs swapped (TabTwF::NEG; +hs holds w^hs = -1,
 * for e = 0).  510 entries for S = 512: with the 32 KB tile a block needs
 * 40 KB of LDS, and four blocks share a CU. */
template <int RLL>
__host__ __device__ constexpr uint32_t mid_off(int l)
{
    return ((1u << l) - (1u << RLL)) + (uint32_t)(l - RLL);
}

/* Stage twiddles of a round from an LDS table: the round's top stage has
 * half-size 2^LH and group spacing d = 2^LD; stage S of the group (half-size
 * 2^(LH-S)) pairs values I and I + half with e = j + (I mod half) * d.
 * MIDRLL > 0: the fused passes' shared table (mid_off<MIDRLL>); there the
 * inverse twiddle w^-e is read as the forward entry at hs - e, which holds
 * w^(hs-e) = -w^-e, so the inverse butterfly's outputs trade places (NEG). */
template <int LH, int LD, int R, bool DIT, int MIDRLL>
struct TabTwF {
    const TwPair *tw;
    uint32_t j;
    static constexpr bool NEG = DIT && MIDRLL != 0;
    template <int S, int I>
    __device__ __forceinline__ TwPair get() const
    {
        constexpr int half = R >> (S + 1);
        constexpr int lhs = LH - S;
        const uint32_t e = j + (uint32_t)((I & (half - 1)) << LD);
        if constexpr (MIDRLL != 0 && DIT) return tw[mid_off<MIDRLL ? MIDRLL : 1>(lhs) + (1u << lhs) - e];
        else if constexpr (MIDRLL != 0) return tw[mid_off<MIDRLL ? MIDRLL : 1>(lhs) + e];
        else return tw[(1u << lhs) - 1 + e];
    }
};

/* ---- radix-R rounds over the LDS tile -------------------------------------
 * A DIF round with top stage half-size h = 2^LH and R = 2^RL covers stages
 * h, h/2, .., h/2^(RL-1): the groups {m0 + i*d}, d = h >> (RL-1),
 * m0 = blk*2h + j, j < d. */

/* MIDRLL > 0: the fused passes' shared table (mid_off<MIDRLL>).  Values in
 * LDS are below 2p before and after the round. */
template <int PI, int LOGS, int LH, int RL, bool XT, int MIDRLL = 0>
__device__ __forceinline__ void dif_round(uint64_t *lds, const TwPair *twf, uint32_t logC, const XtwCtx &xc)
{
    constexpr int R = 1 << RL;
    constexpr int LD = LH - (RL - 1);
    constexpr uint32_t d = 1u << LD;
    constexpr int LNB = LOGS - LH - 1;
    for (uint32_t gid = threadIdx.x; gid < (FT_SIZE >> RL); gid += FTPB) {
        uint32_t c = gid & ((1u << logC) - 1);
        uint32_t u = gid >> logC;
        uint32_t j = u & (d - 1);
        uint32_t v = u >> LD;
        uint32_t blk = v & ((1u << LNB) - 1);
        uint32_t gg = v >> LNB;
        uint32_t m0 = (blk << (LH + 1)) + j;
        uint32_t base = (((gg << LOGS) + m0) << logC) + c;
        uint64_t x[R];
#pragma unroll
        for (int i = 0; i < R; i++) x[i] = lds[LP(base + ((uint32_t)(i << LD) << logC))];
        if constexpr (LD == 0 && XT) {
            /* the chain's Montgomery products take values below 8p to (0, 2p) */
            const_dif<PI, RL, 8>(x);
            if constexpr (xt_ahead<LOGS>()) {
                chain_apply<PI, RL>(x, xc.ta, xc.tb);
            } else {
                /* the outputs of rows m0 + i (m0 = blk << RL) of column col0 + c */
                const uint32_t r = xc.col0 + c;
                chain_twiddle<PI, RL>(x, xc.tw, (r * brev_n(m0, LOGS)) << xc.shift,
                                      (r << (LOGS - RL)) << xc.shift);
            }
        } else if constexpr (LD == 0) {
            const_dif<PI, RL, 2>(x);
        } else {
            lz_group<PI, RL, false, false, 2, 2>(x, TabTwF<LH, LD, R, false, MIDRLL>{twf, j});
        }
#pragma unroll
        for (int i = 0; i < R; i++) lds[LP(base + ((uint32_t)(i << LD) << logC))] = x[i];
    }
}

/* Values leave unreduced, below 16p (what reads inverse-pass output reduces
 * it in one step or multiplies it by a factor below 2p); BIN0: the
 * smallest round's (LD = 0) inputs are below BIN0*p, and every other
 * round's inputs below the bound the rounds under it leave (dit_bound).
 * GSTORE: the round's outputs go to global memory instead of back to LDS,
 * tile word q to gout[q] (a tile contiguous in its row: one column, T = 1). */
template <int PI, int LOGS, int LH, int RL, bool XT, int MIDRLL, int BIN0, bool GSTORE = false>
__device__ __forceinline__ void dit_round(uint64_t *lds, const TwPair *twi, uint32_t logC, const XtwCtx &xc,
                                          uint64_t *gout = nullptr)
{
    constexpr int R = 1 << RL;
    constexpr int LD = LH - (RL - 1);
    constexpr uint32_t d = 1u << LD;
    constexpr int LNB = LOGS - LH - 1;
    for (uint32_t gid = threadIdx.x; gid < (FT_SIZE >> RL); gid += FTPB) {
        uint32_t c = gid & ((1u << logC) - 1);
        uint32_t u = gid >> logC;
        uint32_t j = u & (d - 1);
        uint32_t v = u >> LD;
        uint32_t blk = v & ((1u << LNB) - 1);
        uint32_t gg = v >> LNB;
        uint32_t m0 = (blk << (LH + 1)) + j;
        uint32_t base = (((gg << LOGS) + m0) << logC) + c;
        uint64_t x[R];
#pragma unroll
        for (int i = 0; i < R; i++) x[i] = lds[LP(base + ((uint32_t)(i << LD) << logC))];
        if constexpr (LD == 0) {
            /* the chain's Montgomery products leave any input below 3p */
            if constexpr (XT && xt_ahead<LOGS>()) {
                chain_apply<PI, RL>(x, xc.ta, xc.tb);
            } else if constexpr (XT) {
                /* inputs of rows m0 + i of column col0 + c: inverse twiddles */
                const uint32_t r = xc.col0 + c;
                chain_twiddle<PI, RL>(x, xc.tw, 0u - ((r * brev_n(m0, LOGS)) << xc.shift),
                                      0u - ((r << (LOGS - RL)) << xc.shift));
            }
            const_dit<PI, RL, XT ? 4 : BIN0, LZ_MAX>(x);
        } else {
            lz_group<PI, RL, true, false, dit_bound(LH - RL, BIN0), LZ_MAX>(x, TabTwF<LH, LD, R, true, MIDRLL>{twi, j});
        }
        if constexpr (GSTORE) {
#pragma unroll
            for (int i = 0; i < R; i++) gout[base + ((uint32_t)(i << LD) << logC)] = x[i];
        } else {
#pragma unroll
            for (int i = 0; i < R; i++) lds[LP(base + ((uint32_t)(i << LD) << logC))] = x[i];
        }
    }
}

/* The first DIF round of a first pass whose rows S/2.. are zero (operands
 * of at most L/2 coefficients, LOGS >= 4): the top stage pairs each 16-bit
 * input with a zero, so it keeps x and forms x*w (a small-operand Shoup
 * product, no sum or difference); the round's other two stages as usual.
 * The zero rows are not read. */
template <int PI, int LOGS>
__device__ __forceinline__ void dif_round_hz(uint64_t *lds, const TwPair *twf, uint32_t logC)
{
    constexpr int LH = LOGS - 1, RL = 3, R = 8, LD = LH - (RL - 1);
    constexpr uint32_t d = 1u << LD;
    static_assert(LOGS >= 4, "the first round is not the last");
    for (uint32_t gid = threadIdx.x; gid < (FT_SIZE >> RL); gid += FTPB) {
        const uint32_t c = gid & ((1u << logC) - 1);
        const uint32_t u = gid >> logC;
        const uint32_t j = u & (d - 1);
        const uint32_t gg = u >> LD;                    /* one block per group: the top stage spans S */
        const uint32_t base = (((gg << LOGS) + j) << logC) + c;
        uint64_t x[R];
#pragma unroll
        for (int i = 0; i < R / 2; i++) x[i] = lds[LP(base + ((uint32_t)(i << LD) << logC))];
#pragma unroll
        for (int i = 0; i < R / 2; i++) {
            const TwPair w = twf[(1u << LH) - 1 + j + (uint32_t)(i << LD)];
            x[i + R / 2] = shoup_small<PI>((uint32_t)x[i], w.w, w.wp);
        }
        /* the other two stages: every value below 2p now */
        lz_group<PI, RL, false, false, 2, 2, 1>(x, TabTwF<LH, LD, R, false, 0>{twf, j});
#pragma unroll
        for (int i = 0; i < R; i++) lds[LP(base + ((uint32_t)(i << LD) << logC))] = x[i];
    }
}

/* DIF rounds of an S = 2^LOGS transform from half-size 2^LH down, except the
 * smallest (LD = 0) round. */
template <int PI, int LOGS, int LH>
__device__ __forceinline__ void dif_above_last(uint64_t *lds, const TwPair *twf, uint32_t logC)
{
    constexpr int RL = LH + 1 >= 3 ? 3 : LH + 1;
    if constexpr (LH - RL >= 0) {
        dif_round<PI, LOGS, LH, RL, false>(lds, twf, logC, XtwCtx{});
        lds_sync();
        dif_above_last<PI, LOGS, LH - RL>(lds, twf, logC);
    }
}

/* Thread tid's group n of the smallest DIF round (groups of 2^last_rl(LOGS)
 * rows, tile_geom.h): rows m0 .. m0 + 2^RL - 1 of tile column c, at tile
 * words q0 + i*C; the same rows go to gb + i*gs in the row. */
template <int LOGS>
__device__ __forceinline__ uint32_t last_group(const PassArgs &pa, const FGeom &g, uint32_t tt, uint32_t tid, int n,
                                               uint32_t &c, uint32_t &m0, size_t &gb, size_t &gs)
{
    const uint32_t q0 = last_group_q0<LOGS>(g.logC, tid + (uint32_t)n * FTPB, c, m0);
    uint32_t m, r;
    const size_t a0 = ft_addr(LOGS, g, tt, q0, m, r);
    gb = res_at(a0, pa.pad_log);
    gs = res_at(a0 + ((size_t)1 << g.logT), pa.pad_log) - gb;   /* row m0 + 1: 2^logT residues on */
    return q0;
}

/* The smallest DIF round with the cross-twiddle chain (for an 8-point group
 * and one group per thread, A and B looked up at tile start; otherwise per
 * group), stored from registers: no LDS round trip and no block barrier
 * before the tile leaves (measured on the first pass of a 2^30-bit product:
 * 3.89 -> 3.43 ms, and 78 -> 70 VGPRs). */
template <int PI, int LOGS, bool BTAB = false>
__device__ __forceinline__ void dif_last_store(const PassArgs &pa, const FGeom &g, uint32_t tt, const uint64_t *lds,
                                               uint64_t *x, const XtwCtx &xc)
{
    static_assert(!BTAB || (last_rl<LOGS>() == 3 && xt_ahead<LOGS>()), "ratio tables: 8-point groups, B ahead");
    constexpr int RL = last_rl<LOGS>(), R = 1 << RL, NG = 8 >> RL;
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int n = 0; n < NG; n++) {
        uint32_t c, m0;
        size_t gb, gs;
        const uint32_t q0 = last_group<LOGS>(pa, g, tt, tid, n, c, m0, gb, gs);
        uint64_t v[R];
#pragma unroll
        for (int i = 0; i < R; i++) v[i] = lds[LP(q0 + ((uint32_t)i << g.logC))];
        const_dif<PI, RL, 8>(v);     /* the chain takes values below 8p to (0, 2p) */
        if constexpr (BTAB) {
            chain_apply_tab<PI, 3>(v, xc.ta, xc.tb, xc.tbp);
        } else if constexpr (RL == 3 && xt_ahead<LOGS>()) {
            chain_apply<PI, 3>(v, xc.ta, xc.tb);
        } else {
            const uint32_t r = xc.col0 + c;
            chain_twiddle<PI, RL>(v, xc.tw, (r * brev_n(m0, LOGS)) << xc.shift, (r << (LOGS - RL)) << xc.shift);
        }
        group_store<R>(x, gb, gs, v);
    }
}

/* The top DIT round of a C * S = 4096 tile (radix 8 for every S >= 8),
 * stored from registers: thread tid's group is tile words tid + k*FTPB,
 * which tile_span places at gb + k*gs. */
template <int PI, int LOGS, int BIN>
__device__ __forceinline__ void dit_top_finish(uint64_t *v, const TwPair *twi, uint32_t logC, uint64_t *x, size_t gb,
                                               size_t gs)
{
    constexpr int LH = LOGS - 1, RL = 3, R = 8, LD = LH - (RL - 1);
    const uint32_t j = threadIdx.x >> logC;    /* the group's first row, below S/8 = 2^LD */
    /* inputs below BIN*p, stored unreduced, below 16p (inverse-pass output) */
    lz_group<PI, RL, true, false, BIN, LZ_MAX>(v, TabTwF<LH, LD, R, true, 0>{twi, j});
    group_store<R>(x, gb, gs, v);
}

/* All DIF rounds of an S = 2^LOGS transform, top stage first, radix 8 while
 * three or more stages remain. */
template <int PI, int LOGS, int LH, bool XT = false>
__device__ __forceinline__ void dif_all(uint64_t *lds, const TwPair *twf, uint32_t logC, const XtwCtx &xc = {})
{
    if constexpr (LH >= 0) {
        constexpr int RL = LH + 1 >= 3 ? 3 : LH + 1;
        dif_round<PI, LOGS, LH, RL, XT, 0>(lds, twf, logC, xc);
        lds_sync();
        dif_all<PI, LOGS, LH - RL, XT>(lds, twf, logC, xc);
    }
}

/* DIT rounds in the mirror order (lowest stages first). */
template <int PI, int LOGS, int LH, bool XT = false, bool REDUCEDIN = false>
__device__ __forceinline__ void dit_all(uint64_t *lds, const TwPair *twi, uint32_t logC, const XtwCtx &xc = {})
{
    if constexpr (LH >= 0) {
        constexpr int RL = LH + 1 >= 3 ? 3 : LH + 1;
        dit_all<PI, LOGS, LH - RL, XT, REDUCEDIN>(lds, twi, logC, xc);
        dit_round<PI, LOGS, LH, RL, XT, 0, REDUCEDIN && !XT ? 2 : 4>(lds, twi, logC, xc);
        lds_sync();
    }
}

/* Stage twiddles of an S-point transform into LDS, one contiguous run per
 * stage so a round's lanes read consecutive entries: the stage with
 * half-size hs = 2^l keeps w_(2hs)^e, e < hs, at [hs - 1, 2hs - 1) of twf and
 * w_(2hs)^-e at the same place in twi (S - 1 entries each). */
template <int PI, int LOGS, int DIRS>   /* DIRS: 1 forward, 2 inverse, 3 both */
__device__ __forceinline__ void stage_twiddles(const TwPair *small_all, TwPair *twf, TwPair *twi)
{
    const TwPair *sm = small_all + PI * (FT_SIZE / 2);
    constexpr uint32_t n = (1u << LOGS) - 1;
    for (uint32_t idx = threadIdx.x; idx < n; idx += FTPB) {
        uint32_t l = 31 - __clz(idx + 1);            /* stage: hs = 2^l */
        uint32_t e = idx + 1 - (1u << l);            /* e < hs */
        /* w_(2hs)^e = w_4096^(e * 4096 / (2hs)) */
        uint32_t sh = FT_LOG - 1 - l;
        if (DIRS & 1) twf[idx] = sm[e << sh];
        if (!(DIRS & 2)) continue;
        if (e == 0) {
            twi[idx] = sm[0];
        } else {
            /* w^-e = -w^(hs - e) for a root of order 2hs; the Shoup partner of
             * p - w is 2^64 - 1 - partner(w) (w != 0) */
            TwPair t = sm[((1u << l) - e) << sh];
            twi[idx].w = Prime<PI>::P - t.w;
            twi[idx].wp = ~t.wp;
        }
    }
}

/* ------------------------------------------------------------------------ */

/* Value of forward-pass input element `a` (global index within the row). */
__device__ __forceinline__ uint64_t first_in(const PassArgs &pa, int op, size_t a)
{
    size_t prod = a >> pa.logL;
    if (prod >= pa.nprod) return 0;
    const MulDesc &d = pa.desc ? pa.desc[prod] : pa.one;
    const uint16_t *s = (const uint16_t *)(op ? d.b : d.a);
    size_t n16 = 4 * (size_t)(op ? d.nb : d.na);
    size_t loc = a & (((size_t)1 << pa.logL) - 1);
    return loc < n16 ? (uint64_t)s[loc] : 0;
}

/* Limb of forward-pass input elements a .. a+3 (a a multiple of 4): the
 * operand's u64 limb a/4 of its product, or 0 past the operand. */
__device__ __forceinline__ uint64_t first_limb(const PassArgs &pa, int op, size_t a)
{
    size_t prod = a >> pa.logL;
    if (prod >= pa.nprod) return 0;
    const MulDesc &d = pa.desc ? pa.desc[prod] : pa.one;
    const uint64_t *s = op ? d.b : d.a;
    size_t n = op ? d.nb : d.na;
    size_t j = (a & (((size_t)1 << pa.logL) - 1)) >> 2;
    return j < n ? s[j] : 0;
}

/* The first pass's tile straight from the operands into LDS.  With four or
 * more columns per tile, four consecutive columns of a row are one operand
 * limb: one 8-byte load gives four coefficients. */
template <int LOGS>
__device__ __forceinline__ void fill_first_tile(const PassArgs &pa, int op, const FGeom &g, uint32_t t, uint64_t *lds)
{
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (g.logC >= 2) {
        /* with one group per tile (C * S = 4096), limb share k covers tile
         * elements [2048k, 2048(k+1)), rows [k S/2, (k+1) S/2): the second
         * share is rows S/2.., which the half-zero first round (LOGS >= 4)
         * never reads */
        constexpr int NK = FT_PER / 4;
        static_assert(NK == 2, "two limb shares per thread");
        const int nk = (LOGS >= 4 && pa.half_zero && g.logG == 0) ? NK / 2 : NK;
        uint64_t limb[NK];
#pragma unroll
        for (int k = 0; k < NK; k++) {
            uint32_t q = (tid + k * FTPB) << 2, m, r;
            size_t a = ft_addr(LOGS, g, t, q, m, r);
            limb[k] = k < nk ? first_limb(pa, op, a) : 0;
        }
#pragma unroll
        for (int k = 0; k < NK; k++) {
            if (k >= nk) break;
            uint32_t q = (tid + k * FTPB) << 2;
#pragma unroll
            for (int i = 0; i < 4; i++) lds[LP(q + i)] = (limb[k] >> (16 * i)) & 0xffff;
        }
    } else {
        uint64_t v[FT_PER];
#pragma unroll
        for (int k = 0; k < FT_PER; k++) {
            uint32_t q = tid + k * FTPB, m, r;
            size_t a = ft_addr(LOGS, g, t, q, m, r);
            v[k] = first_in(pa, op, a);
        }
#pragma unroll
        for (int k = 0; k < FT_PER; k++) lds[LP(tid + k * FTPB)] = v[k];
    }
}

/* Asynchronous copy of tile t of row x straight into LDS (no registers):
 * each wave moves FT_SIZE/FTPB KB as 1 KB global_load_lds_dwordx4
 * transfers, lane l carrying LDS words q, q+1 (q = chunk*128 + 2l) from
 * their global places.  Those two are adjacent there too: both columns of
 * one row when C >= 2, consecutive rows of the one column when T = 1 (C = 1
 * otherwise never occurs; the launcher checks).  The caller waits (vmcnt)
 * and syncs before reading the tile. */
typedef __attribute__((address_space(3))) void lds_void_t;
/* SPLIT: word q0 + 2*lane of chunk q0 (a multiple of 128; 2*lane below 128)
 * is addressed as ft_addr(t, q0) + ft_addr(0, 2*lane).  ft_addr's fields c,
 * m and gg are bit ranges of q and the two parts occupy disjoint bits, so
 * the sum is ft_addr(t, q0 + 2*lane).  The first part is wave-uniform
 * (scalar registers: t is the block's tile), the second a per-lane offset:
 * about 4 VALU instructions per transfer instead of 17 (forward table pass
 * 3.24 -> 3.19 ms).  The chain pass with the ratio table measured slower
 * with it (1.81 -> 1.88 ms) and takes the per-lane form (!SPLIT). */
template <int LOGS, bool SPLIT = true>
__device__ __forceinline__ void tile_dma(const PassArgs &pa, const uint64_t *x, const FGeom &g, uint32_t t,
                                         uint64_t *lds)
{
    constexpr int CHUNKS = FT_SIZE / 128 / (FTPB / 64);    /* 1 KB chunks per wave */
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));     /* address math per call, not hoisted out of tile loops */
    const uint32_t w = SPLIT ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
    uint32_t m, r;
    const size_t lo = SPLIT ? ft_addr(LOGS, g, 0, 2 * (tid & 63), m, r) : 0;
#pragma unroll
    for (int j = 0; j < CHUNKS; j++) {
        const uint32_t q0 = (w * CHUNKS + j) * 128;
        const size_t a = SPLIT ? res_at(ft_addr(LOGS, g, t, q0, m, r) + lo, pa.pad_log)
                               : res_at(ft_addr(LOGS, g, t, q0 + 2 * (tid & 63), m, r), pa.pad_log);
        __builtin_amdgcn_global_load_lds((const void *)(x + a), (lds_void_t *)(lds + LP(q0)), 16, 0, 0);
    }
}
/* transfers tile_dma leaves in flight per wave */
#define TILE_DMA_OPS (FT_SIZE / 128 / (FTPB / 64))
/* s_waitcnt vmcnt(N) (gfx9 encoding: vmcnt low bits [3:0], high bits [15:14];
 * expcnt and lgkmcnt left at their maxima) */
#define WAIT_VMCNT(N) __builtin_amdgcn_s_waitcnt(0x0F70 | ((N) & 15) | (((N) >> 4) << 14))


/* XTAB: the cross twiddles come from pa.xtab and multiply the tile as it is
 * stored (element q = tid + k*FTPB sits in row (q >> logC) mod S, column
 * col0 + (q mod C)): one Montgomery product per residue instead of the
 * in-round chain's two, for passes whose S x T block is at most
 * 2^NTT_XTAB_MAX_LOG residues (its table stays in the L2 caches). */
/* BTAB (the first pass, 8-point last-round groups): the chain's ratio B
 * comes from pa.btab with its Shoup partner. */
template <int PI, int LOGS, bool FIRST, bool XTAB, bool BTAB = false>
__device__ void fwd_tiles(const PassArgs &pa, uint64_t *x, int op, uint64_t *lds, TwPair *twf)
{
    static_assert(!BTAB || (FIRST && LOGS >= 4 && xt_ahead<LOGS>()), "ratio tables: first pass, B ahead");
    const FGeom g{pa.logT, pa.logC, pa.logG};
    /* the table pass takes tiles nblk apart (as inv_tiles_dma: neighbouring
     * column tiles run side by side), the others consecutive ones */
    constexpr bool STRIDED = XTAB && !FIRST && LOGS >= 3;   /* the loop that takes them */
    uint32_t t = STRIDED ? xcd_tile(blockIdx.x, gridDim.x) : xcd_tile(blockIdx.x, gridDim.x) * pa.tpb;
    const uint32_t tend = STRIDED ? pa.ntiles : min(t + pa.tpb, pa.ntiles);
    const uint32_t colTiles = g.logT - g.logC;
    uint32_t tt = XTAB ? xtab_tile(t, colTiles, pa.ntiles) : t;   /* the tile worked on */
    XtwCtx xc{pa.tw + (size_t)PI * 2 * NTT_TW_SIZE, 0, NTT_ROOT_LOG - (LOGS + g.logT), 0, 0, 0,
              BTAB ? pa.btab + ((size_t)PI << g.logT) : nullptr};
    /* the first tile and the cross-twiddle look-ups are in flight while
     * the stage twiddles are written */
    stage_twiddles<PI, LOGS, 1>(pa.small_tw, twf, nullptr);
    if constexpr (!FIRST) tile_dma<LOGS>(pa, x, g, tt, lds);
    xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
    if constexpr (xt_ahead<LOGS>() && !XTAB) xt_lookup<PI, LOGS, false, BTAB>(xc, g.logC);
    if constexpr (!FIRST) WAIT_VMCNT(0);
    lds_sync();
    if constexpr (XTAB && !FIRST && LOGS >= 3) {
        /* The smallest round stores from registers, and the next tile's
         * transfer starts once every thread has read its groups, so it is in
         * flight during that round, the table products and the stores
         * (3.51 -> 3.43 ms at L = 2^27, two tiles per block).  The table
         * entries of a group are rows m0 .. m0 + 2^RL - 1 of its column. */
        constexpr int RL = last_rl<LOGS>(), R = 1 << RL, NG = 8 >> RL;
        const uint64_t *xt = pa.xtab + ((size_t)PI << (LOGS + g.logT));
        for (;;) {
            dif_above_last<PI, LOGS, LOGS - 1>(lds, twf, g.logC);
            uint32_t tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            /* the groups and their table entries are read before the next
             * tile's transfer is issued */
            uint64_t v[8], w[8];
            size_t gb0 = 0, gs0 = 0;
#pragma unroll
            for (int n = 0; n < NG; n++) {
                uint32_t c, m0;
                size_t gb, gs;
                const uint32_t q0 = last_group<LOGS>(pa, g, tt, tid, n, c, m0, gb, gs);
                if (n == 0) {
                    gb0 = gb;
                    gs0 = gs;
                }
#pragma unroll
                for (int i = 0; i < R; i++) v[n * R + i] = lds[LP(q0 + ((uint32_t)i << g.logC))];
#pragma unroll
                for (int i = 0; i < R; i++)
                    w[n * R + i] = xt[((size_t)m0 << g.logT) + xc.col0 + c + i * ((size_t)1 << g.logT)];
            }
            const uint32_t tcur = tt;
            const uint32_t tnext = t + (STRIDED ? gridDim.x : 1);
            const bool more = tnext < tend;
            if (more) {
                t = tnext;
                tt = xtab_tile(t, colTiles, pa.ntiles);
                lds_sync();       /* every group read: the tile buffer is free */
                tile_dma<LOGS>(pa, x, g, tt, lds);
            }
#pragma unroll
            for (int n = 0; n < NG; n++) {
                size_t gb = gb0, gs = gs0;
                if (n > 0) {
                    uint32_t c, m0;
                    (void)last_group<LOGS>(pa, g, tcur, tid, n, c, m0, gb, gs);
                }
                /* unreduced (below 16p): the Montgomery product by a table
                 * entry below p returns (0, 2p) */
                const_dif<PI, RL, LZ_MAX>(v + n * R);
#pragma unroll
                for (int i = 0; i < R; i++) v[n * R + i] = mont_mul_lazy<PI>(v[n * R + i], w[n * R + i]);
                group_store<R>(x, gb, gs, v + n * R);
            }
            if (!more) return;
            xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
            WAIT_VMCNT(0);
            lds_sync();
        }
    }
    /* the loop below serves the first pass and the chain passes (the table
     * passes, S >= 8, returned above); the first pass stores its smallest
     * round from registers */
    constexpr bool REGSTORE = FIRST && LOGS >= 4;
    for (;;) {
        uint32_t tid;
        if constexpr (FIRST) {
            fill_first_tile<LOGS>(pa, op, g, tt, lds);
            lds_sync();
        }
        if constexpr (FIRST && LOGS >= 4) {
            /* only the first round differs: the rounds after it are shared
             * code (a second copy of them costs more in instruction fetch
             * than the cheaper round saves) */
            if (pa.half_zero) dif_round_hz<PI, LOGS>(lds, twf, g.logC);   /* a bare twiddle product on top */
            else dif_round<PI, LOGS, LOGS - 1, 3, false>(lds, twf, g.logC, xc);
            lds_sync();
            if constexpr (REGSTORE) {
                dif_above_last<PI, LOGS, LOGS - 4>(lds, twf, g.logC);
                dif_last_store<PI, LOGS, BTAB>(pa, g, tt, lds, x, xc);
            } else {
                dif_all<PI, LOGS, LOGS - 4, true>(lds, twf, g.logC, xc);
            }
        } else {
            /* cross twiddles by chain in the last round */
            dif_all<PI, LOGS, LOGS - 1, true>(lds, twf, g.logC, xc);
        }
        tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        if constexpr (!REGSTORE) {
            size_t base, step;
            tile_span<LOGS>(pa, g, tt, tid, base, step);
#pragma unroll
            for (int k = 0; k < FT_PER; k++) x[base + k * step] = lds[LP(tid + k * FTPB)];
        }
        if (++t >= tend) break;
        tt = t;
        lds_sync();
        if constexpr (!FIRST) tile_dma<LOGS>(pa, x, g, tt, lds);
        xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
        if constexpr (xt_ahead<LOGS>() && !XTAB) xt_lookup<PI, LOGS, false, BTAB>(xc, g.logC);
        if constexpr (!FIRST) {
            WAIT_VMCNT(0);
            lds_sync();
        }
    }
}

/* The inverse passes that apply their cross twiddles at the load (kinds 5, 7
 * and 9; C * S = 4096).  Thread tid's tile words tid + k*FTPB are rows
 * m0 + k*2^(LOGS-3), k < 8, of column c (m0 = tid >> logC < 2^(LOGS-3)), and
 * bitrev(m0 + k*2^(LOGS-3)) = bitrev(m0) + bitrev3(k), so their cross
 * twiddles are one table entry each (kind 7) or a chain A*B^bitrev3(k)
 * (kinds 5 and 9), applied in LDS before the rounds, which then carry no
 * twiddle code.  Tiles reach LDS by DMA, and the next tile's transfer
 * starts as soon as the top round has read its groups, so it is in flight
 * during that round's butterflies and stores (the top round stores from
 * registers).  A block takes tiles j, j + nblk, j + 2 nblk, ... (j its
 * XCD-ordered index): neighbouring column tiles, which share 128-byte lines,
 * still run side by side on one XCD. */
template <int PI, int LOGS, bool XTAB, bool BTAB>
__device__ void inv_tiles_dma(const PassArgs &pa, uint64_t *x, uint64_t *lds, TwPair *twi)
{
    static_assert(!(XTAB && BTAB), "one source of cross twiddles");
    const FGeom g{pa.logT, pa.logC, pa.logG};
    const uint32_t nblk = gridDim.x;
    uint32_t t = xcd_tile(blockIdx.x, nblk);
    const uint32_t colTiles = g.logT - g.logC;
    uint32_t tt = XTAB ? xtab_tile(t, colTiles, pa.ntiles) : t;   /* the tile worked on */
    XtwCtx xc{pa.tw + (size_t)PI * 2 * NTT_TW_SIZE, 0, NTT_ROOT_LOG - (LOGS + g.logT), 0, 0};
    /* the first tile is in flight while the stage twiddles are written */
    tile_dma<LOGS, !BTAB>(pa, x, g, tt, lds);
    stage_twiddles<PI, LOGS, 2>(pa.small_tw, nullptr, twi);
    for (;;) {
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));     /* recompute index math per tile, not hoisted */
        xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
        const uint32_t r = xc.col0 + (tid & ((1u << g.logC) - 1)), m0 = tid >> g.logC;
        uint64_t v[FT_PER];
        if constexpr (XTAB) {
            /* one table entry per word, loaded while the tile arrives */
            const uint64_t *xt = pa.xtab + ((size_t)PI << (LOGS + g.logT)) + ((size_t)m0 << g.logT) + r;
            uint64_t w[FT_PER];
#pragma unroll
            for (int k = 0; k < FT_PER; k++) w[k] = xt[k * ((size_t)1 << (LOGS - 3 + g.logT))];
            WAIT_VMCNT(0);
            lds_sync();
#pragma unroll
            for (int k = 0; k < FT_PER; k++) v[k] = mont_mul_lazy<PI>(lds[LP(tid + k * FTPB)], w[k]);
        } else {
            constexpr uint32_t MASK = (1u << NTT_ROOT_LOG) - 1;
            uint64_t ta = tw_mont<PI>(xc.tw, (0u - ((r * brev_n(m0, LOGS)) << xc.shift)) & MASK);
            if (pa.iscale) ta = mont_mul_lazy<PI>(ta, pa.scale_m[PI]);   /* the 1/L rides on the chain */
            if constexpr (BTAB) {
                /* B = w_N^-r and its Shoup partner from the ratio table */
                const TwPair b = pa.btab[((size_t)PI << g.logT) + r];
                WAIT_VMCNT(0);
                lds_sync();
#pragma unroll
                for (int k = 0; k < FT_PER; k++) v[k] = lds[LP(tid + k * FTPB)];
                chain_apply_tab<PI, 3>(v, ta, b.w, b.wp);
            } else {
                const uint64_t tb = tw_mont<PI>(xc.tw, (0u - (r << xc.shift)) & MASK);
                WAIT_VMCNT(0);
                lds_sync();
#pragma unroll
                for (int k = 0; k < FT_PER; k++) v[k] = lds[LP(tid + k * FTPB)];
                chain_apply<PI, 3>(v, ta, tb);
            }
        }
        /* each thread rewrites only the words it read */
#pragma unroll
        for (int k = 0; k < FT_PER; k++) lds[LP(tid + k * FTPB)] = v[k];
        lds_sync();
        /* the table's Montgomery products (an entry below p) leave the tile
         * below 2p, the chain's below 3p */
        dit_all<PI, LOGS, LOGS - 4, false, XTAB>(lds, twi, g.logC, xc);
        tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        size_t base, step;
        tile_span<LOGS>(pa, g, tt, tid, base, step);
#pragma unroll
        for (int i = 0; i < FT_PER; i++) v[i] = lds[LP(tid + i * FTPB)];
        t += nblk;
        const bool more = t < pa.ntiles;
        if (more) {
            tt = XTAB ? xtab_tile(t, colTiles, pa.ntiles) : t;
            lds_sync();           /* every group read: the tile buffer is free */
            tile_dma<LOGS, !BTAB>(pa, x, g, tt, lds);
        }
        dit_top_finish<PI, LOGS, dit_bound(LOGS - 4, XTAB ? 2 : 4)>(v, twi, g.logC, x, base, step);
        if (!more) return;
    }
}

/* Inverse passes without load-time cross twiddles (kind 1): the tile goes
 * straight into LDS (in flight while the stage twiddles are written) and
 * the inverse cross twiddles go into the first round, by chain. */
template <int PI, int LOGS>
__device__ void inv_tiles(const PassArgs &pa, uint64_t *x, uint64_t *lds, TwPair *twi)
{
    const FGeom g{pa.logT, pa.logC, pa.logG};
    uint32_t t = xcd_tile(blockIdx.x, gridDim.x) * pa.tpb;
    const uint32_t tend = min(t + pa.tpb, pa.ntiles);
    const uint32_t colTiles = g.logT - g.logC;
    uint32_t tt = t;
    XtwCtx xc{pa.tw + (size_t)PI * 2 * NTT_TW_SIZE, 0, NTT_ROOT_LOG - (LOGS + g.logT), 0, 0};
    stage_twiddles<PI, LOGS, 2>(pa.small_tw, nullptr, twi);
    xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
    tile_dma<LOGS>(pa, x, g, tt, lds);
    if constexpr (xt_ahead<LOGS>()) xt_lookup<PI, LOGS, true>(xc, g.logC);
    WAIT_VMCNT(0);
    lds_sync();
    for (;;) {
        dit_all<PI, LOGS, LOGS - 1, true, false>(lds, twi, g.logC, xc);
        uint32_t tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        size_t base, step;
        tile_span<LOGS>(pa, g, tt, tid, base, step);
#pragma unroll
        for (int k = 0; k < FT_PER; k++) x[base + k * step] = lds[LP(tid + k * FTPB)];
        if (++t >= tend) break;
        tt = t;
        lds_sync();
        xc.col0 = (tt & ((1u << colTiles) - 1)) << g.logC;
        tile_dma<LOGS>(pa, x, g, tt, lds);
        if constexpr (xt_ahead<LOGS>()) xt_lookup<PI, LOGS, true>(xc, g.logC);
        WAIT_VMCNT(0);
        lds_sync();
    }
}

/* DIF rounds of an S = 2^LOGS transform except the last (smallest) one, and
 * DIT rounds except the first: the fused middle pass does that round itself.
 * WL: the rounds are wave-local (see mid_tile) and need no block barrier. */
template <int PI, int LOGS, int LH, int RLL, bool WL = false>
__device__ __forceinline__ void dif_upper(uint64_t *lds, const TwPair *tw)
{
    constexpr int RL = LH + 1 >= 3 ? 3 : LH + 1;
    if constexpr (LH - RL >= 0) {
        dif_round<PI, LOGS, LH, RL, false, RLL>(lds, tw, 0, XtwCtx{});
        round_sync<WL>();
        dif_upper<PI, LOGS, LH - RL, RLL, WL>(lds, tw);
    }
}
/* GSTORE: the last (top) round writes the tile to gout (see dit_round). */
template <int PI, int LOGS, int LH, int RLL, bool WL = false, bool GSTORE = false>
__device__ __forceinline__ void dit_upper(uint64_t *lds, const TwPair *tw, uint64_t *gout = nullptr)
{
    constexpr int RL = LH + 1 >= 3 ? 3 : LH + 1;
    if constexpr (LH - RL >= 0) {
        dit_upper<PI, LOGS, LH - RL, RLL, WL, false>(lds, tw);
        dit_round<PI, LOGS, LH, RL, false, RLL, 4, GSTORE>(lds, tw, 0, XtwCtx{}, gout);
        if constexpr (!GSTORE) round_sync<WL>();
    }
}

/* The fused passes' table (layout above): entry k of stage l is
 * w_(2hs)^k = w_4096^(k * 4096 / (2hs)) for k < hs, and -1 for k = hs. */
template <int PI, int LOGS, int RLL>
__device__ __forceinline__ void stage_twiddles_mid(const TwPair *small_all, TwPair *tw)
{
    const TwPair *sm = small_all + PI * (FT_SIZE / 2);
    constexpr uint32_t n = mid_off<RLL>(LOGS);
    for (uint32_t idx = threadIdx.x; idx < n; idx += FTPB) {
        int l = RLL;
        while (l + 1 < LOGS && idx >= mid_off<RLL>(l + 1)) l++;
        const uint32_t k = idx - mid_off<RLL>(l);
        if (k < (1u << l)) {
            tw[idx] = sm[k << (FT_LOG - 1 - l)];
        } else {       /* w^hs = -1: p - 1, Shoup partner 2^64 - 1 - partner(1) */
            const TwPair one = sm[0];
            tw[idx].w = Prime<PI>::P - one.w;
            tw[idx].wp = ~one.wp;
        }
    }
}

/* Last forward pass of a and b, pointwise product (x 2^64/L via Montgomery and
 * a Shoup scale), first inverse pass; result into a's row.  T = 1 here.
 * The smallest DIF round and the smallest DIT round work on the same groups
 * of 2^RLL consecutive points, so a's spectrum is kept in registers in that
 * round's layout, and b's last DIF round, the product and the first DIT
 * round run on the same registers without an LDS round trip. */
template <int PI, int LOGS, bool FIRST>
__device__ void mid_tile(const PassArgs &pa, uint64_t *xa, const uint64_t *xb, uint64_t *lds, TwPair *tw)
{
    constexpr int RLL = ((LOGS - 1) % 3) + 1;       /* the smallest round: 2^RLL points per group */
    constexpr int RR = 1 << RLL;
    constexpr int NG = (FT_SIZE >> RLL) / FTPB;      /* its groups per thread */
    constexpr int LNB = LOGS - RLL;
    static_assert(NG * RR == FT_PER, "group layout");
    /* Wave-local rounds: with S = 2^LOGS, LOGS a multiple of 3 (radix-8
     * rounds only) and one column (T = 1), wave w's groups in every round lie
     * in tile words [512w, 512w + 512), and so do the tile transfers it
     * issues (its four 1 KB chunks) and, below, the words it stores.  After
     * the stage twiddles are written, no wave reads what another wrote, so
     * the rounds need no block barrier, only the wave's own LDS wait. */
    constexpr bool WL = !FIRST && LOGS % 3 == 0;
    static_assert(!WL || (FT_SIZE / 128 / (FTPB / 64)) * 128 == 512, "a wave's tile chunks are its 512 words");
    FGeom g{0, 0, pa.logG};
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    uint64_t ra[FT_PER];
    size_t base = 0, step = 0;
    if constexpr (FIRST) {
        stage_twiddles_mid<PI, LOGS, RLL>(pa.small_tw, tw);
#pragma unroll
        for (int k = 0; k < FT_PER; k++) {
            uint32_t q = threadIdx.x + k * FTPB, m, r;
            lds[LP(q)] = first_in(pa, 0, ft_addr(LOGS, g, t, q, m, r));
        }
    } else {
        /* a's tile is in flight while the stage twiddles are written */
        tile_dma<LOGS>(pa, xa, g, t, lds);
        stage_twiddles_mid<PI, LOGS, RLL>(pa.small_tw, tw);
        WAIT_VMCNT(0);
    }
    lds_sync();      /* the stage twiddles (and, if not wave-local, a's tile) */
    dif_upper<PI, LOGS, LOGS - 1, RLL, WL>(lds, tw);
    /* a's smallest round, kept in registers: its groups leave LDS first */
#pragma unroll
    for (int n = 0; n < NG; n++) {
        const uint32_t gid = threadIdx.x + n * FTPB;
        const uint32_t base = ((gid >> LNB) << LOGS) + ((gid & ((1u << LNB) - 1)) << RLL);
#pragma unroll
        for (int i = 0; i < RR; i++) ra[n * RR + i] = lds[LP(base + i)];
    }
    /* b's tile into the same LDS once a's is in registers (a second buffer
     * would let it arrive during a's rounds, but at 80 KB per block two
     * blocks share a CU instead of three, which was measured slower); the
     * transfer is in flight during a's register round */
    if (!pa.square) {
        round_sync<WL>();   /* a's reads are done before b's tile overwrites them */
        if constexpr (FIRST) {
#pragma unroll
            for (int k = 0; k < FT_PER; k++) {
                uint32_t q = threadIdx.x + k * FTPB, m, r;
                lds[LP(q)] = first_in(pa, 1, ft_addr(LOGS, g, t, q, m, r));
            }
        } else {
            tile_dma<LOGS>(pa, xb, g, pa.b_tiles ? t % pa.b_tiles : t, lds);
        }
    }
    /* a's round between the transfer's issue and its wait: the empty asm
     * statements pin it there (the scheduler would otherwise sink it to the
     * pointwise product, after b's rounds) */
#pragma unroll
    for (int i = 0; i < FT_PER; i++) asm volatile("" : "+v"(ra[i]));
    /* a's values stay unreduced (below 16p): the Montgomery product takes
     * any first factor when the second is below 2p, and returns (0, 3p) */
#pragma unroll
    for (int n = 0; n < NG; n++) const_dif<PI, RLL, LZ_MAX>(ra + n * RR);
#pragma unroll
    for (int i = 0; i < FT_PER; i++) asm volatile("" : "+v"(ra[i]));
    if (!pa.square) {
        if constexpr (!FIRST) WAIT_VMCNT(0);
        round_sync<WL>();
        dif_upper<PI, LOGS, LOGS - 1, RLL, WL>(lds, tw);
    }
    const uint64_t s = pa.scale[PI], sp = pa.scale_p[PI];
#pragma unroll
    for (int n = 0; n < NG; n++) {
        const uint32_t gid = threadIdx.x + n * FTPB;
        const uint32_t base = ((gid >> LNB) << LOGS) + ((gid & ((1u << LNB) - 1)) << RLL);
        uint64_t x[RR];
        if (pa.square) {      /* the second factor below 2p */
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = reduce_any<PI>(ra[n * RR + i]);
        } else {
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = lds[LP(base + i)];
            const_dif<PI, RLL, 2>(x);
        }
        if (pa.scale_late) {    /* the last inverse pass applies the 1/L scale */
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = mont_mul_lazy<PI>(ra[n * RR + i], x[i]);
        } else {
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = shoup_lazy<PI>(mont_mul_lazy<PI>(ra[n * RR + i], x[i]), s, sp);
        }
        const_dit<PI, RLL, 4, LZ_MAX>(x);     /* the products are below 3p */
#pragma unroll
        for (int i = 0; i < RR; i++) lds[LP(base + i)] = x[i];
    }
    round_sync<WL>();
    if constexpr (WL && LOGS > RLL) {
        /* the top inverse round stores straight from registers (T = 1: the
         * tile is contiguous in its row, pads fall between tiles; for
         * S = 512, lane l of wave w holds words 512w + l + 64k, 512-byte runs) */
        uint32_t m, r;
        dit_upper<PI, LOGS, LOGS - 1, RLL, true, true>(lds, tw, xa + res_at(ft_addr(LOGS, g, t, 0, m, r), pa.pad_log));
    } else if constexpr (WL) {
        /* one round only (S = 8), done above: wave w stores its own words
         * 512w + lane + 64k from LDS */
        const uint32_t q0 = ((threadIdx.x >> 6) << 9) + (threadIdx.x & 63);
        uint32_t m, r;
        base = res_at(ft_addr(LOGS, g, t, q0, m, r), pa.pad_log);
        step = res_at(ft_addr(LOGS, g, t, q0 + 64, m, r), pa.pad_log) - base;
#pragma unroll
        for (int k = 0; k < FT_PER; k++) xa[base + k * step] = lds[LP(q0 + 64 * k)];
    } else {
        dit_upper<PI, LOGS, LOGS - 1, RLL>(lds, tw);
        tile_span<LOGS>(pa, g, t, threadIdx.x, base, step);
#pragma unroll
        for (int k = 0; k < FT_PER; k++) xa[base + k * step] = lds[LP(threadIdx.x + k * FTPB)];
    }
}

/* The wave-local fused pass over several tiles (kind 2, S = 2^6 or 2^9; the
 * rounds, transfers and stores of mid_tile's wave-local form).  A block takes
 * tiles j, j + nblk, ... (j its XCD-ordered index) and stages the stage
 * twiddles once.  A wave issues its part of the next tile's transfer of a as
 * soon as it has read its groups of the top inverse round (its own 512 tile
 * words: no block barrier), so the transfer is in flight during that round's
 * butterflies and stores. */
template <int PI, int LOGS>
__device__ void mid_tiles_wl(const PassArgs &pa, uint64_t *xa, const uint64_t *xb, uint64_t *lds, TwPair *tw)
{
    constexpr int RLL = 3, RR = 8, LNB = LOGS - RLL;
    static_assert(LOGS % 3 == 0 && LOGS >= 6 && (FT_SIZE >> RLL) == FTPB, "8-point groups, one per thread");
    static_assert((FT_SIZE / 128 / (FTPB / 64)) * 128 == 512, "a wave's tile chunks are its 512 words");
    const FGeom g{0, 0, pa.logG};
    const uint32_t nblk = gridDim.x;
    uint32_t t = xcd_tile(blockIdx.x, nblk);
    /* a's first tile is in flight while the stage twiddles are written */
    tile_dma<LOGS>(pa, xa, g, t, lds);
    stage_twiddles_mid<PI, LOGS, RLL>(pa.small_tw, tw);
    WAIT_VMCNT(0);
    lds_sync();      /* the stage twiddles and a's tile */
    const uint64_t s = pa.scale[PI], sp = pa.scale_p[PI];
    for (;;) {
        dif_upper<PI, LOGS, LOGS - 1, RLL, true>(lds, tw);
        /* a's smallest round, kept in registers: its group leaves LDS first */
        uint32_t gid = threadIdx.x;
        asm volatile("" : "+v"(gid));
        const uint32_t gbase = ((gid >> LNB) << LOGS) + ((gid & ((1u << LNB) - 1)) << RLL);
        uint64_t ra[RR];
#pragma unroll
        for (int i = 0; i < RR; i++) ra[i] = lds[LP(gbase + i)];
        if (!pa.square) {
            wave_lds_sync();   /* a's reads are done before b's tile overwrites them */
            tile_dma<LOGS>(pa, xb, g, pa.b_tiles ? t % pa.b_tiles : t, lds);
        }
        /* a's round between the transfer's issue and its wait (pinned there) */
#pragma unroll
        for (int i = 0; i < RR; i++) asm volatile("" : "+v"(ra[i]));
        const_dif<PI, RLL, LZ_MAX>(ra);
#pragma unroll
        for (int i = 0; i < RR; i++) asm volatile("" : "+v"(ra[i]));
        uint64_t x[RR];
        if (pa.square) {      /* the second factor below 2p */
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = reduce_any<PI>(ra[i]);
        } else {
            WAIT_VMCNT(0);
            wave_lds_sync();
            dif_upper<PI, LOGS, LOGS - 1, RLL, true>(lds, tw);
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = lds[LP(gbase + i)];
            const_dif<PI, RLL, 2>(x);
        }
        if (pa.scale_late) {    /* the last inverse pass applies the 1/L scale */
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = mont_mul_lazy<PI>(ra[i], x[i]);
        } else {
#pragma unroll
            for (int i = 0; i < RR; i++) x[i] = shoup_lazy<PI>(mont_mul_lazy<PI>(ra[i], x[i]), s, sp);
        }
        const_dit<PI, RLL, 4, LZ_MAX>(x);     /* the products are below 3p */
#pragma unroll
        for (int i = 0; i < RR; i++) lds[LP(gbase + i)] = x[i];
        wave_lds_sync();
        /* the inverse rounds above the smallest but for the top one */
        dit_upper<PI, LOGS, LOGS - 4, RLL, true, false>(lds, tw);
        /* the top round (half-size 2^(LOGS-1), groups of rows j + 2^(LOGS-3) i
         * of sub-transform gg), stored from registers (T = 1: the tile is
         * contiguous in its row) */
        constexpr int LH = LOGS - 1, LD = LH - 2;
        gid = threadIdx.x;
        asm volatile("" : "+v"(gid));
        const uint32_t j = gid & ((1u << LD) - 1);
        const uint32_t tbase = ((gid >> LD) << LOGS) + j;
        uint64_t v[RR];
#pragma unroll
        for (int i = 0; i < RR; i++) v[i] = lds[LP(tbase + ((uint32_t)i << LD))];
        lz_group<PI, 3, true, false, dit_bound(LH - 3, 4), LZ_MAX>(v, TabTwF<LH, LD, RR, true, RLL>{tw, j});
        uint32_t m, r;
        uint64_t *gout = xa + res_at(ft_addr(LOGS, g, t, 0, m, r), pa.pad_log);
        t += nblk;
        const bool more = t < pa.ntiles;
        if (more) {
            wave_lds_sync();   /* this wave's words read: its part of the buffer is free */
            tile_dma<LOGS>(pa, xa, g, t, lds);
        }
#pragma unroll
        for (int i = 0; i < RR; i++) gout[tbase + ((uint32_t)i << LD)] = v[i];
        if (!more) return;
        WAIT_VMCNT(0);
        wave_lds_sync();
    }
}

#define TW_LDS_MAX 1024  /* stage-twiddle entries per direction held in LDS (S - 1) */

/* kind 0: forward pass, 1: inverse pass, 2: fused middle pass,
 * 3: first forward pass (reads the operands), 4: single fused pass,
 * 5: inverse pass with the cross twiddles applied at load (C * S = 4096),
 * 6: kind 0 and 7: kind 5 with the cross twiddles from pa.xtab,
 * 8: kind 3 (S = 2^6, 2^9) and 9: kind 5 with the chains' ratios from
 * pa.btab. */
template <int PI, int LOGS, int KIND>
__device__ __forceinline__ void pass_body(const PassArgs &pa, uint64_t *xa, const uint64_t *xb, int op,
                                          uint64_t *lds, TwPair *tw)
{
    /* each routine stages its stage twiddles itself, after starting its
     * first tile's transfer */
    if constexpr (KIND == 0) fwd_tiles<PI, LOGS, false, false>(pa, xa, op, lds, tw);
    else if constexpr (KIND == 6) fwd_tiles<PI, LOGS, false, true>(pa, xa, op, lds, tw);
    else if constexpr (KIND == 3) fwd_tiles<PI, LOGS, true, false>(pa, xa, op, lds, tw);
    else if constexpr (KIND == 1) inv_tiles<PI, LOGS>(pa, xa, lds, tw);
    else if constexpr (KIND == 5) inv_tiles_dma<PI, LOGS, false, false>(pa, xa, lds, tw);
    else if constexpr (KIND == 7) inv_tiles_dma<PI, LOGS, true, false>(pa, xa, lds, tw);
    else if constexpr (KIND == 8) fwd_tiles<PI, LOGS, true, false, true>(pa, xa, op, lds, tw);
    else if constexpr (KIND == 9) inv_tiles_dma<PI, LOGS, false, true>(pa, xa, lds, tw);
    else if constexpr (KIND == 2 && LOGS % 3 == 0 && LOGS >= 6) mid_tiles_wl<PI, LOGS>(pa, xa, xb, lds, tw);
    else if constexpr (KIND == 2) mid_tile<PI, LOGS, false>(pa, xa, xb, lds, tw);
    else mid_tile<PI, LOGS, true>(pa, xa, xb, lds, tw);
}

/* Stage-twiddle entries in LDS: one direction's S - 1 (forward or inverse
 * passes), or the fused passes' shared table. */
template <int LOGS, int KIND>
constexpr uint32_t tw_lds_entries()
{
    /* (at least one entry: transforms of up to 8 points have no LDS stages) */
    return (KIND == 2 || KIND == 4) ? (mid_off<((LOGS - 1) % 3) + 1>(LOGS) ? mid_off<((LOGS - 1) % 3) + 1>(LOGS) : 1)
                                    : (1u << LOGS);
}

/* Every kind keeps the compiler's register allocation.  The 512-point passes
 * at 64 VGPRs or fewer (the inverse passes, 62) run four blocks per CU (40 KB
 * of LDS each); capped at 64 the others spill, among them the multi-tile
 * fused pass (18 registers), which runs faster at 70 VGPRs and three blocks
 * per CU than the single-tile one did at 64 and four (3.02-3.05 against
 * 3.10-3.11 ms at L = 2^27). */
template <int LOGS, int KIND>
__global__ void __launch_bounds__(FTPB)
k_fast_pass(PassArgs pa)
{
    __shared__ uint64_t lds[FT_LDS];
    __shared__ TwPair tws[tw_lds_entries<LOGS, KIND>()];

    static_assert((1 << LOGS) <= TW_LDS_MAX, "stage twiddles exceed LDS table");
    const uint32_t pi = blockIdx.y, op = blockIdx.z + pa.op0;
    uint64_t *xa = pa.res + ((size_t)op * NTT_NPRIMES + pi) * pa.len;
    const uint64_t *xb = pa.bres ? pa.bres + (size_t)pi * pa.blen : pa.res + ((size_t)NTT_NPRIMES + pi) * pa.len;
    if (pi == 0) pass_body<0, LOGS, KIND>(pa, xa, xb, op, lds, tws);
    else if (pi == 1) pass_body<1, LOGS, KIND>(pa, xa, xb, op, lds, tws);
    else pass_body<2, LOGS, KIND>(pa, xa, xb, op, lds, tws);
}

/* Host-side launcher: one pass of a class (grid.x tiles, grid.y primes,
 * grid.z operands for forward passes). */
hipError_t launch_fast_pass(int logS, int kind, const PassArgs &pa, size_t ntiles, int nops, hipStream_t st)
{
    /* forward and inverse passes: tpb tiles per block; fused passes: one */
    const bool looped = kind == 0 || kind == 1 || kind == 3 || kind == 5 || kind == 6 || kind == 7 || kind == 8 ||
                        kind == 9 || (kind == 2 && (logS == 6 || logS == 9));
    if (looped && (pa.tpb < 1 || pa.ntiles != ntiles)) return hipErrorInvalidValue;
    if ((kind == 5 || kind == 7 || kind == 9) && (logS < 3 || pa.logC + logS != 12)) return hipErrorInvalidValue;
    /* ratio tables: the first pass with 8-point last-round groups, and the
     * load chain */
    if ((kind == 8 || kind == 9) && !pa.btab) return hipErrorInvalidValue;
    if (kind == 8 && (!pa.first || (logS != 6 && logS != 9))) return hipErrorInvalidValue;
    /* tile transfers move column pairs (C >= 2) or row pairs (T = 1) */
    if ((kind == 0 || kind == 1 || kind == 6) && pa.logC == 0 && pa.logT != 0) return hipErrorInvalidValue;
    /* table twiddles: a table of this pass's S x T block */
    if ((kind == 6 || kind == 7) && (!pa.xtab || logS + (int)pa.logT > NTT_XTAB_MAX_LOG)) return hipErrorInvalidValue;
    const size_t nblk = looped ? (ntiles + pa.tpb - 1) / pa.tpb : ntiles;
    const bool fwd = kind == 0 || kind == 3 || kind == 6 || kind == 8;
    dim3 grid((unsigned)nblk, NTT_NPRIMES, fwd ? nops : 1);
    dim3 block(FTPB);
#define L1(S)                                                                                  \
    case S:                                                                                    \
        switch (kind) {                                                                        \
        case 0: hipLaunchKernelGGL((k_fast_pass<S, 0>), grid, block, 0, st, pa); break;        \
        case 1: hipLaunchKernelGGL((k_fast_pass<S, 1>), grid, block, 0, st, pa); break;        \
        case 2: hipLaunchKernelGGL((k_fast_pass<S, 2>), grid, block, 0, st, pa); break;        \
        case 3: hipLaunchKernelGGL((k_fast_pass<S, 3>), grid, block, 0, st, pa); break;        \
        case 4: hipLaunchKernelGGL((k_fast_pass<S, 4>), grid, block, 0, st, pa); break;        \
        default: return hipErrorInvalidValue;                                                  \
        }                                                                                      \
        break;
#define L5(S)                                                                                  \
    case S:                                                                                    \
        switch (kind) {                                                                        \
        case 0: hipLaunchKernelGGL((k_fast_pass<S, 0>), grid, block, 0, st, pa); break;        \
        case 1: hipLaunchKernelGGL((k_fast_pass<S, 1>), grid, block, 0, st, pa); break;        \
        case 2: hipLaunchKernelGGL((k_fast_pass<S, 2>), grid, block, 0, st, pa); break;        \
        case 3: hipLaunchKernelGGL((k_fast_pass<S, 3>), grid, block, 0, st, pa); break;        \
        case 4: hipLaunchKernelGGL((k_fast_pass<S, 4>), grid, block, 0, st, pa); break;        \
        case 5: hipLaunchKernelGGL((k_fast_pass<S, 5>), grid, block, 0, st, pa); break;        \
        case 6: hipLaunchKernelGGL((k_fast_pass<S, 6>), grid, block, 0, st, pa); break;        \
        case 7: hipLaunchKernelGGL((k_fast_pass<S, 7>), grid, block, 0, st, pa); break;        \
        case 9: hipLaunchKernelGGL((k_fast_pass<S, 9>), grid, block, 0, st, pa); break;        \
        default: return hipErrorInvalidValue;                                                  \
        }                                                                                      \
        break;
    if (kind == 8) {   /* checked above: S = 2^6 or 2^9 */
        if (logS == 6) hipLaunchKernelGGL((k_fast_pass<6, 8>), grid, block, 0, st, pa);
        else hipLaunchKernelGGL((k_fast_pass<9, 8>), grid, block, 0, st, pa);
        return hipGetLastError();
    }
    switch (logS) {
        L1(1) L1(2) L5(3) L5(4) L5(5) L5(6) L5(7) L5(8) L5(9) L5(10)
    default:
        return hipErrorInvalidValue;
    }
#undef L1
#undef L5
    return hipGetLastError();
}

hipError_t launch_xtab_gen(const TwPair *tw, uint64_t *out, int logS, int logT, const XtabScale &sc, hipStream_t st)
{
    const int logN = logS + logT;
    if (logS < 1 || logT < 0 || logN > NTT_XTAB_MAX_LOG || logN < 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_xtab_gen, dim3(1u << (logN - 8), NTT_NPRIMES, NTT_XTAB_DIRS), dim3(256), 0, st, tw, out,
                       (uint32_t)logS, (uint32_t)logT, sc);
    return hipGetLastError();
}

/* grid (T / 256, primes, NTT_BTAB_DIRS): the chain ratios (layout in ntt_kernels.h) */
template <int PI>
__device__ __forceinline__ TwPair btab_entry(const TwPair *tw, uint32_t logS, uint32_t logT, uint32_t r, bool inv)
{
    const uint32_t shift = NTT_ROOT_LOG - logS - logT;
    const uint32_t E = (inv ? 0u - (r << shift) : (r << (logS - 3)) << shift) & ((1u << NTT_ROOT_LOG) - 1);
    const TwPair lo = tw[E & (NTT_TW_SIZE - 1)], hi = tw[NTT_TW_SIZE + (E >> NTT_TW_LOG)];
    const uint64_t w = reduce1<PI>(shoup_lazy<PI>(lo.w, hi.w, hi.wp));
    /* floor(w * 2^64 / p) bit by bit (w < p < 2^60: no overflow) */
    uint64_t rem = w, q = 0;
    for (int i = 63; i >= 0; i--) {
        rem <<= 1;
        if (rem >= Prime<PI>::P) {
            rem -= Prime<PI>::P;
            q |= (uint64_t)1 << i;
        }
    }
    return TwPair{w, q};
}
__global__ void __launch_bounds__(256) k_btab_gen(const TwPair *tw, TwPair *out, uint32_t logS, uint32_t logT)
{
    const uint32_t T = 1u << logT;
    const uint32_t r = blockIdx.x * 256 + threadIdx.x, pi = blockIdx.y;
    const bool inv = blockIdx.z != 0;
    if (r >= T) return;
    const TwPair *t = tw + (size_t)pi * 2 * NTT_TW_SIZE;
    out[((size_t)blockIdx.z * NTT_NPRIMES + pi) * T + r] = pi == 0   ? btab_entry<0>(t, logS, logT, r, inv)
                                                           : pi == 1 ? btab_entry<1>(t, logS, logT, r, inv)
                                                                     : btab_entry<2>(t, logS, logT, r, inv);
}

hipError_t launch_btab_gen(const TwPair *tw, TwPair *out, int logS, int logT, hipStream_t st)
{
    if (logS < 3 || logS > FT_MAX_LOGS || logT < 0 || logS + logT > NTT_MAX_LOG_LEN) return hipErrorInvalidValue;
    const unsigned T = 1u << logT;
    hipLaunchKernelGGL(k_btab_gen, dim3((T + 255) / 256, NTT_NPRIMES, NTT_BTAB_DIRS), dim3(256), 0, st, tw, out,
                       (uint32_t)logS, (uint32_t)logT);
    return hipGetLastError();
}

hipError_t launch_small_twiddles(const TwPair *tw, TwPair *small, hipStream_t st)
{
    hipLaunchKernelGGL(k_small_twiddles, dim3(FT_SIZE / 2 / 256, NTT_NPRIMES), dim3(256), 0, st, tw, small);
    return hipGetLastError();
}
