/*
 * ntt_kernels.hip — gfx950 kernels of the 3-prime CRT-NTT big-integer multiply.
 *
 * Twiddle-table generation, Garner CRT with the 16-bit repack and the
 * carry-lookahead carry, and the small-operand schoolbook.  The transform
 * passes are in ntt_fast.hip.
 *
 * Data layout (HBM): a size class of nprod products with transform length
 * L = 2^logL keeps residues in res[op][prime][row], row = nprod*L rounded up
 * to a tile (op 0 = a, op 1 = b), product i at [i*L, (i+1)*L).
 *
 * Reference counterparts: ntt/lib/garner.hip:187-260 (garner_kernel),
 * ntt/lib/transfer_kernels.hip:55-313 (gather / carry-lookahead kernels).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.h"
#include "ntt_kernels.h"

using namespace nttc;

/* ------------------------------------------------------------------------ */
/* Twiddle tables: tw[prime][0][j] = w^j, tw[prime][1][j] = w^(j*2^15) for a  */
/* root w of order 2^30, each with its Shoup partner floor(t*2^64/p).         */
/* ------------------------------------------------------------------------ */

template <int PI>
__device__ uint64_t mulmod_dev(uint64_t a, uint64_t b, uint64_t r2)
{
    /* a*b*2^-64 then *2^128*2^-64: exact a*b mod p for a, b < p */
    return mont_mul<PI>(mont_mul<PI>(a, b), r2);
}

template <int PI>
__device__ uint64_t powmod_dev(uint64_t b, uint64_t e, uint64_t r2)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod_dev<PI>(r, b, r2);
        b = mulmod_dev<PI>(b, b, r2);
        e >>= 1;
    }
    return r;
}

/* floor(w * 2^64 / p) for w < p < 2^60 by restoring long division. */
template <int PI>
__device__ uint64_t shoup_partner(uint64_t w)
{
    const uint64_t p = Prime<PI>::P;
    uint64_t r = w, q = 0;
    for (int i = 63; i >= 0; i--) {
        r <<= 1;                       /* r < p < 2^60: no overflow */
        if (r >= p) { r -= p; q |= (uint64_t)1 << i; }
    }
    return q;
}

template <int PI>
__device__ void twiddle_gen_one(TwPair *tw, uint32_t j, uint64_t r2)
{
    const uint64_t p = Prime<PI>::P;
    uint64_t w = powmod_dev<PI>(prime_gen(PI), (p - 1) >> NTT_ROOT_LOG, r2);
    uint64_t lo = powmod_dev<PI>(w, j, r2);
    uint64_t hi = powmod_dev<PI>(w, (uint64_t)j << NTT_TW_LOG, r2);
    TwPair *t = tw + (size_t)PI * 2 * NTT_TW_SIZE;
    t[j].w = lo;
    t[j].wp = shoup_partner<PI>(lo);
    t[NTT_TW_SIZE + j].w = hi;
    t[NTT_TW_SIZE + j].wp = shoup_partner<PI>(hi);
}

__global__ void __launch_bounds__(256) k_twiddle_gen(TwPair *tw, uint64_t r2_0, uint64_t r2_1, uint64_t r2_2)
{
    uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= NTT_TW_SIZE) return;
    switch (blockIdx.y) {
    case 0: twiddle_gen_one<0>(tw, j, r2_0); break;
    case 1: twiddle_gen_one<1>(tw, j, r2_1); break;
    default: twiddle_gen_one<2>(tw, j, r2_2); break;
    }
}

/* ------------------------------------------------------------------------ */
/* Garner CRT + 16-bit repack + carry lookahead                               */
/* ------------------------------------------------------------------------ */

/* x = a0 + a1*p0 + a2*p0*p1 in the reference's mixed-radix form
 * (garner.hip:135-173).  A product coefficient sums at most L/2 products of
 * 16-bit coefficients, so it is below (L/2)(2^16-1)^2: below p0 (> 2^60 -
 * 2^45) up to L = 2^29, below 2 p0 at L = 2^30.  An intact product
 * therefore has a2 = 0 and a1 = 0 or 1: the value is x0 or x0 + p0.  Any
 * other digits are a corrupted product and raise the integrity flag; the
 * test needs no product (crt_below_2p0). */
__device__ __forceinline__ uint64_t garner3(uint64_t r0, uint64_t r1, uint64_t r2, int &bad)
{
    /* inverse-pass output: any 64-bit value */
    const uint64_t x0 = reduce1<0>(reduce_any<0>(r0));
    const uint64_t x1 = reduce1<1>(reduce_any<1>(r1));
    const uint64_t x2 = reduce1<2>(reduce_any<2>(r2));
    bool ok;
    const uint64_t X = crt_below_2p0(x0, x1, x2, ok);
    bad |= !ok;
    return X;
}

/* Inclusive scan of 256 (or 1024) cells in LDS, lower thread index first. */
template <int N>
__device__ __forceinline__ void cla_block_scan(uint8_t *sg)
{
    for (int s = 1; s < N; s <<= 1) {
        uint8_t cur = sg[threadIdx.x];
        uint8_t lo = (threadIdx.x >= (unsigned)s) ? sg[threadIdx.x - s] : (uint8_t)2;
        __syncthreads();
        sg[threadIdx.x] = cla_compose(lo, cur);
        __syncthreads();
    }
}

/* ---- Garner + repack + carry: wavefront layout ------------------------------
 * A block is 4 wavefronts; wavefront w owns CW_LIMBS consecutive output limbs,
 * taken in CW_ROUNDS rounds of 64 (lane l of round e: limb base + 64e + l), so
 * every load and store is a contiguous 64-lane access.  Within a round the
 * thin carries follow from two ballots: with X = G|P and Y = G (generate and
 * propagate bits of the 64 limbs), the carry into lane i is bit i of
 * (X + Y + cin) ^ X ^ Y and the round's carry out is the overflow bit — the
 * generate/propagate recurrence of ntt/lib/transfer_core.h:165-252 evaluated
 * by the adder itself. */

/* (lo, hi) = sum_{t<4} X_{4j+t} 2^(16t) for limb j, with Garner per coefficient */
/* Residues of prime i for the product at offset off: res + i*len + res_at(off + k).
 * The 4 coefficients of a limb never straddle a pad (pads follow every
 * 2^pad_log >= 4096 residues, products start at multiples of 4). */
__device__ __forceinline__ void limb_sum(const GarnerArgs &ga, size_t off, size_t j, size_t n_coef, int &bad,
                                         uint64_t &lo, uint64_t &hi)
{
    lo = 0;
    hi = 0;
    size_t k0 = 4 * j;
    if (k0 >= n_coef) return;
    const size_t at = res_at(off + k0, ga.pad_log);
    const uint64_t *r0 = ga.res + at, *r1 = ga.res + ga.len + at, *r2 = ga.res + 2 * ga.len + at;
    uint64_t v0[4], v1[4], v2[4];
    if (k0 + 4 <= n_coef) {
        const ulonglong2 *p0 = (const ulonglong2 *)r0, *p1 = (const ulonglong2 *)r1,
                         *p2 = (const ulonglong2 *)r2;
        ulonglong2 a0 = p0[0], a1 = p0[1], b0 = p1[0], b1 = p1[1], c0 = p2[0], c1 = p2[1];
        v0[0] = a0.x; v0[1] = a0.y; v0[2] = a1.x; v0[3] = a1.y;
        v1[0] = b0.x; v1[1] = b0.y; v1[2] = b1.x; v1[3] = b1.y;
        v2[0] = c0.x; v2[1] = c0.y; v2[2] = c1.x; v2[3] = c1.y;
    } else {
        for (int t = 0; t < 4; t++) {
            bool in = k0 + t < n_coef;
            v0[t] = in ? r0[t] : 0;
            v1[t] = in ? r1[t] : 0;
            v2[t] = in ? r2[t] : 0;
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        uint64_t X = garner3(v0[t], v1[t], v2[t], bad);
        if (t == 0) {
            lo = X;
        } else {
            uint64_t sl = X << (16 * t), sh = X >> (64 - 16 * t);
            uint64_t nl = lo + sl;
            hi += sh + (nl < lo);
            lo = nl;
        }
    }
}

/* K_a (grid: blocks x products).  Writes res_j (pre-carry digit) to c[j], the
 * round generate masks to gen[], the block's (G,P) aggregate to blk[]. */
__global__ void __launch_bounds__(256) k_garner_pack(GarnerArgs ga)
{
    __shared__ uint8_t wcell[4];
    const uint32_t prod = blockIdx.y;
    const MulDesc d = ga.desc ? ga.desc[prod] : ga.one;
    const size_t L = (size_t)1 << ga.logL;
    const size_t off = (size_t)prod << ga.logL;
    const size_t n_out = d.n_out;
    const size_t n_coef = 4 * ((size_t)d.na + d.nb) < L ? 4 * ((size_t)d.na + d.nb) : L;
    const size_t n_span = (n_coef + 3) / 4;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t wave = (size_t)blockIdx.x * 4 + w;
    const size_t base = wave * CW_LIMBS;
    uint64_t *gmask = (uint64_t *)ga.gen + ((size_t)prod * ga.blk_per_prod * 4 + wave) * CW_ROUNDS;
    int bad = 0;
    uint64_t prev_hi = 0;   /* hi of limb base-1, valid in lane 0 */
    if (lane == 0 && base > 0 && base - 1 < n_span) {
        uint64_t lo;
        limb_sum(ga, off, base - 1, n_coef, bad, lo, prev_hi);
        if (d.add && base - 1 < d.nadd) {
            uint64_t nl = lo + d.add[base - 1];
            prev_hi += nl < lo;
        }
    }
    uint8_t agg = 2;    /* identity (G=0, P=1) */
#pragma unroll 1
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        uint64_t lo = 0, hi = 0;
        if (j < n_span) limb_sum(ga, off, j, n_coef, bad, lo, hi);
        if (d.add && j < d.nadd) {
            uint64_t nl = lo + d.add[j];
            hi += nl < lo;
            lo = nl;
        }
        uint64_t up = __shfl_up(hi, 1);
        uint64_t hprev = lane ? up : prev_hi;
        prev_hi = __shfl(hi, 63);
        uint64_t res = lo + hprev;
        bool g = res < lo, p = res == ~0ull;
        if (j < n_out) {
            d.c[j] = res;
            if (j == n_out - 1) bad |= hi != 0;          /* nothing may spill past the top limb */
        } else {
            bad |= (lo | hi | (j == n_out ? hprev : 0)) != 0;
            g = false;
            p = true;                                       /* identity cells */
        }
        uint64_t G = __ballot(g), P = __ballot(p);
        if (lane == 0) gmask[e] = G;
        uint8_t cell = round_cell(G, P);
        agg = cla_compose(agg, cell);
    }
    if (bad) ga.flag[0] = 1;
    if (lane == 0) wcell[w] = agg;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint8_t b = 2;
        for (int k = 0; k < 4; k++) b = cla_compose(b, wcell[k]);
        ga.blk[(size_t)prod * ga.blk_per_prod + blockIdx.x] = b;
    }
}

/* K_b (grid: products): exclusive carry-in of every block of one product. */
/* Cells of up to CS_LDS blocks are staged in LDS: read and written back with
 * coalesced, unrolled byte accesses, so each thread's serial walk over its
 * segment reads LDS instead of issuing dependent global loads. */
#define CS_LDS 32768
__global__ void __launch_bounds__(1024) k_carry_scan(uint8_t *blkall, uint32_t nblk, uint8_t cin0)
{
    __shared__ uint8_t sg[1024];
    __shared__ uint8_t cells[CS_LDS];
    uint8_t *gblk = blkall + (size_t)blockIdx.x * nblk;
    const bool staged = nblk <= CS_LDS;
    if (staged) {
#pragma unroll 8
        for (uint32_t i = threadIdx.x; i < nblk; i += 1024) cells[i] = gblk[i];
        __syncthreads();
    }
    uint8_t *blk = staged ? cells : gblk;
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t s0 = threadIdx.x * per;
    uint8_t agg = 2;
    for (uint32_t i = s0; i < s0 + per && i < nblk; i++) agg = cla_compose(agg, blk[i]);
    sg[threadIdx.x] = agg;
    __syncthreads();
    cla_block_scan<1024>(sg);
    /* carry into this thread's segment: the prefix before it, fed by cin0 */
    uint8_t cin = threadIdx.x ? (uint8_t)((sg[threadIdx.x - 1] & 1) | ((sg[threadIdx.x - 1] >> 1) & cin0)) : cin0;
    __syncthreads();
    for (uint32_t i = s0; i < s0 + per && i < nblk; i++) {
        uint8_t c = blk[i];
        blk[i] = cin;                 /* becomes the block's carry-in */
        cin = (c & 1) | (((c >> 1) & 1) & cin);
    }
    if (staged) {
        __syncthreads();
#pragma unroll 8
        for (uint32_t i = threadIdx.x; i < nblk; i += 1024) gblk[i] = cells[i];
    }
}

/* K_c (grid: blocks x products): carry-in of each wavefront = block carry-in
 * composed with the earlier wavefronts' aggregates; each round then adds its
 * carry vector (X + Y + cin) ^ X ^ Y and passes the overflow on. */
__global__ void __launch_bounds__(256) k_carry_apply(GarnerArgs ga)
{
    __shared__ uint8_t wcell[4];
    const uint32_t prod = blockIdx.y;
    const MulDesc d = ga.desc ? ga.desc[prod] : ga.one;
    const size_t n_out = d.n_out;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t wave = (size_t)blockIdx.x * 4 + w;
    const size_t base = wave * CW_LIMBS;
    const uint64_t *gmask = (const uint64_t *)ga.gen + ((size_t)prod * ga.blk_per_prod * 4 + wave) * CW_ROUNDS;
    uint64_t v[CW_ROUNDS], G[CW_ROUNDS], P[CW_ROUNDS];
    uint8_t agg = 2;
#pragma unroll
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        v[e] = j < n_out ? d.c[j] : 0;
        G[e] = gmask[e];
        P[e] = __ballot(j >= n_out || v[e] == ~0ull);
        agg = cla_compose(agg, round_cell(G[e], P[e]));
    }
    if (lane == 0) wcell[w] = agg;
    __syncthreads();
    uint8_t c = ga.blk[(size_t)prod * ga.blk_per_prod + blockIdx.x];   /* block carry-in */
    for (uint32_t k = 0; k < w; k++) {
        uint8_t x = wcell[k];
        c = (x & 1) | ((x >> 1) & c);
    }
    uint64_t cin = c & 1;
#pragma unroll
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        uint64_t X = G[e] | P[e], Y = G[e];
        uint64_t s = X + Y;
        uint64_t s2 = s + cin;
        uint64_t carries = s2 ^ X ^ Y;
        uint64_t cout = (s < X) | (s2 < s);
        uint64_t cj = (carries >> lane) & 1;
        if (j < n_out) d.c[j] = v[e] + cj;
        else if (j == n_out && cj) ga.flag[0] = 1;        /* carry out of the top limb */
        cin = cout;
    }
    if (cin && base + CW_LIMBS == (size_t)n_out) ga.flag[0] = 1;
}

/* ------------------------------------------------------------------------ */
/* Schoolbook for small operands (na + nb <= SB_MAX_OUT): one wavefront per    */
/* product, 4 products per block.  Column sums in 3 words, then lane 0        */
/* ripples the carry (ntt/lib/arith/multiply.c:35-67 semantics, on the GPU).  */
/* ------------------------------------------------------------------------ */

__global__ void __launch_bounds__(256) k_schoolbook(const MulDesc *desc, MulDesc one, uint32_t nprod, int *flag)
{
    __shared__ uint64_t colall[4][3 * SB_MAX_OUT];
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t prod = blockIdx.x * 4 + w;
    const bool valid = prod < nprod;
    MulDesc d = one;
    if (valid && desc) d = desc[prod];
    uint64_t *col = colall[w];
    const size_t na = d.na, nb = d.nb, nout = na + nb;
    if (valid) {
        for (size_t j = lane; j < nout; j += 64) {
            uint64_t w0 = 0, w1 = 0, w2 = 0;
            size_t ilo = j >= nb ? j - nb + 1 : 0;
            size_t ihi = j < na ? j : na - 1;
            for (size_t i = ilo; i <= ihi && j + 1 < nout; i++) {
                uint64_t x = d.a[i], y = d.b[j - i];
                uint64_t lo = x * y, hi = __umul64hi(x, y);
                uint64_t s0 = w0 + lo;
                uint64_t c0 = s0 < w0;
                w0 = s0;
                uint64_t s1 = w1 + hi;
                uint64_t c1 = s1 < w1;
                uint64_t s2 = s1 + c0;
                c1 += s2 < s1;
                w1 = s2;
                w2 += c1;
            }
            if (d.add && j < d.nadd) {
                uint64_t s0 = w0 + d.add[j];
                uint64_t c0 = s0 < w0;
                w0 = s0;
                uint64_t s1 = w1 + c0;
                w2 += s1 < w1;
                w1 = s1;
            }
            col[3 * j] = w0;
            col[3 * j + 1] = w1;
            col[3 * j + 2] = w2;
        }
    }
    __syncthreads();
    if (valid && lane == 0) {
        uint64_t c0 = 0, c1 = 0;   /* running carry (two words) */
        const size_t nspan = nout > d.n_out ? nout : d.n_out;
        uint64_t spill = 0;
        for (size_t j = 0; j < nspan; j++) {
            uint64_t x0 = j < nout ? col[3 * j] : 0, x1 = j < nout ? col[3 * j + 1] : 0,
                     x2 = j < nout ? col[3 * j + 2] : 0;
            if (j >= nout && d.add && j < d.nadd) x0 = d.add[j];
            uint64_t s = x0 + c0;
            uint64_t k = s < c0;
            uint64_t n1 = x1 + c1;
            uint64_t k1 = n1 < c1;
            uint64_t n1b = n1 + k;
            k1 += n1b < n1;
            if (j < d.n_out) d.c[j] = s;
            else spill |= s;
            c0 = n1b;
            c1 = x2 + k1;
        }
        if (c0 | c1 | spill) flag[0] = 1;   /* product does not fit n_out limbs */
    }
}
