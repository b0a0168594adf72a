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

/* floor(w * 2^64 / p) for w < p < 2^62 by restoring long division. */
template <int PI>
__device__ uint64_t shoup_partner(uint64_t w)
{
    const uint64_t p = Prime<PI>::P;
    uint64_t r = w, q = 0;
    for (int i = 63; i >= 0; i--) {
        r <<= 1;                       /* r < p < 2^62: no overflow */
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

/* x = a0 + a1*p0 + a2*p0*p1 in the reference's order (garner.hip:135-173);
 * the product coefficients are < L*(2^16-1)^2 < 2^62 < p0, so a1 = a2 = 0 for
 * an intact product and the value is a0.  Any nonzero a1/a2 is a corrupted
 * product and raises the integrity flag. */
__device__ __forceinline__ uint64_t garner_coeff(const uint64_t *r0, const uint64_t *r1, const uint64_t *r2,
                                                 size_t k, const GarnerConsts &gc, int &bad)
{
    const uint64_t p0 = Prime<0>::P, p1 = Prime<1>::P, p2 = Prime<2>::P;
    uint64_t x0 = reduce1<0>(r0[k]);
    uint64_t x1 = reduce1<1>(r1[k]);
    uint64_t x2 = reduce1<2>(r2[k]);
    uint64_t a0 = x0;
    uint64_t a0m1 = a0 >= p1 ? a0 - p1 : a0;            /* a0 < p0 < 2*p1 */
    uint64_t t1 = x1 >= a0m1 ? x1 - a0m1 : x1 + p1 - a0m1;
    uint64_t a1 = reduce1<1>(shoup_lazy<1>(t1, gc.inv01, gc.inv01p));
    uint64_t a0m2 = a0 >= p2 ? a0 - p2 : a0;
    uint64_t b2 = a0m2 + reduce1<2>(shoup_lazy<2>(a1, gc.p0m2, gc.p0m2p));
    b2 = b2 >= p2 ? b2 - p2 : b2;
    uint64_t t2 = x2 >= b2 ? x2 - b2 : x2 + p2 - b2;
    uint64_t a2 = reduce1<2>(shoup_lazy<2>(t2, gc.inv012, gc.inv012p));
    (void)p0;
    bad |= (a1 | a2) != 0;
    return a0;
}

/* (G,P) cell composition: lower-index cell `lo` then higher `hi`;
 * identity is (0,1), encoded G | P<<1 = 2. */
__device__ __forceinline__ uint8_t cla_compose(uint8_t lo, uint8_t hi)
{
    uint8_t g = (hi & 1) | ((hi >> 1) & lo & 1);
    uint8_t p = (uint8_t)((hi >> 1) & (lo >> 1) & 1);
    return (uint8_t)(g | (p << 1));
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

/* K_a (grid: carry blocks x products): each thread owns CARRY_E consecutive
 * output limbs of one product.  s_j = sum_t X_{4j+t} 2^(16t) (+ addend_j) is
 * split into lo_j + hi_j 2^64; res_j = lo_j + hi_{j-1} mod 2^64 is written to
 * out[j], its overflow bit (generate) to gen[], and the block's (G,P)
 * aggregate to blk[]. */
__global__ void __launch_bounds__(256) k_garner_pack(GarnerArgs ga)
{
    __shared__ uint8_t sg[256];
    const uint32_t prod = blockIdx.y;
    const MulDesc d = ga.desc ? ga.desc[prod] : ga.one;
    const size_t L = (size_t)1 << ga.logL;
    const size_t off = (size_t)prod << ga.logL;
    const uint64_t *r0 = ga.res + off, *r1 = ga.res + ga.len + off, *r2 = ga.res + 2 * ga.len + off;
    const size_t n_out = d.n_out;
    const size_t n_coef = 4 * ((size_t)d.na + d.nb) < L ? 4 * ((size_t)d.na + d.nb) : L;
    const size_t thr = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t j0 = thr * CARRY_E;
    int bad = 0;
    uint8_t G = 0, P = 1, gbits = 0;
    /* limbs at or above n_out that the convolution reaches must be zero */
    const size_t n_span = (n_coef + 3) / 4;
    if (j0 >= n_out && j0 < n_span) {
        for (size_t k = 4 * j0; k < 4 * (j0 + CARRY_E) && k < n_coef; k++)
            bad |= garner_coeff(r0, r1, r2, k, ga.gc, bad) != 0;
    }
    if (j0 < n_out) {
        uint64_t prev_hi = 0;
        for (int jj = -1; jj < CARRY_E; jj++) {
            if (jj == -1 && j0 == 0) continue;
            size_t j = j0 + jj;
            if (j >= n_out) break;
            uint64_t lo = 0, hi = 0;
            for (int t = 0; t < 4; t++) {
                size_t k = 4 * j + t;
                if (k >= n_coef) break;
                uint64_t X = garner_coeff(r0, r1, r2, k, ga.gc, bad);
                if (t == 0) {
                    lo = X;
                } else {
                    uint64_t sl = X << (16 * t), sh = X >> (64 - 16 * t);
                    uint64_t nl = lo + sl;
                    hi += sh + (nl < lo);
                    lo = nl;
                }
            }
            if (d.add && j < d.nadd) {
                uint64_t nl = lo + d.add[j];
                hi += nl < lo;
                lo = nl;
            }
            if (jj >= 0) {
                uint64_t res = lo + prev_hi;
                uint8_t g = res < lo;
                uint8_t p = res == ~0ull;
                d.c[j] = res;
                gbits |= (uint8_t)(g << jj);
                G = g | (p & G);
                P = P & p;
            }
            prev_hi = hi;
        }
        /* the high part of the top limb must vanish */
        if (j0 + CARRY_E >= n_out) bad |= prev_hi != 0;
        ga.gen[(size_t)prod * ga.thr_per_prod + thr] = gbits;
    }
    if (bad) ga.flag[0] = 1;
    sg[threadIdx.x] = (uint8_t)(G | (P << 1));
    __syncthreads();
    cla_block_scan<256>(sg);
    if (threadIdx.x == 255) ga.blk[(size_t)prod * ga.blk_per_prod + blockIdx.x] = sg[255];
}

/* K_b (grid: products): exclusive carry-in of every block of one product. */
__global__ void __launch_bounds__(1024) k_carry_scan(uint8_t *blkall, uint32_t nblk)
{
    __shared__ uint8_t sg[1024];
    uint8_t *blk = blkall + (size_t)blockIdx.x * nblk;
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t s0 = threadIdx.x * per;
    uint8_t agg = 2;
    for (uint32_t i = s0; i < s0 + per && i < nblk; i++) agg = cla_compose(agg, blk[i]);
    sg[threadIdx.x] = agg;
    __syncthreads();
    cla_block_scan<1024>(sg);
    uint8_t cin = threadIdx.x ? (sg[threadIdx.x - 1] & 1) : 0;
    __syncthreads();
    for (uint32_t i = s0; i < s0 + per && i < nblk; i++) {
        uint8_t c = blk[i];
        blk[i] = cin;                 /* becomes the block's carry-in */
        cin = (c & 1) | (((c >> 1) & 1) & cin);
    }
}

/* K_c (grid: carry blocks x products): thread carry-in = block carry-in
 * composed with the in-block exclusive prefix; then ripple through the
 * thread's limbs. */
__global__ void __launch_bounds__(256) k_carry_apply(GarnerArgs ga)
{
    __shared__ uint8_t sg[256];
    const uint32_t prod = blockIdx.y;
    const MulDesc d = ga.desc ? ga.desc[prod] : ga.one;
    const size_t n_out = d.n_out;
    const size_t thr = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t j0 = thr * CARRY_E;
    uint8_t G = 0, P = 1, gbits = 0;
    uint64_t v[CARRY_E];
    if (j0 < n_out) {
        gbits = ga.gen[(size_t)prod * ga.thr_per_prod + thr];
        for (int jj = 0; jj < CARRY_E; jj++) {
            size_t j = j0 + jj;
            v[jj] = j < n_out ? d.c[j] : 0;
            uint8_t g = (gbits >> jj) & 1;
            uint8_t p = j < n_out && v[jj] == ~0ull;
            G = g | (p & G);
            P = P & p;
        }
    }
    sg[threadIdx.x] = (uint8_t)(G | (P << 1));
    __syncthreads();
    cla_block_scan<256>(sg);
    uint8_t bc = ga.blk[(size_t)prod * ga.blk_per_prod + blockIdx.x];
    uint8_t cin;
    if (threadIdx.x == 0) cin = bc;
    else {
        uint8_t e = sg[threadIdx.x - 1];
        cin = (e & 1) | ((e >> 1) & bc);
    }
    if (j0 < n_out) {
        for (int jj = 0; jj < CARRY_E; jj++) {
            size_t j = j0 + jj;
            if (j >= n_out) {
                if (cin) ga.flag[0] = 1;     /* carry out of the top limb */
                cin = 0;
                break;
            }
            uint64_t r = v[jj];
            uint8_t g = (gbits >> jj) & 1;
            uint8_t p = r == ~0ull;
            d.c[j] = r + cin;
            cin = g | (p & cin);
        }
        if (cin && j0 + CARRY_E == n_out) ga.flag[0] = 1;
    }
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
