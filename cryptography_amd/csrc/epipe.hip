/*
 * epipe.hip — Euler's number on the device: product tree, Newton reciprocal,
 * fixed-point division and radix conversion, all around the CRT-NTT engine.
 *
 * The value printed is the reference's (ntt/app/compute_e/main.c:167-345):
 * with n_total = D + 20 guard digits and the tree of sum_{k=1}^{N} 1/k!
 * (N = split_terms_needed(n_total)), x = (B + P) / B and the output is the
 * first D significant digits of x, truncated ("e = 2.ddd...").  The
 * reference forms floor((B+P) 10^n_total / B) with an integer Newton division
 * and converts it by divide and conquer; here the fraction f = x - 2 =
 * (P - B) / B is formed in fixed point to n_total digits plus two guard limbs
 * (Newton reciprocal of B, one product), and its digits are produced by a
 * fixed-point radix-conversion tree that only multiplies: a node holding a
 * fraction y that needs m digits hands y itself (truncated) to its left child
 * for the first k digits and frac(y * 10^k) to its right child for the rest,
 * with k = 19 * 2^j; leaves of at most 19 digits use 128-bit arithmetic.
 * Every node at one depth is one batched multiply.  The approximations are
 * far below 10^-n_total; the 20 guard digits absorb them, and a guard block
 * of all 0s or all 9s (where a carry could reach the printed digits) is
 * reported as an error instead of printed.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "../../include/binary_split.h"
#include "../../include/compute_e.h"
#include "../../include/multiply.h"
#include "engine_internal.h"

using namespace ntte;

/* ---- O(n) big-number kernels ------------------------------------------------ */

struct AddArgs {
    const uint64_t *x, *y;
    uint64_t *z;
    size_t nx, ny, nz;
    int sub;            /* z = x - y (x >= y) instead of x + y */
    uint64_t *gen;      /* one generate mask per 64-limb round */
    uint8_t *blk;       /* block cells, then block carry-ins */
    int *flag;
};

/* z_j = x_j + y_j (or x_j + ~y_j); the carries follow in k_add_apply. */
__global__ void __launch_bounds__(256) k_add_local(AddArgs a)
{
    __shared__ uint8_t wcell[4];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t wave = (size_t)blockIdx.x * 4 + w;
    const size_t base = wave * CW_LIMBS;
    uint8_t agg = 2;
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        uint64_t xv = j < a.nx ? a.x[j] : 0;
        uint64_t yv = j < a.ny ? a.y[j] : 0;
        if (a.sub) yv = ~yv;
        uint64_t r = xv + yv;
        bool g = r < xv, p = r == ~0ull;
        if (j < a.nz) a.z[j] = r;
        else g = false, p = true;
        uint64_t G = __ballot(g), P = __ballot(p);
        if (lane == 0) a.gen[wave * CW_ROUNDS + e] = G;
        agg = cla_compose(agg, round_cell(G, P));
    }
    if (lane == 0) wcell[w] = agg;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint8_t b = 2;
        for (int k = 0; k < 4; k++) b = cla_compose(b, wcell[k]);
        a.blk[blockIdx.x] = b;
    }
}

__global__ void __launch_bounds__(256) k_add_apply(AddArgs a, uint32_t nblk)
{
    __shared__ uint8_t wcell[4];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const size_t wave = (size_t)blockIdx.x * 4 + w;
    const size_t base = wave * CW_LIMBS;
    uint64_t v[CW_ROUNDS], G[CW_ROUNDS], P[CW_ROUNDS];
    uint8_t agg = 2;
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        v[e] = j < a.nz ? a.z[j] : 0;
        G[e] = a.gen[wave * CW_ROUNDS + e];
        P[e] = __ballot(j >= a.nz || v[e] == ~0ull);
        agg = cla_compose(agg, round_cell(G[e], P[e]));
    }
    if (lane == 0) wcell[w] = agg;
    __syncthreads();
    uint8_t c = a.blk[blockIdx.x];
    for (uint32_t k = 0; k < w; k++) c = (wcell[k] & 1) | ((wcell[k] >> 1) & c);
    uint64_t cin = c & 1;
    for (int e = 0; e < CW_ROUNDS; e++) {
        size_t j = base + 64 * e + lane;
        uint64_t X = G[e] | P[e], Y = G[e];
        uint64_t s = X + Y, s2 = s + cin;
        uint64_t cj = ((s2 ^ X ^ Y) >> lane) & 1;
        if (j < a.nz) a.z[j] = v[e] + cj;
        if (j == a.nz - 1 && lane < 63) {
            /* carry out of the top limb: bit lane+1 of the carry vector */
            uint64_t out = ((s2 ^ X ^ Y) >> (lane + 1)) & 1;
            if (out != (uint64_t)a.sub) a.flag[0] = 1;
        }
        uint64_t cout = (s < X) | (s2 < s);
        if (j == a.nz - 1 && lane == 63 && cout != (uint64_t)a.sub) a.flag[0] = 1;
        cin = cout;
    }
    (void)nblk;
}

/* z[0..nz) = x << (64*off + bits), bits < 64, x zero-extended. */
__global__ void k_shl(uint64_t *z, size_t nz, const uint64_t *x, size_t nx, size_t off, int bits)
{
    size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= nz) return;
    uint64_t v = 0;
    if (j >= off) {
        size_t i = j - off;
        uint64_t cur = i < nx ? x[i] : 0;
        uint64_t prev = (i >= 1 && i - 1 < nx) ? x[i - 1] : 0;
        v = bits ? (cur << bits) | (prev >> (64 - bits)) : cur;
    }
    z[j] = v;
}

__global__ void k_fill(uint64_t *z, size_t n, uint64_t v)
{
    size_t j = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (j < n) z[j] = v;
}

/* A leaf of the conversion tree: m <= LEAF_DIGITS digits of the fraction
 * y / 2^(64 p) held in p <= LEAF_LIMBS limbs (y[0] lowest). */
#define LEAF_STEPS 16
#define LEAF_DIGITS (19 * LEAF_STEPS)
#define LEAF_LIMBS 18          /* frac_limbs(LEAF_DIGITS) */
struct DigitLeaf {
    const uint64_t *y;
    uint64_t pos;          /* first digit's index in the output */
    uint32_t m, p;
};

/* One thread per leaf: y * 10^r splits exactly into an integer part (the
 * next r <= 19 digits) and a p-limb fraction, repeated until m digits are out. */
__global__ void __launch_bounds__(256) k_leaf_digits(const DigitLeaf *leaves, size_t n, char *out)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const DigitLeaf L = leaves[i];
    uint64_t y[LEAF_LIMBS];
#pragma unroll
    for (int j = 0; j < LEAF_LIMBS; j++) y[j] = j < (int)L.p ? L.y[j] : 0;
    uint64_t pos = L.pos;
    for (uint32_t left = L.m; left > 0;) {
        const uint32_t r = left < 19 ? left : 19;
        uint64_t p10 = 1;
        for (uint32_t k = 0; k < r; k++) p10 *= 10;
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < LEAF_LIMBS; j++) {
            if (j < (int)L.p) {
                uint64_t lo = y[j] * p10, hi = __umul64hi(y[j], p10);
                lo += carry;
                hi += lo < carry;
                y[j] = lo;
                carry = hi;
            }
        }
        for (int k = (int)r - 1; k >= 0; k--) {
            out[pos + k] = (char)('0' + carry % 10);
            carry /= 10;
        }
        pos += r;
        left -= r;
    }
}

/* Exact digits across the radix tree's truncations.  A child's fraction is
 * its node's, truncated to the child's precision, so it lies at most one unit
 * of its last limb below the exact value; its digits are still exact unless
 * the value beyond its digit range (its tail, in units of its last digit) is
 * smaller than that error, where the truncation would borrow from the last
 * digit.  The left child's tail is its right sibling's fraction; the right
 * child's tail is its node's.  A child with a tail below 2^-110 is rounded up
 * by one unit instead: then it lies at most one unit above, which cannot
 * carry into its digits because the tail is far below 1.  (Otherwise the
 * tail is at least 2^-110 and the truncation error at most 2^-128 in the same
 * units.)  One thread per split node, after the level's products. */
struct TailFix {
    uint64_t *l, *r;             /* the children's fractions (views) */
    uint32_t pl, pr;
    uint32_t parent, lslot, rslot;
};

__device__ __forceinline__ void add_ulp(uint64_t *x, uint32_t n)
{
    for (uint32_t k = 0; k < n; k++)
        if (++x[k] != 0) break;
}

__global__ void __launch_bounds__(256) k_tail_fix(const TailFix *fx, size_t n, uint8_t *flag)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const TailFix f = fx[i];
    const uint64_t r1 = f.r[f.pr - 1], r2 = f.pr >= 2 ? f.r[f.pr - 2] : 0;
    const bool ltiny = r1 == 0 && r2 < ((uint64_t)1 << 18);
    const bool rtiny = flag[f.parent] != 0;
    flag[f.lslot] = ltiny;
    flag[f.rslot] = rtiny;
    if (ltiny) add_ulp(f.l, f.pl);
    if (rtiny) add_ulp(f.r, f.pr);
}

/* ---- host side -------------------------------------------------------------- */

namespace {

double now_s()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Bump allocator over one device allocation. */
struct Arena {
    uint64_t *base = nullptr;
    size_t cap = 0, used = 0;
    int init(size_t words)
    {
        if (hipMalloc((void **)&base, words * 8) != hipSuccess)
            return fail(NTT_ERR_NOMEM, "e pipeline arena (%zu bytes)", words * 8);
        cap = words;
        return NTT_OK;
    }
    uint64_t *take(size_t n)
    {
        n = (n + 1) & ~(size_t)1;   /* keep 16-byte alignment */
        if (used + n > cap) return nullptr;
        uint64_t *p = base + used;
        used += n;
        return p;
    }
    void release()
    {
        if (base) (void)hipFree(base);
        base = nullptr;
    }
};

struct Scratch {
    uint64_t *gen = nullptr;
    uint8_t *blk = nullptr;
    size_t limbs = 0;
    int reserve(size_t n)
    {
        if (n <= limbs) return NTT_OK;
        if (gen || blk) (void)hipStreamSynchronize(g.stream);
        if (gen) (void)hipFree(gen);
        if (blk) (void)hipFree(blk);
        size_t nblk = (n + CB_LIMBS - 1) / CB_LIMBS;
        if (hipMalloc((void **)&gen, nblk * 4 * CW_ROUNDS * 8) != hipSuccess ||
            hipMalloc((void **)&blk, nblk + 64) != hipSuccess)
            return fail(NTT_ERR_NOMEM, "add scratch");
        limbs = nblk * CB_LIMBS;
        return NTT_OK;
    }
    void release()
    {
        if (gen) (void)hipFree(gen);
        if (blk) (void)hipFree(blk);
        gen = nullptr;
        blk = nullptr;
        limbs = 0;
    }
};

/* z = x + y or x - y on the engine stream (nz limbs; sub requires x >= y). */
int dev_addsub(Scratch &sc, uint64_t *z, size_t nz, const uint64_t *x, size_t nx, const uint64_t *y, size_t ny,
               int sub)
{
    int rc = sc.reserve(nz);
    if (rc) return rc;
    uint32_t nblk = (uint32_t)((nz + CB_LIMBS - 1) / CB_LIMBS);
    AddArgs a{x, y, z, nx, ny, nz, sub, sc.gen, sc.blk, g.flag};
    hipLaunchKernelGGL(k_add_local, dim3(nblk), dim3(256), 0, g.stream, a);
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, g.stream, sc.blk, nblk, (uint8_t)sub);
    hipLaunchKernelGGL(k_add_apply, dim3(nblk), dim3(256), 0, g.stream, a, nblk);
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

int dev_shl(uint64_t *z, size_t nz, const uint64_t *x, size_t nx, size_t off, int bits)
{
    hipLaunchKernelGGL(k_shl, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, g.stream, z, nz, x, nx, off, bits);
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

/* Trimmed length of a device number whose value fits `cap` limbs and whose
 * true length is at least cap - probe (reads only the top limbs). */
int dev_trim_len(const uint64_t *x, size_t cap, size_t probe, size_t *n)
{
    size_t k = std::min(cap, probe);
    std::vector<uint64_t> top(k);
    int rc = copy_d2h(top.data(), x + (cap - k), k * 8);
    if (rc) return rc;
    size_t len = cap;
    for (size_t i = k; i-- > 0;) {
        if (top[i]) break;
        len--;
    }
    if (len < cap - k && cap > k) return fail(NTT_ERR_INTEGRITY, "value shorter than its bound by more than %zu", k);
    *n = len;
    return NTT_OK;
}

int mul_dev(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    return ntt_mul_device_async(c, a, na, b, nb);
}

/* Newton reciprocal: X ~ 2^(64 t) / b with b = B'/2^(64 nB) in [1/2, 1) and
 * t = tgoal limbs (X has t + 1 limbs; x = X / 2^(64t) in (1, 2]). */
int newton_recip(Arena &ar, Scratch &sc, const uint64_t *Bn, size_t nB, size_t tgoal, uint64_t **Xout)
{
    uint64_t b1;
    int rc0 = copy_d2h(&b1, Bn + nB - 1, 8);
    if (rc0) return rc0;
    if (!(b1 >> 63)) return fail(NTT_ERR_ARG, "divisor not normalized");
    unsigned __int128 x0 = (~(unsigned __int128)0) / b1;   /* ~ 2^128 / b1, < 2^65 */
    uint64_t h[2] = {(uint64_t)x0, (uint64_t)(x0 >> 64)};
    size_t t0 = 1;
    uint64_t *X = ar.take(2);
    if (!X) return fail(NTT_ERR_NOMEM, "arena");
    rc0 = copy_h2d(X, h, 16);
    if (rc0) return rc0;
    /* scratch for the iterations: sized for the last one */
    uint64_t *S = ar.take(2 * tgoal + 4), *U = ar.take(3 * tgoal + 8), *W = ar.take(tgoal + 4),
             *Xn = ar.take(tgoal + 4), *X2 = ar.take(tgoal + 4);
    if (!S || !U || !W || !Xn || !X2) return fail(NTT_ERR_NOMEM, "arena");
    uint64_t *cur = X, *next = Xn;
    while (t0 < tgoal) {
        size_t t = std::min(2 * t0, tgoal);
        size_t nBt = std::min(t + 1, nB);
        const uint64_t *Bt = Bn + (nB - nBt);
        int rc = ntt_sqr_device_async(S, cur, t0 + 1);                       /* x^2: 2t0+2 limbs */
        if (!rc) rc = mul_dev(U, Bt, nBt, S, 2 * t0 + 2);                     /* b x^2 */
        if (rc) return rc;
        size_t sh = nBt + 2 * t0 - t;                                         /* U >> 64 sh */
        size_t nV = nBt + 2 * t0 + 2 - sh;                                    /* = t + 2 */
        rc = dev_shl(W, t + 2, cur, t0 + 1, t - t0, 1);                       /* 2x at precision t */
        if (!rc) rc = dev_addsub(sc, next, t + 2, W, t + 2, U + sh, nV, 1);  /* 2x - b x^2 */
        if (rc) return rc;
        t0 = t;
        std::swap(cur, next);
        if (cur == X) std::swap(cur, next);   /* keep the 2-limb seed buffer out of rotation */
        if (next == X) next = X2;
    }
    *Xout = cur;
    return NTT_OK;
}

struct RNode {
    uint64_t *y;         /* fraction limbs (p of them, value y / 2^(64 p)) */
    size_t p;
    size_t m, pos;
    int depth;
    uint32_t slot;       /* its tail flag (see k_tail_fix) */
};

size_t frac_limbs(size_t digits) { return (size_t)ceil((double)digits * 3.3219280948873623 / 64.0) + 2; }

}  // namespace

/* Digits 1..M of the fraction F / 2^(64 tgoal) (device, tgoal limbs),
 * truncated, into d_digits[0, M) on the device: the fixed-point radix tree
 * (a node with fraction y and m digits gives its left child y for the first
 * k = 19 * 2^j digits and its right child frac(y * 10^k) for the rest; one
 * batched multiply per depth), then 19-digit leaves in 128-bit arithmetic.
 * Powers of 10 come from the arena; the tree's products from one of its own. */
int fraction_digits(Arena &ar, uint64_t *F, size_t tgoal, size_t M, char *d_digits)
{
    int rc = NTT_OK;
    Arena zr;   /* products of the conversion tree */
    {
        /* powers 10^(19 * 2^j) */
        std::vector<uint64_t *> pw;
        std::vector<size_t> pwn;
        uint64_t *t1p = ar.take(2);
        uint64_t ten19 = 10000000000000000000ull;
        if (!t1p) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = copy_h2d(t1p, &ten19, 8);
        pw.push_back(t1p);
        pwn.push_back(1);
        size_t leaves = (M + 18) / 19;
        while (!rc && ((size_t)1 << pw.size()) < leaves) {
            size_t n = pwn.back();
            uint64_t *q = ar.take(2 * n);
            if (!q) {
                rc = fail(NTT_ERR_NOMEM, "arena");
                break;
            }
            rc = ntt_sqr_device_async(q, pw.back(), n);
            pw.push_back(q);
            pwn.push_back(2 * n);
        }
        /* size the conversion tree's products, then plan it level by level */
        if (!rc) {
            size_t words = 0;
            std::vector<std::pair<size_t, size_t>> lv{{tgoal, M}}, nx;   /* (p, m) */
            while (!lv.empty()) {
                nx.clear();
                for (auto &pm : lv) {
                    if (pm.second <= LEAF_DIGITS) continue;
                    size_t nl = (pm.second + 18) / 19, K = 1;
                    int j = 0;
                    while (K * 2 < nl) K *= 2, j++;
                    size_t k = 19 * K;
                    words += ((pm.first + pwn[j]) + 1) & ~(size_t)1;
                    nx.push_back({std::min(frac_limbs(k), pm.first), k});
                    nx.push_back({std::min(frac_limbs(pm.second - k), pm.first), pm.second - k});
                }
                lv.swap(nx);
            }
            rc = zr.init(words + 64);
        }
        /* tail flags: the root's is 0 (beyond its M digits lies only the
         * approximation error, which the caller's guard digits absorb) */
        uint8_t *d_flag = nullptr;
        const size_t max_nodes = 4 * ((M + LEAF_DIGITS - 1) / LEAF_DIGITS) + 8;
        if (hipMalloc((void **)&d_flag, max_nodes) != hipSuccess) rc = fail(NTT_ERR_NOMEM, "tail flags");
        const uint8_t zero = 0;
        if (!rc) rc = copy_h2d(d_flag, &zero, 1);
        uint32_t nslots = 1;
        std::vector<std::vector<TailFix>> fixes;   /* host copies, kept until the final sync */
        TailFix *d_fix = nullptr;                  /* every level's fixes, one after another */
        size_t fix_used = 0;
        const size_t fix_cap = max_nodes / 2 + 8;
        if (!rc && hipMalloc((void **)&d_fix, fix_cap * sizeof(TailFix)) != hipSuccess)
            rc = fail(NTT_ERR_NOMEM, "tail fixes");
        std::vector<RNode> level{{F, tgoal, M, 0, 0, 0}};
        std::vector<DigitLeaf> leafv;
        while (!rc && !level.empty()) {
            std::vector<RNode> next;
            std::vector<ntt_mul_desc> descs;
            fixes.emplace_back();
            std::vector<TailFix> &fx = fixes.back();
            for (const RNode &nd : level) {
                if (nd.m <= LEAF_DIGITS) {
                    if (nd.p > LEAF_LIMBS) {
                        rc = fail(NTT_ERR_ARG, "leaf of %zu limbs (internal)", nd.p);
                        break;
                    }
                    leafv.push_back({nd.y, nd.pos, (uint32_t)nd.m, (uint32_t)nd.p});
                    continue;
                }
                size_t nl = (nd.m + 18) / 19, K = 1;
                int j = 0;
                while (K * 2 < nl) {
                    K *= 2;
                    j++;
                }
                size_t k = 19 * K;
                size_t pl = std::min(frac_limbs(k), nd.p), pr = std::min(frac_limbs(nd.m - k), nd.p);
                uint64_t *Z = zr.take(nd.p + pwn[j]);
                if (!Z || nslots + 2 > max_nodes) {
                    rc = fail(NTT_ERR_NOMEM, "arena");
                    break;
                }
                const uint32_t ls = nslots++, rs = nslots++;
                next.push_back({nd.y + (nd.p - pl), pl, k, nd.pos, nd.depth + 1, ls});
                ntt_mul_desc dd{nd.y, pw[j], Z, nullptr, (uint32_t)nd.p, (uint32_t)pwn[j], 0,
                                (uint32_t)(nd.p + pwn[j])};
                descs.push_back(dd);
                next.push_back({Z + (nd.p - pr), pr, nd.m - k, nd.pos + k, nd.depth + 1, rs});
                fx.push_back({nd.y + (nd.p - pl), Z + (nd.p - pr), (uint32_t)pl, (uint32_t)pr, nd.slot, ls, rs});
            }
            if (!rc && !descs.empty()) rc = ntt_mul_batch_device_async(descs.data(), descs.size());
            if (!rc && !fx.empty()) {
                if (fix_used + fx.size() > fix_cap) rc = fail(NTT_ERR_NOMEM, "tail fixes (internal)");
                TailFix *dst = d_fix + fix_used;
                if (!rc && hipMemcpyAsync(dst, fx.data(), fx.size() * sizeof(TailFix), hipMemcpyHostToDevice,
                                          g.stream) != hipSuccess)
                    rc = fail(NTT_ERR_HIP, "tail fix upload");
                if (!rc) {
                    hipLaunchKernelGGL(k_tail_fix, dim3((unsigned)((fx.size() + 255) / 256)), dim3(256), 0, g.stream,
                                       dst, fx.size(), d_flag);
                    if (hipGetLastError() != hipSuccess) rc = fail(NTT_ERR_HIP, "k_tail_fix");
                }
                fix_used += fx.size();
            }
            level.swap(next);
        }
        DigitLeaf *d_leaves = nullptr;
        if (!rc && hipMalloc((void **)&d_leaves, leafv.size() * sizeof(DigitLeaf)) != hipSuccess)
            rc = fail(NTT_ERR_NOMEM, "leaves");
        if (!rc && hipMemcpyAsync(d_leaves, leafv.data(), leafv.size() * sizeof(DigitLeaf), hipMemcpyHostToDevice,
                                  g.stream) != hipSuccess)
            rc = fail(NTT_ERR_HIP, "leaf upload");
        if (!rc) {
            hipLaunchKernelGGL(k_leaf_digits, dim3((unsigned)((leafv.size() + 255) / 256)), dim3(256), 0, g.stream,
                               d_leaves, leafv.size(), d_digits);
            if (hipGetLastError() != hipSuccess) rc = fail(NTT_ERR_HIP, "leaf kernel");
        }
        if (!rc) rc = sync_check();
        if (rc) (void)hipStreamSynchronize(g.stream);
        if (d_leaves) (void)hipFree(d_leaves);
        if (d_fix) (void)hipFree(d_fix);
        if (d_flag) (void)hipFree(d_flag);
    }
    zr.release();
    return rc;
}

/* ---- integer to decimal on the device --------------------------------------
 * The digits of N are the first D digits of the fraction f = (N + 1/2) / 10^D
 * (D = 19 K >= digits of N), produced by the same fixed-point radix tree.
 * The half unit makes the digits after position D read 5000... exactly, so
 * the approximation error (far below 10^-(D+20)) shows up only as 5000.. or
 * 4999.. in the guard digits and can never reach digit D; anything else in
 * the guard digits, or a mismatch of the last 19 digits against N mod 10^19
 * (computed on the host), is reported as an integrity error. */

/* N mod 10^19 from the host limbs, top down. */
static uint64_t mod_ten19(const uint64_t *h, size_t n)
{
    const uint64_t T = 10000000000000000000ull;
    unsigned __int128 r = 0;
    for (size_t i = n; i-- > 0;) r = ((r << 64) | h[i]) % T;
    return (uint64_t)r;
}

__global__ void k_or_limb0(uint64_t *x, uint64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) x[0] |= v;
}

extern "C" int ntt_to_decimal(const uint64_t *h, size_t n, char *out, size_t cap, size_t *len)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((n && !h) || !out) return fail(NTT_ERR_ARG, "null buffer");
    while (n && !h[n - 1]) n--;
    if (n == 0) {
        if (cap < 2) return fail(NTT_ERR_ARG, "out_cap too small");
        out[0] = '0';
        out[1] = 0;
        if (len) *len = 1;
        return NTT_OK;
    }
    const size_t bits = 64 * (n - 1) + 64 - __builtin_clzll(h[n - 1]);
    const size_t dmax = (size_t)((double)bits * 0.30103) + 1;      /* >= digits of N (0.30103 > log10 2) */
    if (cap < dmax + 1) return fail(NTT_ERR_ARG, "out_cap %zu below the %zu-digit bound + 1", cap, dmax);
    const size_t K = (dmax + 18) / 19, D = 19 * K, G = 20, M = D + G;
    const size_t tgoal = frac_limbs(M);
    Arena ar;
    Scratch sc;
    char *d_digits = nullptr;
    int rc = ar.init(40 * (n + tgoal) + 4096);
    if (!rc) rc = sc.reserve(2 * (n + tgoal) + 64);
    uint64_t *dN = rc ? nullptr : ar.take(n);
    if (!rc && !dN) rc = fail(NTT_ERR_NOMEM, "arena");
    if (!rc) rc = copy_h2d(dN, h, n * 8);
    /* T = 10^D = prod of 10^(19 * 2^j) over the set bits j of K */
    uint64_t *T = nullptr;
    size_t nT = 0;
    if (!rc) {
        const uint64_t ten19 = 10000000000000000000ull;
        uint64_t *pw = ar.take(2);
        size_t npw = 1;
        if (!pw) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = copy_h2d(pw, &ten19, 8);
        for (size_t k = K; !rc && k; k >>= 1) {
            if (k & 1) {
                if (!T) {
                    T = pw;
                    nT = npw;
                } else {
                    uint64_t *q = ar.take(nT + npw);
                    if (!q) rc = fail(NTT_ERR_NOMEM, "arena");
                    if (!rc) rc = mul_dev(q, T, nT, pw, npw);
                    T = q;
                    nT += npw;
                }
            }
            if (!rc && k > 1) {
                uint64_t *q = ar.take(2 * npw);
                if (!q) rc = fail(NTT_ERR_NOMEM, "arena");
                if (!rc) rc = ntt_sqr_device_async(q, pw, npw);
                pw = q;
                npw *= 2;
            }
        }
        /* the powers carry up to ~1.5 % leading zero limbs */
        if (!rc) rc = dev_trim_len(T, nT, nT / 48 + 8, &nT);
    }
    /* normalized divisor 2 * 10^D and its reciprocal */
    uint64_t *Bn = nullptr, *X = nullptr;
    size_t nB = 0;
    int sh = 0;
    if (!rc) {
        uint64_t *M2 = ar.take(nT + 1);
        Bn = ar.take(nT + 1);
        if (!M2 || !Bn) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = dev_shl(M2, nT + 1, T, nT, 0, 1);
        uint64_t top = 0;
        if (!rc) rc = copy_d2h(&top, M2 + nT, 8);
        nB = top ? nT + 1 : nT;
        if (!rc && !top) rc = copy_d2h(&top, M2 + nT - 1, 8);
        if (!rc) {
            sh = __builtin_clzll(top);
            rc = dev_shl(Bn, nB, M2, nB, 0, sh);
        }
    }
    if (!rc) rc = newton_recip(ar, sc, Bn, nB, tgoal, &X);
    /* F = ((2N + 1) << sh) * X >> 64 nB: f to tgoal limbs */
    uint64_t *F = nullptr;
    if (!rc) {
        uint64_t *Y = ar.take(n + 2), *Fz = ar.take(n + 2 + tgoal + 1);
        if (!Y || !Fz) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = dev_shl(Y, n + 2, dN, n, (size_t)(sh + 1) >> 6, (sh + 1) & 63);   /* 2N << sh */
        if (!rc) {
            hipLaunchKernelGGL(k_or_limb0, dim3(1), dim3(64), 0, g.stream, Y, (uint64_t)1 << sh);
            if (hipGetLastError() != hipSuccess) rc = fail(NTT_ERR_HIP, "k_or_limb0");
        }
        if (!rc) rc = mul_dev(Fz, Y, n + 2, X, tgoal + 1);
        F = Fz ? Fz + nB : nullptr;
    }
    if (!rc && hipMalloc((void **)&d_digits, M + 64) != hipSuccess) rc = fail(NTT_ERR_NOMEM, "digit buffer");
    if (!rc) rc = fraction_digits(ar, F, tgoal, M, d_digits);
    std::vector<char> dig;
    if (!rc) {
        dig.resize(M);
        rc = copy_d2h(dig.data(), d_digits, M);
    }
    if (!rc) {
        /* guard digits: 5 then 0s, or 4 then 9s */
        const char g0 = dig[D], fill = g0 == '5' ? '0' : '9';
        bool ok = g0 == '5' || g0 == '4';
        for (size_t k = D + 1; ok && k < D + 12; k++) ok = dig[k] == fill;
        if (!ok) rc = fail(NTT_ERR_INTEGRITY, "decimal conversion: guard digits %.12s", dig.data() + D);
    }
    if (!rc) {
        size_t z = 0;
        while (z + 1 < D && dig[z] == '0') z++;
        const size_t nd = D - z;
        /* the last 19 digits against N mod 10^19 */
        uint64_t r = 0;
        for (size_t k = D >= 19 ? D - 19 : 0; k < D; k++) r = r * 10 + (uint64_t)(dig[k] - '0');
        if (r != mod_ten19(h, n) || nd > dmax || dig[z] == '0')
            rc = fail(NTT_ERR_INTEGRITY, "decimal conversion: digits disagree with N mod 10^19");
        if (!rc) {
            memcpy(out, dig.data() + z, nd);
            out[nd] = 0;
            if (len) *len = nd;
        }
    }
    if (d_digits) (void)hipFree(d_digits);
    if (rc) (void)hipStreamSynchronize(g.stream);
    sc.release();
    ar.release();
    return rc;
}

extern "C" int ntt_compute_e(long n_digits, int batched_tree, char *out, size_t out_cap, ntt_e_stats *st)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) {
        int rc = ntt_engine_init(0);
        if (rc) return rc;
    }
    if (n_digits < 1 || !out || out_cap < (size_t)n_digits + 2)
        return fail(NTT_ERR_ARG, "bad arguments (need out_cap >= digits + 2)");
    ntt_e_stats stats{};
    double t0 = now_s();
    const long n_total = n_digits + 20;
    const long n_terms = split_terms_needed(n_total);
    stats.n_terms = n_terms;
    unsigned long disp0 = g.dispatches;
    int rc = NTT_OK;
    Arena ar;
    Scratch sc;
    ntt_split_tree *tree = nullptr;
    uint64_t *dP = nullptr, *dB = nullptr;
    size_t capP = 0, capB = 0;
    std::vector<uint64_t> hP, hB;

    /* 1. the product tree */
    if (batched_tree) {
        rc = ntt_split_tree_create(1, n_terms + 1, 1, &tree);
        if (!rc) rc = ntt_split_tree_run(tree);
        if (!rc) rc = ntt_split_tree_device_result(tree, (const uint64_t **)&dP, &capP, (const uint64_t **)&dB,
                                                   &capB);
    } else {
        SplitResult sr;
        binary_split(&sr, 1, n_terms + 1, NULL);
        capP = sr.P.n_limbs;
        capB = sr.B.n_limbs;
        hP.assign(sr.P.limbs, sr.P.limbs + capP);
        hB.assign(sr.B.limbs, sr.B.limbs + capB);
        split_free(&sr);
    }
    double t1 = now_s();
    stats.t_tree = t1 - t0;
    const size_t tgoal = frac_limbs((size_t)n_total);
    const size_t M = (size_t)n_total - 1;            /* fractional digits computed */
    const size_t n_leaves = (M + 18) / 19;
    if (!rc) rc = ar.init(capP + capB + 4 * (capP + capB) + 12 * tgoal + 4 * n_leaves + 4096);
    if (!rc && !batched_tree) {
        dP = ar.take(capP);
        dB = ar.take(capB);
        if (!dP || !dB) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = copy_h2d(dP, hP.data(), capP * 8);
        if (!rc) rc = copy_h2d(dB, hB.data(), capB * 8);
    }
    /* 2. normalized divisor and Newton reciprocal */
    size_t nB = 0, nP = 0;
    if (!rc) rc = dev_trim_len(dB, capB, 8, &nB);
    if (!rc) rc = dev_trim_len(dP, capP, 8, &nP);
    if (!rc) rc = sc.reserve(std::max(nP, tgoal) + 8);
    int s = 0;
    uint64_t *Bn = nullptr, *X = nullptr;
    if (!rc) {
        uint64_t top;
        rc = copy_d2h(&top, dB + nB - 1, 8);
        s = __builtin_clzll(top);
        Bn = ar.take(nB);
        if (!Bn) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = dev_shl(Bn, nB, dB, nB, 0, s);
    }
    if (!rc) rc = newton_recip(ar, sc, Bn, nB, tgoal, &X);
    double t2 = now_s();
    stats.t_recip = t2 - t1;
    /* 3. F = ((P - B) << s) * X >> 64 nB: the fraction f = x - 2 to tgoal limbs */
    uint64_t *Fz = nullptr;
    if (!rc) {
        uint64_t *Dd = ar.take(nP + 1), *M = ar.take(nP + 2);
        Fz = ar.take(nP + 2 + tgoal + 1);
        if (!Dd || !M || !Fz) rc = fail(NTT_ERR_NOMEM, "arena");
        if (!rc) rc = dev_addsub(sc, Dd, nP, dP, nP, dB, nB, 1);
        if (!rc) rc = dev_shl(M, nP + 1, Dd, nP, 0, s);
        if (!rc) rc = mul_dev(Fz, M, nP + 1, X, tgoal + 1);
        if (!rc) rc = sync_check();
    }
    uint64_t *F = Fz ? Fz + nB : nullptr;   /* tgoal limbs of fraction */
    double t3 = now_s();
    stats.t_div = t3 - t2;
    /* 4. radix conversion of the n_total - 1 fractional digits */
    char *d_digits = nullptr;
    if (!rc && hipMalloc((void **)&d_digits, M + 64) != hipSuccess) rc = fail(NTT_ERR_NOMEM, "digit buffer");
    if (!rc) rc = fraction_digits(ar, F, tgoal, M, d_digits);
    std::vector<char> digits;
    if (!rc) {
        digits.resize(M);
        rc = copy_d2h(digits.data(), d_digits, M);
    }
    double t4 = now_s();
    stats.t_radix = t4 - t3;
    if (!rc) {
        /* the 20 guard digits must not be all 0s or all 9s */
        const char *gd = digits.data() + (n_digits - 1);
        bool all0 = true, all9 = true;
        for (int k = 0; k < 20; k++) {
            all0 &= gd[k] == '0';
            all9 &= gd[k] == '9';
        }
        if (all0 || all9) rc = fail(NTT_ERR_INTEGRITY, "guard digits ambiguous (all %c)", all0 ? '0' : '9');
    }
    if (!rc) {
        out[0] = '2';
        out[1] = '.';
        memcpy(out + 2, digits.data(), (size_t)n_digits - 1);
        out[n_digits + 1] = 0;
    }
    if (d_digits) (void)hipFree(d_digits);
    if (tree) ntt_split_tree_destroy(tree);
    sc.release();
    ar.release();
    stats.t_total = now_s() - t0;
    stats.products = g.dispatches - disp0;
    if (st) *st = stats;
    return rc;
}
