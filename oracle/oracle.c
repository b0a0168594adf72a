/*
 * oracle.c — CPU restatement of the reference's CRT-NTT multiply path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the product library never uses it.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* The library's three primes 2^60 - c*2^32 + 1 (c = 96, 120, 168), each
 * c'*2^k + 1 with c' odd (lazy values below 16p fit a u64 word).
 * Generators were found by testing g^((p-1)/q) != 1 for every prime q | p-1;
 * tests/test_oracle.py (test_primes_form_and_generators) re-proves primality, the factorisation of p-1
 * and the generator property from scratch. */
static const uint64_t PR[ORC_NPRIMES]  = {0x0fffffa000000001ull, 0x0fffff8800000001ull,
                                          0x0fffff5800000001ull};
static const int      PK[ORC_NPRIMES]  = {37, 35, 35};
static const uint64_t PG[ORC_NPRIMES]  = {3, 3, 6};

uint64_t orc_prime(int i) { return PR[i]; }
int orc_prime_k(int i) { return PK[i]; }
uint64_t orc_prime_gen(int i) { return PG[i]; }

/* Modular primitives: same semantics as the reference's addmod/submod/mulmod
 * (ntt/lib/primes.h:193-217) and modpow (primes.h:268-279), written with a
 * plain u128 remainder (the textbook definition). */
uint64_t orc_addmod(uint64_t a, uint64_t b, uint64_t p)
{
    uint64_t s = a + b;             /* a,b < p < 2^60: no wrap */
    return s >= p ? s - p : s;
}
uint64_t orc_submod(uint64_t a, uint64_t b, uint64_t p) { return a >= b ? a - b : a + p - b; }
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1) r = orc_mulmod(r, a, p);
        a = orc_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}
uint64_t orc_invmod(uint64_t a, uint64_t p) { return orc_powmod(a, p - 2, p); }

uint64_t orc_root(int i, int log_n)
{
    if (log_n < 0 || log_n > PK[i]) return 0;
    return orc_powmod(PG[i], (PR[i] - 1) >> log_n, PR[i]);
}

/* Shoup multiply (ntt/lib/shoup.h:45-57 make_shoup, :69-112 shoup_mul):
 * q = hi64(a*w'), r = a*w - q*p in [0,2p), one conditional subtract. */
uint64_t orc_shoup_pre(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }
uint64_t orc_shoup_mul_lazy(uint64_t a, uint64_t w, uint64_t wp, uint64_t p)
{
    uint64_t q = (uint64_t)(((u128)a * wp) >> 64);
    return a * w - q * p;
}
uint64_t orc_shoup_mul(uint64_t a, uint64_t w, uint64_t wp, uint64_t p)
{
    uint64_t r = orc_shoup_mul_lazy(a, w, wp, p);
    return r >= p ? r - p : r;
}

/* Bit reversal then radix-2 CT-DIT stages, twiddle tw[j*stride] = w^(j*stride),
 * addmod/submod after every butterfly: ntt/ref/ntt_cpu.c:128-202
 * (bit_reverse_perm, ct_dit_loop, ntt_forward, ntt_inverse). */
static void bit_reverse(uint64_t *a, size_t n, int log_n)
{
    for (size_t i = 0; i < n; i++) {
        size_t r = 0, x = i;
        for (int b = 0; b < log_n; b++) { r = (r << 1) | (x & 1); x >>= 1; }
        if (i < r) { uint64_t t = a[i]; a[i] = a[r]; a[r] = t; }
    }
}

static void ct_dit(uint64_t *a, int log_n, uint64_t w, uint64_t p)
{
    size_t n = (size_t)1 << log_n, half_n = n >> 1;
    if (n < 2) return;
    /* tw[k] = w^k for k < n/2 (ntt_cpu.c ntt_alloc_twiddles), with Shoup
     * partners so the butterfly multiply is exact without a u128 remainder. */
    uint64_t *tw = (uint64_t *)malloc(half_n * sizeof(uint64_t));
    uint64_t *twp = (uint64_t *)malloc(half_n * sizeof(uint64_t));
    tw[0] = 1;
    for (size_t k = 1; k < half_n; k++) tw[k] = orc_mulmod(tw[k - 1], w, p);
    for (size_t k = 0; k < half_n; k++) twp[k] = orc_shoup_pre(tw[k], p);
    bit_reverse(a, n, log_n);
    for (size_t len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1, stride = n / len;
        for (size_t k = 0; k < n; k += len)
            for (size_t j = 0; j < half; j++) {
                uint64_t u = a[k + j];
                uint64_t t = orc_shoup_mul(a[k + j + half], tw[j * stride], twp[j * stride], p);
                a[k + j] = orc_addmod(u, t, p);
                a[k + j + half] = orc_submod(u, t, p);
            }
    }
    free(tw);
    free(twp);
}

/* The same CT-DIT transform on all host cores, for the CPU baseline at full
 * size (ntt_cpu.c:158-202's arithmetic unchanged: bit reversal, then for
 * len = 2, 4, .., n the butterflies (u + t, u - t), t = a[k+j+half] *
 * w^(j n/len) by Shoup with a full reduction).  Differences are of schedule
 * only: the twiddle table w^k (k < n/2) with its Shoup partners is built once
 * by the caller (tw_build) and reused across transforms; the stages with
 * len <= ORC_BLK run block by block while a block sits in cache; every larger
 * stage splits its butterflies over the threads. */
#define ORC_BLK ((size_t)1 << 14)

typedef struct { uint64_t *w, *wp; size_t half; } TwTab;

static int tw_build(TwTab *t, uint64_t w, uint64_t p, int log_n, int nt)
{
    const size_t half = ((size_t)1 << log_n) >> 1;
    t->half = half;
    t->w = (uint64_t *)malloc((half ? half : 1) * sizeof(uint64_t));
    t->wp = (uint64_t *)malloc((half ? half : 1) * sizeof(uint64_t));
    if (!t->w || !t->wp) { free(t->w); free(t->wp); return -1; }
    const uint64_t wsh = orc_shoup_pre(w, p);
    const size_t chunk = 1 << 16;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (size_t c0 = 0; c0 < half; c0 += chunk) {
        uint64_t x = orc_powmod(w, c0, p);           /* w^c0, then successive products */
        const size_t c1 = c0 + chunk < half ? c0 + chunk : half;
        for (size_t k = c0; k < c1; k++) {
            t->w[k] = x;
            t->wp[k] = orc_shoup_pre(x, p);
            x = orc_shoup_mul(x, w, wsh, p);
        }
    }
    (void)nt;
    return 0;
}

static void tw_free(TwTab *t) { free(t->w); free(t->wp); }

static inline void dit_stage_range(uint64_t *a, size_t k0, size_t k1, size_t len, size_t stride, const TwTab *t,
                                   uint64_t p)
{
    const size_t half = len >> 1;
    for (size_t k = k0; k < k1; k += len)
        for (size_t j = 0; j < half; j++) {
            const uint64_t u = a[k + j];
            const uint64_t v = orc_shoup_mul(a[k + j + half], t->w[j * stride], t->wp[j * stride], p);
            a[k + j] = orc_addmod(u, v, p);
            a[k + j + half] = orc_submod(u, v, p);
        }
}

static void ct_dit_par(uint64_t *a, int log_n, const TwTab *t, uint64_t p, int nt)
{
    const size_t n = (size_t)1 << log_n;
    if (n < 2) return;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (size_t i = 0; i < n; i++) {            /* each pair swapped once, by its smaller index */
        size_t r = 0, x = i;
        for (int b = 0; b < log_n; b++) { r = (r << 1) | (x & 1); x >>= 1; }
        if (i < r) { uint64_t v = a[i]; a[i] = a[r]; a[r] = v; }
    }
    const size_t blk = n < ORC_BLK ? n : ORC_BLK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (size_t b0 = 0; b0 < n; b0 += blk)
        for (size_t len = 2; len <= blk; len <<= 1) dit_stage_range(a, b0, b0 + blk, len, n / len, t, p);
    for (size_t len = blk << 1; len <= n; len <<= 1) {
        const size_t half = len >> 1, stride = n / len;
        /* butterflies (k, j) over all groups k and offsets j, in pieces */
        const size_t piece = half < 4096 ? half : 4096;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
        for (size_t q = 0; q < n / 2; q += piece) {
            const size_t k = (q / half) * len, j0 = q % half;
            for (size_t j = j0; j < j0 + piece; j++) {
                const uint64_t u = a[k + j];
                const uint64_t v = orc_shoup_mul(a[k + j + half], t->w[j * stride], t->wp[j * stride], p);
                a[k + j] = orc_addmod(u, v, p);
                a[k + j + half] = orc_submod(u, v, p);
            }
        }
    }
    (void)nt;
}

void orc_ntt_forward(uint64_t *a, int log_n, int pi)
{
    ct_dit(a, log_n, orc_root(pi, log_n), PR[pi]);
}

void orc_ntt_inverse(uint64_t *a, int log_n, int pi)
{
    uint64_t p = PR[pi];
    ct_dit(a, log_n, orc_invmod(orc_root(pi, log_n), p), p);
    size_t n = (size_t)1 << log_n;
    uint64_t ninv = orc_invmod(n % p, p), ninvp = orc_shoup_pre(ninv, p);
    for (size_t i = 0; i < n; i++) a[i] = orc_shoup_mul(a[i], ninv, ninvp, p);
}

void orc_dft_naive(const uint64_t *in, uint64_t *out, int n, uint64_t w, uint64_t p)
{
    for (int k = 0; k < n; k++) {
        uint64_t wk = orc_powmod(w, (uint64_t)k, p), acc = 0, wkj = 1;
        for (int j = 0; j < n; j++) {
            acc = orc_addmod(acc, orc_mulmod(in[j], wkj, p), p);
            wkj = orc_mulmod(wkj, wk, p);
        }
        out[k] = acc;
    }
}

/* Garner mixed-radix CRT in the reference's order (ntt/lib/garner.hip:135-173,
 * garner_one): a0 = r0; a1 = (r1 - a0 mod p1) * p0^-1 mod p1;
 * a2 = (r2 - (a0 + a1*p0) mod p2) * (p0*p1)^-1 mod p2; x = a0 + a1*p0 + a2*p0*p1.
 * The reference runs the same recurrence over four residues into a U256. */
typedef struct { uint64_t inv01, inv02, inv12; } GarnerK;

static GarnerK garner_consts(void)
{
    GarnerK k;
    k.inv01 = orc_invmod(PR[0] % PR[1], PR[1]);
    k.inv02 = orc_invmod(PR[0] % PR[2], PR[2]);
    k.inv12 = orc_invmod(PR[1] % PR[2], PR[2]);
    return k;
}

static void garner3_k(uint64_t r0, uint64_t r1, uint64_t r2, const GarnerK *K, uint64_t out[3]);

void orc_garner3(uint64_t r0, uint64_t r1, uint64_t r2, uint64_t out[3])
{
    const GarnerK K = garner_consts();
    garner3_k(r0, r1, r2, &K, out);
}

static void garner3_k(uint64_t r0, uint64_t r1, uint64_t r2, const GarnerK *K, uint64_t out[3])
{
    const uint64_t p0 = PR[0], p1 = PR[1], p2 = PR[2];
    const uint64_t inv01 = K->inv01, inv02 = K->inv02, inv12 = K->inv12;
    uint64_t a0 = r0;
    uint64_t a1 = orc_mulmod(orc_submod(r1, a0 % p1, p1), inv01, p1);
    uint64_t base2 = orc_addmod(a0 % p2, orc_mulmod(a1, p0 % p2, p2), p2);
    uint64_t a2 = orc_mulmod(orc_mulmod(orc_submod(r2, base2, p2), inv02, p2), inv12, p2);
    /* x = a0 + a1*p0 + a2*(p0*p1): schoolbook on 64-bit words. */
    u128 p01 = (u128)p0 * p1;
    u128 t = (u128)a1 * p0 + a0;
    uint64_t w0 = (uint64_t)t, w1 = (uint64_t)(t >> 64), w2 = 0;
    u128 m_lo = (u128)a2 * (uint64_t)p01;
    u128 m_hi = (u128)a2 * (uint64_t)(p01 >> 64);
    u128 s = (u128)w0 + (uint64_t)m_lo;
    w0 = (uint64_t)s;
    s = (s >> 64) + w1 + (uint64_t)(m_lo >> 64) + (uint64_t)m_hi;
    w1 = (uint64_t)s;
    w2 = (uint64_t)(s >> 64) + (uint64_t)(m_hi >> 64);
    out[0] = w0; out[1] = w1; out[2] = w2;
}

/* 16-bit split: the build spec's coefficient base (ntt/lib/arith/bigint.c:286-297
 * bigint_scatter writes one whole limb per coefficient; here a u64 limb is read
 * as four little-endian 16-bit pieces, i.e. two 32-bit limbs split in two). */
void orc_split16(const uint64_t *limbs, size_t n_limbs, uint64_t *coeffs, size_t n_coeffs)
{
    for (size_t k = 0; k < n_coeffs; k++) {
        size_t li = k >> 2;
        coeffs[k] = li < n_limbs ? (limbs[li] >> (16 * (k & 3))) & 0xffff : 0;
    }
}

/* Carry assembly (ntt/lib/arith/bigint.c:333-357, bigint_gather): every
 * coefficient is added at its offset with a running carry.  The reference adds
 * a U256 at limb offset i; here a 3-word coefficient lands at bit 16k. */
uint64_t orc_pack16(const uint64_t *X, size_t n_coeffs, uint64_t *out, size_t n_out)
{
    memset(out, 0, n_out * sizeof(uint64_t));
    uint64_t spill = 0;
    for (size_t k = 0; k < n_coeffs; k++) {
        size_t li = k >> 2;
        int sh = 16 * (int)(k & 3);
        /* x << sh spans up to 4 words */
        uint64_t w[4];
        w[0] = X[3 * k] << sh;
        w[1] = sh ? (X[3 * k + 1] << sh) | (X[3 * k] >> (64 - sh)) : X[3 * k + 1];
        w[2] = sh ? (X[3 * k + 2] << sh) | (X[3 * k + 1] >> (64 - sh)) : X[3 * k + 2];
        w[3] = sh ? (X[3 * k + 2] >> (64 - sh)) : 0;
        u128 carry = 0;
        size_t pos = li;
        for (int t = 0; t < 4; t++, pos++) {
            if (pos >= n_out) { spill |= w[t] | (uint64_t)carry; carry = 0; continue; }
            u128 s = (u128)out[pos] + w[t] + carry;
            out[pos] = (uint64_t)s;
            carry = s >> 64;
        }
        while (carry) {
            if (pos >= n_out) { spill |= (uint64_t)carry; break; }
            u128 s = (u128)out[pos] + carry;
            out[pos++] = (uint64_t)s;
            carry = s >> 64;
        }
    }
    return spill;
}

/* Carry-lookahead form (ntt/lib/transfer_core.h:165-252): per limb j the local
 * sum s_j = sum_t X[4j+t] << 16t is split into lo_j + hi_j*2^64; t_j = lo_j +
 * hi_{j-1} gives a thin generate bit, prop_j = (res_j == 2^64-1), and the carry
 * into limb j is the prefix of (g,p) under (Glo,Plo)o(Ghi,Phi) = (Ghi|Phi&Glo,
 * Plo&Phi).  The scan here is sequential; the monoid is what is pinned. */
int orc_pack16_cla(const uint64_t *X, size_t n_coeffs, uint64_t *out, size_t n_out)
{
    uint64_t prev_hi = 0;
    unsigned char cin = 0;  /* running prefix G (carry into limb j) */
    for (size_t j = 0; j < n_out; j++) {
        u128 s = 0;
        for (int t = 0; t < 4; t++) {
            size_t k = 4 * j + t;
            if (k >= n_coeffs) break;
            if (X[3 * k + 1] | X[3 * k + 2]) return -1;
            s += (u128)X[3 * k] << (16 * t);
        }
        uint64_t lo = (uint64_t)s, hi = (uint64_t)(s >> 64);
        u128 tt = (u128)lo + prev_hi;
        unsigned char g = (unsigned char)(tt >> 64);
        uint64_t res = (uint64_t)tt;
        unsigned char pp = res == ~0ull;
        out[j] = res + cin;
        /* compose prefix with this limb's cell */
        cin = (unsigned char)(g | (pp & cin));
        prev_hi = hi;
    }
    return (prev_hi || cin) ? 1 : 0;
}

int orc_log_len(size_t na, size_t nb)
{
    size_t need = 4 * (na + nb);   /* >= 4(na+nb)-1 coefficients */
    int l = 1;
    while (((size_t)1 << l) < need) l++;
    return l;
}

static int orc_mul_finish(uint64_t **ra, size_t na, size_t nb, size_t n, uint64_t *out);

/* Pipeline order of ntt_bigint_mul (ntt/lib/arith/ntt_mul.hip:995-1235):
 * split -> forward(a), forward(b) per prime -> pointwise -> inverse ->
 * Garner -> carry assembly.  The product has at most na+nb limbs. */
int orc_mul_ntt3(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                 uint64_t *out, int n_threads)
{
    if (na == 0 || nb == 0) { memset(out, 0, (na + nb) * sizeof(uint64_t)); return 0; }
    int log_n = orc_log_len(na, nb);
    if (log_n > 30) return -1;
    size_t n = (size_t)1 << log_n;
    uint64_t *ra[ORC_NPRIMES];
    int ok = 1;
    for (int i = 0; i < ORC_NPRIMES; i++) {
        ra[i] = (uint64_t *)malloc(n * sizeof(uint64_t));
        if (!ra[i]) ok = 0;
    }
    if (!ok) { for (int i = 0; i < ORC_NPRIMES; i++) free(ra[i]); return -1; }
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : 1;
#pragma omp parallel for num_threads(nt) schedule(static, 1) if (nt > 1)
#endif
    for (int i = 0; i < ORC_NPRIMES; i++) {
        uint64_t *fb = (uint64_t *)malloc(n * sizeof(uint64_t));
        orc_split16(a, na, ra[i], n);
        orc_split16(b, nb, fb, n);
        orc_ntt_forward(ra[i], log_n, i);
        orc_ntt_forward(fb, log_n, i);
        for (size_t k = 0; k < n; k++) ra[i][k] = orc_mulmod(ra[i][k], fb[k], PR[i]);
        orc_ntt_inverse(ra[i], log_n, i);
        free(fb);
    }
    (void)n_threads;
    return orc_mul_finish(ra, na, nb, n, out);
}

/* Garner and the carry of orc_mul_ntt3 over its three residue rows. */
static int orc_mul_finish(uint64_t **ra, size_t na, size_t nb, size_t n, uint64_t *out)
{
    size_t ncoef = 4 * (na + nb);
    if (ncoef > n) ncoef = n;
    uint64_t *X = (uint64_t *)malloc(3 * ncoef * sizeof(uint64_t));
    const GarnerK K = garner_consts();
    for (size_t k = 0; k < ncoef; k++) garner3_k(ra[0][k], ra[1][k], ra[2][k], &K, X + 3 * k);
    uint64_t spill = orc_pack16(X, ncoef, out, na + nb);
    free(X);
    for (int i = 0; i < ORC_NPRIMES; i++) free(ra[i]);
    return spill ? -1 : 0;
}

/* orc_mul_ntt3 with every transform spread over n_threads host threads (the
 * all-core CPU baseline): the primes one after the other, each prime's
 * forward table built once for both operands, Garner and the split in
 * parallel, the carry sequential.  Same products as orc_mul_ntt3. */
int orc_mul_ntt3_par(const uint64_t *a, size_t na, const uint64_t *b, size_t nb, uint64_t *out, int n_threads)
{
    if (na == 0 || nb == 0) { memset(out, 0, (na + nb) * sizeof(uint64_t)); return 0; }
    const int log_n = orc_log_len(na, nb);
    if (log_n > 30) return -1;
    const int nt = n_threads > 0 ? n_threads : 1;
    const size_t n = (size_t)1 << log_n;
    uint64_t *ra[ORC_NPRIMES] = {0};
    uint64_t *fb = (uint64_t *)malloc(n * sizeof(uint64_t));
    int ok = fb != NULL;
    for (int i = 0; i < ORC_NPRIMES; i++) {
        ra[i] = (uint64_t *)malloc(n * sizeof(uint64_t));
        if (!ra[i]) ok = 0;
    }
    for (int i = 0; ok && i < ORC_NPRIMES; i++) {
        const uint64_t p = PR[i], w = orc_root(i, log_n);
        uint64_t *x = ra[i];
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
        for (size_t k = 0; k < n; k++) {
            const size_t li = k >> 2;
            x[k] = li < na ? (a[li] >> (16 * (k & 3))) & 0xffff : 0;
            fb[k] = li < nb ? (b[li] >> (16 * (k & 3))) & 0xffff : 0;
        }
        TwTab t;
        if (tw_build(&t, w, p, log_n, nt)) { ok = 0; break; }
        ct_dit_par(x, log_n, &t, p, nt);
        ct_dit_par(fb, log_n, &t, p, nt);
        tw_free(&t);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
        for (size_t k = 0; k < n; k++) x[k] = orc_mulmod(x[k], fb[k], p);
        if (tw_build(&t, orc_invmod(w, p), p, log_n, nt)) { ok = 0; break; }
        ct_dit_par(x, log_n, &t, p, nt);
        tw_free(&t);
        const uint64_t ninv = orc_invmod(n % p, p), ninvp = orc_shoup_pre(ninv, p);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
        for (size_t k = 0; k < n; k++) x[k] = orc_shoup_mul(x[k], ninv, ninvp, p);
    }
    free(fb);
    int rc = -1;
    if (ok) {
        size_t ncoef = 4 * (na + nb);
        if (ncoef > n) ncoef = n;
        uint64_t *X = (uint64_t *)malloc(3 * ncoef * sizeof(uint64_t));
        if (X) {
            const GarnerK K = garner_consts();
#ifdef _OPENMP
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
            for (size_t k = 0; k < ncoef; k++) garner3_k(ra[0][k], ra[1][k], ra[2][k], &K, X + 3 * k);
            rc = orc_pack16(X, ncoef, out, na + nb) ? -1 : 0;
            free(X);
        }
    }
    for (int i = 0; i < ORC_NPRIMES; i++) free(ra[i]);
    return rc;
}

/* ntt/lib/arith/multiply.c:35-67 (mul_schoolbook): per row a u128 carry runs
 * across b, then propagates past the b span. */
void orc_mul_schoolbook(const uint64_t *a, size_t na, const uint64_t *b, size_t nb,
                        uint64_t *out)
{
    memset(out, 0, (na + nb) * sizeof(uint64_t));
    for (size_t i = 0; i < na; i++) {
        u128 carry = 0;
        for (size_t j = 0; j < nb; j++) {
            u128 pr = (u128)a[i] * b[j];
            u128 acc = (u128)out[i + j] + (uint64_t)pr + (uint64_t)carry;
            out[i + j] = (uint64_t)acc;
            carry = (pr >> 64) + (acc >> 64) + (carry >> 64);
        }
        size_t k = i + nb;
        while (carry) {
            u128 s = (u128)out[k] + (uint64_t)carry;
            out[k++] = (uint64_t)s;
            carry = (s >> 64) + (carry >> 64);
        }
    }
}

/* ---- binary splitting (ntt/app/compute_e/binary_split.c:74-155) ---- */

typedef struct { uint64_t *d; size_t n; } Num;

static void num_trim(Num *x) { while (x->n && x->d[x->n - 1] == 0) x->n--; }

static Num num_mul(const Num *a, const Num *b)
{
    Num r;
    r.n = a->n + b->n;
    r.d = (uint64_t *)calloc(r.n ? r.n : 1, sizeof(uint64_t));
    if (a->n && b->n) {
        if (a->n >= 64 && b->n >= 64) orc_mul_ntt3(a->d, a->n, b->d, b->n, r.d, 1);
        else orc_mul_schoolbook(a->d, a->n, b->d, b->n, r.d);
    }
    num_trim(&r);
    return r;
}

static Num num_add(const Num *a, const Num *b)
{
    Num r;
    size_t n = (a->n > b->n ? a->n : b->n) + 1;
    r.d = (uint64_t *)calloc(n, sizeof(uint64_t));
    r.n = n;
    u128 c = 0;
    for (size_t i = 0; i < n; i++) {
        c += (u128)(i < a->n ? a->d[i] : 0) + (i < b->n ? b->d[i] : 0);
        r.d[i] = (uint64_t)c;
        c >>= 64;
    }
    num_trim(&r);
    return r;
}

static void split_rec(long a, long b, Num *P, Num *B)
{
    if (b <= a) {   /* empty range: P = 0, B = 1 (binary_split.c:84-90) */
        P->d = (uint64_t *)calloc(1, 8); P->n = 0;
        B->d = (uint64_t *)calloc(1, 8); B->d[0] = 1; B->n = 1;
        return;
    }
    if (b - a == 1) {   /* leaf: P = 1, B = a (binary_split.c:91-97) */
        P->d = (uint64_t *)calloc(1, 8); P->d[0] = 1; P->n = 1;
        B->d = (uint64_t *)calloc(1, 8); B->d[0] = (uint64_t)a; B->n = 1;
        return;
    }
    long m = a + (b - a) / 2;   /* binary_split.c:99 */
    Num PL, BL, PR_, BR;
    split_rec(a, m, &PL, &BL);
    split_rec(m, b, &PR_, &BR);
    *B = num_mul(&BL, &BR);           /* B = B_L * B_R */
    Num t = num_mul(&PL, &BR);        /* P = P_L * B_R + P_R */
    *P = num_add(&t, &PR_);
    free(t.d); free(PL.d); free(BL.d); free(PR_.d); free(BR.d);
}

void orc_binary_split(long a, long b, uint64_t **P, size_t *np, uint64_t **B, size_t *nb)
{
    Num p, q;
    split_rec(a, b, &p, &q);
    *P = p.d; *np = p.n; *B = q.d; *nb = q.n;
}

/* binary_split.c:143-155 (split_terms_needed): Stirling fixed point, x1.05 + 64. */
long orc_split_terms_needed(long n_digits)
{
    double target = (double)n_digits * log(10.0);
    double N = target / log(target);
    for (int i = 0; i < 50; i++) {
        double logN = log(N);
        if (logN <= 1.0) { N = target; break; }
        N = target / (logN - 1.0);
    }
    return (long)(N * 1.05) + 64;
}

void orc_splitmix_fill(uint64_t seed, uint64_t *out, size_t n)
{
    uint64_t s = seed;
    for (size_t i = 0; i < n; i++) {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        out[i] = z ^ (z >> 31);
    }
}
