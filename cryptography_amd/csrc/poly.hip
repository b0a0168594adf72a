/*
 * poly.hip — cyclic and negacyclic products of integer polynomials on the
 * CRT-NTT passes: a second client of the transform kernels.
 *
 * f * g mod (X^n - 1) is the length-n cyclic convolution: the coefficients
 * are reduced mod each prime into the residue rows, the forward passes, the
 * fused pointwise pass and the inverse passes run as for an integer product,
 * and a full three-digit Garner rebuilds every coefficient modulo
 * P = p0 p1 p2 (> 2^179).  f * g mod (X^n + 1) twists the inputs by psi^k
 * (psi a primitive 2n-th root of unity) and untwists the result by psi^-k.
 * Coefficients are < 2^64, so |c_k| < n 2^128 <= 2^156 < P / 2: the residue
 * modulo P is exact (a negative negacyclic coefficient c is returned as
 * P + c).
 *
 * The reference runs the same algorithm across four GPUs with four primes
 * (ntt/lib/main.hip:434-683; twist / untwist kernels ntt/lib/ntt_kernel.hip:
 * 517-564, psi tables main.hip:238-260).
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include "../../include/ntt_poly.h"
#include "engine_internal.h"
#include "modarith.h"

using namespace nttc;
using namespace ntte;

namespace {

/* x * w^E mod p for the 2^30-order root w (E < 2^30), x < 2^64; result in [0, 2p). */
template <int PI>
__device__ __forceinline__ uint64_t times_root_pow(uint64_t x, const TwPair *tw, uint32_t E)
{
    const TwPair *t = tw + (size_t)PI * 2 * NTT_TW_SIZE;
    TwPair lo = t[E & (NTT_TW_SIZE - 1)], hi = t[NTT_TW_SIZE + (E >> NTT_TW_LOG)];
    return shoup_lazy<PI>(shoup_lazy<PI>(x, lo.w, lo.wp), hi.w, hi.wp);
}

template <int PI>
__device__ __forceinline__ uint64_t load_one(uint64_t x, size_t k, size_t n, int logn, int nega, const TwPair *tw)
{
    if (k >= n) return 0;
    uint64_t r = x % Prime<PI>::P;
    /* psi^k = w^(k 2^(29 - logn)) */
    return nega ? times_root_pow<PI>(r, tw, (uint32_t)(k << (29 - logn))) : r;
}

/* grid: (rows / 256, prime, operand).  Residues of f (operand 0) and g
 * (operand 1), twisted for the negacyclic product, zero past n. */
__global__ void __launch_bounds__(256) k_poly_load(uint64_t *res, size_t len, uint32_t pad_log, const uint64_t *f,
                                                   const uint64_t *g, size_t n, size_t nrow, int logn, int nega,
                                                   const TwPair *tw)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= nrow) return;
    const uint32_t pi = blockIdx.y, op = blockIdx.z;
    const uint64_t x = k < n ? (op ? g : f)[k] : 0;
    uint64_t r = pi == 0 ? load_one<0>(x, k, n, logn, nega, tw)
                         : pi == 1 ? load_one<1>(x, k, n, logn, nega, tw) : load_one<2>(x, k, n, logn, nega, tw);
    res[((size_t)op * NTT_NPRIMES + pi) * len + res_at(k, pad_log)] = r;
}

/* r: inverse-pass output, any 64-bit value. */
template <int PI>
__device__ __forceinline__ uint64_t store_one(uint64_t r, size_t k, int logn, int nega, const TwPair *tw)
{
    if (nega) r = times_root_pow<PI>(r, tw, (0u - (uint32_t)(k << (29 - logn))) & ((1u << NTT_ROOT_LOG) - 1));
    return reduce1<PI>(reduce_any<PI>(r));
}

/* out[3k .. 3k+2] = the value in [0, P) with residues r0, r1, r2: mixed radix
 * a0 + a1 p0 + a2 p0 p1 (the integer product's Garner order), then the
 * 192-bit sum. */
__global__ void __launch_bounds__(256) k_poly_store(uint64_t *out, const uint64_t *res, size_t len, uint32_t pad_log,
                                                    size_t n, int logn, int nega, const TwPair *tw, GarnerConsts gc,
                                                    uint64_t pp_lo, uint64_t pp_hi)
{
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const size_t at = res_at(k, pad_log);
    const uint64_t x0 = store_one<0>(res[at], k, logn, nega, tw);
    const uint64_t x1 = store_one<1>(res[len + at], k, logn, nega, tw);
    const uint64_t x2 = store_one<2>(res[2 * len + at], k, logn, nega, tw);
    const uint64_t p0 = Prime<0>::P, p1 = Prime<1>::P, p2 = Prime<2>::P;
    const uint64_t a0 = x0;
    const uint64_t a0m1 = a0 >= p1 ? a0 - p1 : a0;
    const uint64_t t1 = x1 >= a0m1 ? x1 - a0m1 : x1 + p1 - a0m1;
    const uint64_t a1 = reduce1<1>(shoup_lazy<1>(t1, gc.inv01, gc.inv01p));
    const uint64_t a0m2 = a0 >= p2 ? a0 - p2 : a0;
    uint64_t b2 = a0m2 + reduce1<2>(shoup_lazy<2>(a1, gc.p0m2, gc.p0m2p));
    b2 = b2 >= p2 ? b2 - p2 : b2;
    const uint64_t t2 = x2 >= b2 ? x2 - b2 : x2 + p2 - b2;
    const uint64_t a2 = reduce1<2>(shoup_lazy<2>(t2, gc.inv012, gc.inv012p));
    /* X = a0 + a1 p0 + a2 (pp_hi 2^64 + pp_lo) */
    uint64_t w0 = a1 * p0, w1 = mulhi64(a1, p0), w2 = 0;
    w0 += a0;
    w1 += w0 < a0;
    uint64_t l = a2 * pp_lo, h = mulhi64(a2, pp_lo);
    w0 += l;
    uint64_t c = w0 < l;
    uint64_t m = a2 * pp_hi, mh = mulhi64(a2, pp_hi);
    uint64_t s = w1 + h;
    uint64_t c1 = s < h;
    s += c;
    c1 += s < c;
    w1 = s + m;
    c1 += w1 < m;
    w2 = mh + c1;
    out[3 * k] = w0;
    out[3 * k + 1] = w1;
    out[3 * k + 2] = w2;
}

int poly_enqueue(uint64_t *d_out, const uint64_t *d_f, const uint64_t *d_g, size_t n, int nega)
{
    int logn = 0;
    while (((size_t)1 << logn) < n) logn++;
    if (n < 2 || ((size_t)1 << logn) != n || logn > NTT_POLY_MAX_LOG)
        return fail(NTT_ERR_ARG, "polynomial length %zu is not a power of two in [2, 2^%d]", n, NTT_POLY_MAX_LOG);
    if (!d_out || !d_f || !d_g) return fail(NTT_ERR_ARG, "null polynomial buffer");
    size_t rw, sb;
    class_needs(logn, 1, 0, 0, rw, sb);
    int rc = ws_reserve(g_ws, rw, sb);
    if (rc) return rc;
    const size_t len = row_len(logn, 1), nrow = row_residues(logn, 1);
    const uint32_t pad_log = pad_log_for(logn);
    hipLaunchKernelGGL(k_poly_load, dim3((unsigned)((nrow + 255) / 256), NTT_NPRIMES, 2), dim3(256), 0, g.stream,
                       g_ws.res, len, pad_log, d_f, d_g, n, nrow, logn, nega, g.tw);
    HIP_TRY(hipGetLastError());
    HIP_TRY(mark("k_poly_load"));
    rc = enqueue_ntt_class(g_ws, nullptr, MulDesc{}, 1, logn, 0, 0, g.stream, 1);
    if (rc) return rc;
    const unsigned __int128 pp = (unsigned __int128)prime(0) * prime(1);
    hipLaunchKernelGGL(k_poly_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g.stream, d_out, g_ws.res,
                       len, pad_log, n, logn, nega, g.tw, g.gc, (uint64_t)pp, (uint64_t)(pp >> 64));
    HIP_TRY(hipGetLastError());
    HIP_TRY(mark("k_poly_store"));
    g.dispatches++;
    return NTT_OK;
}

}  // namespace

extern "C" {

int ntt_poly_mul_device(uint64_t *d_out, const uint64_t *d_f, const uint64_t *d_g, size_t n, int negacyclic)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    int rc = poly_enqueue(d_out, d_f, d_g, n, negacyclic ? 1 : 0);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g.stream));
    return NTT_OK;
}

int ntt_poly_mul(uint64_t *out, const uint64_t *f, const uint64_t *g_, size_t n, int negacyclic)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (!out || !f || !g_) return fail(NTT_ERR_ARG, "null polynomial buffer");
    if (n < 2 || (n & (n - 1)) || n > ((size_t)1 << NTT_POLY_MAX_LOG))
        return fail(NTT_ERR_ARG, "polynomial length %zu is not a power of two in [2, 2^%d]", n, NTT_POLY_MAX_LOG);
    uint64_t *d = nullptr;
    if (hipMalloc((void **)&d, 5 * n * sizeof(uint64_t)) != hipSuccess)
        return fail(NTT_ERR_NOMEM, "polynomial buffers (%zu words)", 5 * n);
    uint64_t *df = d, *dg = d + n, *dout = d + 2 * n;
    int rc = copy_h2d(df, f, n * 8);
    if (!rc) rc = copy_h2d(dg, g_, n * 8);
    if (!rc) rc = poly_enqueue(dout, df, dg, n, negacyclic ? 1 : 0);
    if (!rc) rc = copy_d2h(out, dout, 3 * n * 8);
    (void)hipStreamSynchronize(g.stream);
    (void)hipFree(d);
    return rc;
}

}  // extern "C"
