/*
 * ntt_engine.hip — host orchestration of the gfx950 CRT-NTT multiply.
 *
 * Replaces the reference's engine (ntt/lib/arith/ntt_mul.hip:379-531 init,
 * :995-1235 ntt_bigint_mul).  Differences by design:
 *   - the transform length is sized per product (the reference fixes N at init);
 *   - twiddle and Shoup tables are built on the device, two 2^15-entry tables
 *     per prime instead of N-entry tables;
 *   - split, Garner and the carry run on the device; only the limbs cross PCIe.
 */
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "../../include/ntt_engine.h"
#include "modarith.h"
#include "ntt_kernels.h"

using namespace nttc;

#define NTT_MAX_MARKS 40

namespace {

struct Engine {
    int init = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    TwPair *tw = nullptr;
    /* workspace (grow-only) */
    uint64_t *res = nullptr;
    size_t res_words = 0;
    uint8_t *scratch = nullptr;     /* gen bytes + block bytes */
    size_t scratch_bytes = 0;
    int *flag = nullptr;
    /* host staging for the host-buffer API */
    uint64_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;
    size_t d_a_words = 0, d_b_words = 0, d_c_words = 0;
    int threshold = 64;
    unsigned long dispatches = 0;
    int timing = 0;
    /* timing marks: one event after each launch of the last multiply */
    hipEvent_t ev[NTT_MAX_MARKS] = {};
    const char *mark_name[NTT_MAX_MARKS] = {};
    int n_marks = 0, ntt_mark = 0;
    double mark_ms[NTT_MAX_MARKS] = {};
    double last_ntt_ms = 0, last_total_ms = 0;
    int timing_pending = 0;
    GarnerConsts gc;
    char err[512] = "";
};

Engine g;
std::recursive_mutex g_lock;

/* Record a timing event named for the launch that precedes it. */
hipError_t mark(const char *name)
{
    if (!g.timing || g.n_marks >= NTT_MAX_MARKS) return hipSuccess;
    g.mark_name[g.n_marks] = name;
    return hipEventRecord(g.ev[g.n_marks++], g.stream);
}

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g.err, sizeof g.err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(x)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(NTT_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

typedef unsigned __int128 u128;

uint64_t h_mulmod(uint64_t a, uint64_t b, uint64_t p) { return (uint64_t)((u128)a * b % p); }
uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1;
    a %= p;
    while (e) {
        if (e & 1) r = h_mulmod(r, a, p);
        a = h_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}
uint64_t h_inv(uint64_t a, uint64_t p) { return h_powmod(a, p - 2, p); }
uint64_t h_shoup(uint64_t w, uint64_t p) { return (uint64_t)(((u128)w << 64) / p); }

int grow_dev(uint64_t **buf, size_t *have, size_t words)
{
    if (*have >= words) return NTT_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    hipError_t e = hipMalloc((void **)buf, words * 8);
    if (e != hipSuccess) {
        *buf = nullptr;
        return fail(NTT_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", words * 8, hipGetErrorString(e));
    }
    *have = words;
    return NTT_OK;
}

int ensure_scratch(size_t bytes)
{
    if (g.scratch_bytes >= bytes) return NTT_OK;
    if (g.scratch) (void)hipFree(g.scratch);
    g.scratch = nullptr;
    g.scratch_bytes = 0;
    hipError_t e = hipMalloc((void **)&g.scratch, bytes);
    if (e != hipSuccess) return fail(NTT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    g.scratch_bytes = bytes;
    return NTT_OK;
}

int engine_init(int device)
{
    if (g.init) return NTT_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device)
        return fail(NTT_ERR_NOT_INIT, "no HIP device %d (device count %d)", device, n);
    HIP_TRY(hipSetDevice(device));
    g.device = device;
    HIP_TRY(hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc((void **)&g.tw, sizeof(TwPair) * NTT_NPRIMES * 2 * NTT_TW_SIZE));
    HIP_TRY(hipMalloc((void **)&g.flag, sizeof(int)));
    HIP_TRY(hipMemsetAsync(g.flag, 0, sizeof(int), g.stream));
    for (int i = 0; i < NTT_MAX_MARKS; i++) HIP_TRY(hipEventCreate(&g.ev[i]));
    uint64_t r2[3];
    for (int i = 0; i < NTT_NPRIMES; i++) {
        uint64_t p = prime(i);
        uint64_t r = (uint64_t)(((u128)1 << 64) % p);
        r2[i] = h_mulmod(r, r, p);
    }
    hipLaunchKernelGGL(k_twiddle_gen, dim3(NTT_TW_SIZE / 256, NTT_NPRIMES), dim3(256), 0, g.stream, g.tw,
                       r2[0], r2[1], r2[2]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g.stream));
    /* Garner constants (garner.hip:105-132 restated for three primes) */
    const uint64_t p0 = prime(0), p1 = prime(1), p2 = prime(2);
    g.gc.inv01 = h_inv(p0 % p1, p1);
    g.gc.inv01p = h_shoup(g.gc.inv01, p1);
    g.gc.p0m2 = p0 % p2;
    g.gc.p0m2p = h_shoup(g.gc.p0m2, p2);
    g.gc.inv012 = h_mulmod(h_inv(p0 % p2, p2), h_inv(p1 % p2, p2), p2);
    g.gc.inv012p = h_shoup(g.gc.inv012, p2);
    g.init = 1;
    return NTT_OK;
}

void engine_teardown()
{
    if (!g.init) return;
    (void)hipStreamSynchronize(g.stream);
    (void)hipFree(g.tw);
    (void)hipFree(g.flag);
    if (g.res) (void)hipFree(g.res);
    if (g.scratch) (void)hipFree(g.scratch);
    if (g.d_a) (void)hipFree(g.d_a);
    if (g.d_b) (void)hipFree(g.d_b);
    if (g.d_c) (void)hipFree(g.d_c);
    for (int i = 0; i < NTT_MAX_MARKS; i++) (void)hipEventDestroy(g.ev[i]);
    (void)hipStreamDestroy(g.stream);
    int thr = g.threshold;
    g.~Engine();
    new (&g) Engine();
    g.threshold = thr;
}

int log_len(size_t na, size_t nb)
{
    size_t need = 4 * (na + nb);
    int l = 1;
    while (((size_t)1 << l) < need) l++;
    return l;
}

struct PassPlan {
    int npass;
    uint32_t logS[8];
};

PassPlan plan_passes(int logL)
{
    PassPlan pp{};
    if (logL <= 12) {
        pp.npass = 1;
        pp.logS[0] = logL;
        return pp;
    }
    pp.npass = (logL + 8) / 9;
    int base = logL / pp.npass, extra = logL % pp.npass;
    for (int i = 0; i < pp.npass; i++) pp.logS[i] = base + (i < extra ? 1 : 0);
    return pp;
}

__global__ void k_fill_zero(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

/* Enqueue c = a*b (or a^2) on the engine stream.  Device pointers. */
int enqueue_mul(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, int square)
{
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((na && !d_a) || (nb && !d_b) || ((na + nb) && !d_c)) return fail(NTT_ERR_ARG, "null operand pointer");
    const size_t nout = na + nb;
    g.dispatches++;
    if (na == 0 || nb == 0) {
        if (nout)
            hipLaunchKernelGGL(k_fill_zero, dim3((nout + 255) / 256), dim3(256), 0, g.stream, d_c, nout);
        HIP_TRY(hipGetLastError());
        return NTT_OK;
    }
    if ((int64_t)na <= g.threshold && (int64_t)nb <= g.threshold && nout <= 2048) {
        hipLaunchKernelGGL(k_schoolbook, dim3(1), dim3(256), 3 * nout * 8, g.stream, d_a, na, d_b, nb, d_c);
        HIP_TRY(hipGetLastError());
        return NTT_OK;
    }
    const int logL = log_len(na, nb);
    if (logL > NTT_MAX_LOG_LEN)
        return fail(NTT_ERR_TOO_LARGE, "product of %zu x %zu limbs needs 2^%d coefficients (max 2^%d)", na, nb,
                    logL, NTT_MAX_LOG_LEN);
    const size_t L = (size_t)1 << logL;
    const int nops = square ? 1 : 2;
    int rc = grow_dev(&g.res, &g.res_words, (size_t)nops * NTT_NPRIMES * L);
    if (rc) return rc;
    const size_t nthr = (nout + CARRY_E - 1) / CARRY_E;
    const size_t nblk = (nthr + 255) / 256;
    rc = ensure_scratch(nthr + nblk + 64);
    if (rc) return rc;

    const PassPlan pp = plan_passes(logL);
    const uint32_t tile_log = logL < 12 ? logL : 12;
    const size_t ntiles = L >> tile_log;
    const size_t lds = (size_t)8 << tile_log;

    g.n_marks = 0;
    HIP_TRY(mark("start"));
    PassArgs pa{};
    pa.res = g.res;
    pa.tw = g.tw;
    pa.src16[0] = (const uint16_t *)d_a;
    pa.src16[1] = (const uint16_t *)(square ? d_a : d_b);
    pa.n16[0] = 4 * na;
    pa.n16[1] = 4 * (square ? na : nb);
    pa.len = L;
    uint32_t done = 0;
    for (int i = 0; i < pp.npass; i++) {
        pa.logS = pp.logS[i];
        pa.logT = logL - done - pa.logS;
        pa.logC = pa.logT < tile_log - pa.logS ? pa.logT : tile_log - pa.logS;
        pa.logG = tile_log - pa.logS - pa.logC;
        pa.first = i == 0;
        hipLaunchKernelGGL(k_fwd_pass, dim3(ntiles, NTT_NPRIMES, nops), dim3(256), lds, g.stream, pa);
        HIP_TRY(mark("k_fwd_pass"));
        done += pa.logS;
    }
    PointwiseArgs pw{};
    pw.res = g.res;
    pw.len = L;
    pw.square = square;
    for (int i = 0; i < NTT_NPRIMES; i++) {
        uint64_t p = prime(i);
        uint64_t r = (uint64_t)(((u128)1 << 64) % p);
        pw.scale[i] = h_mulmod(r, h_inv(L % p, p), p);
        pw.scale_p[i] = h_shoup(pw.scale[i], p);
    }
    size_t pwb = L / 256 < 8192 ? (L + 255) / 256 : 8192;
    hipLaunchKernelGGL(k_pointwise, dim3(pwb, NTT_NPRIMES), dim3(256), 0, g.stream, pw);
    HIP_TRY(mark("k_pointwise"));
    for (int i = pp.npass - 1; i >= 0; i--) {
        done -= pp.logS[i];
        pa.logS = pp.logS[i];
        pa.logT = logL - done - pa.logS;
        pa.logC = pa.logT < tile_log - pa.logS ? pa.logT : tile_log - pa.logS;
        pa.logG = tile_log - pa.logS - pa.logC;
        pa.first = 0;
        hipLaunchKernelGGL(k_inv_pass, dim3(ntiles, NTT_NPRIMES, 1), dim3(256), lds, g.stream, pa);
        HIP_TRY(mark("k_inv_pass"));
    }
    g.ntt_mark = g.n_marks - 1;
    GarnerArgs ga{};
    ga.res = g.res;
    ga.len = L;
    ga.n_coef = 4 * nout < L ? 4 * nout : L;
    ga.n_out = nout;
    ga.out = d_c;
    ga.gen = g.scratch;
    ga.blk = g.scratch + nthr;
    ga.flag = g.flag;
    ga.gc = g.gc;
    hipLaunchKernelGGL(k_garner_pack, dim3(nblk), dim3(256), 0, g.stream, ga);
    HIP_TRY(mark("k_garner_pack"));
    hipLaunchKernelGGL(k_carry_scan, dim3(1), dim3(1024), 0, g.stream, ga.blk, nblk);
    HIP_TRY(mark("k_carry_scan"));
    hipLaunchKernelGGL(k_carry_apply, dim3(nblk), dim3(256), 0, g.stream, ga);
    HIP_TRY(mark("k_carry_apply"));
    g.timing_pending = g.timing;
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

int sync_check()
{
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (g.timing_pending && g.n_marks > 1) {
        g.mark_ms[0] = 0;
        for (int i = 1; i < g.n_marks; i++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, g.ev[i - 1], g.ev[i]));
            g.mark_ms[i] = ms;
        }
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, g.ev[0], g.ev[g.ntt_mark]));
        HIP_TRY(hipEventElapsedTime(&b, g.ev[0], g.ev[g.n_marks - 1]));
        g.last_ntt_ms = a;
        g.last_total_ms = b;
    }
    g.timing_pending = 0;
    int flag = 0;
    HIP_TRY(hipMemcpy(&flag, g.flag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        HIP_TRY(hipMemset(g.flag, 0, sizeof(int)));
        return fail(NTT_ERR_INTEGRITY, "CRT integrity check failed: a product coefficient exceeded its bound");
    }
    return NTT_OK;
}

int host_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square)
{
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((na && !a) || (nb && !b) || ((na + nb) && !c)) return fail(NTT_ERR_ARG, "null operand pointer");
    const size_t nout = na + nb;
    if (nout == 0) return NTT_OK;
    int rc = grow_dev(&g.d_a, &g.d_a_words, na ? na : 1);
    if (!rc && !square) rc = grow_dev(&g.d_b, &g.d_b_words, nb ? nb : 1);
    if (!rc) rc = grow_dev(&g.d_c, &g.d_c_words, nout);
    if (rc) return rc;
    if (na) HIP_TRY(hipMemcpyAsync(g.d_a, a, na * 8, hipMemcpyHostToDevice, g.stream));
    if (!square && nb) HIP_TRY(hipMemcpyAsync(g.d_b, b, nb * 8, hipMemcpyHostToDevice, g.stream));
    rc = enqueue_mul(g.d_c, g.d_a, na, square ? g.d_a : g.d_b, nb, square);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c, g.d_c, nout * 8, hipMemcpyDeviceToHost, g.stream));
    return sync_check();
}

}  // namespace

extern "C" {

const char *ntt_engine_last_error(void) { return g.err; }

int ntt_engine_init(int device)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return engine_init(device);
}

void ntt_engine_teardown(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    engine_teardown();
}

int ntt_engine_is_init(void) { return g.init; }

int ntt_engine_log_len(size_t na, size_t nb) { return log_len(na, nb); }

int ntt_mul_limbs(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return host_mul(c, a, na, b, nb, 0);
}

int ntt_sqr_limbs(uint64_t *c, const uint64_t *a, size_t na)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return host_mul(c, a, na, a, na, 1);
}

int ntt_mul_device_async(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return enqueue_mul(d_c, d_a, na, d_b, nb, 0);
}

int ntt_sqr_device_async(uint64_t *d_c, const uint64_t *d_a, size_t na)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return enqueue_mul(d_c, d_a, na, d_a, na, 1);
}

int ntt_engine_sync(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    return sync_check();
}

int ntt_mul_device(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int rc = enqueue_mul(d_c, d_a, na, d_b, nb, 0);
    return rc ? rc : sync_check();
}

int ntt_sqr_device(uint64_t *d_c, const uint64_t *d_a, size_t na)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int rc = enqueue_mul(d_c, d_a, na, d_a, na, 1);
    return rc ? rc : sync_check();
}

int ntt_engine_get_threshold(void) { return g.threshold; }
void ntt_engine_set_threshold(int limbs) { g.threshold = limbs; }
unsigned long ntt_engine_dispatch_count(void) { return g.dispatches; }
void ntt_engine_set_timing(int on) { g.timing = on; }
double ntt_engine_last_ntt_ms(void) { return g.last_ntt_ms; }
double ntt_engine_last_total_ms(void) { return g.last_total_ms; }

int ntt_engine_kernel_times(char *names, double *ms, int max)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int n = 0;
    for (int i = 1; i < g.n_marks && n < max; i++, n++) {
        snprintf(names + 32 * n, 32, "%s", g.mark_name[i]);
        ms[n] = g.mark_ms[i];
    }
    return n;
}

void *ntt_device_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    return p;
}
void ntt_device_free(void *p)
{
    if (p) (void)hipFree(p);
}
int ntt_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return NTT_OK;
}
int ntt_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return NTT_OK;
}

/* bigint_alloc_managed storage (cf. ntt/lib/arith/bigint_hip_alloc.hip). */
uint64_t *ntt_managed_alloc_limbs(int cap)
{
    void *p = nullptr;
    if (hipMallocManaged(&p, (size_t)cap * 8) != hipSuccess) return nullptr;
    memset(p, 0, (size_t)cap * 8);
    return (uint64_t *)p;
}
void ntt_managed_free_limbs(uint64_t *p)
{
    if (p) (void)hipFree(p);
}

/* Signal-handler hook (ntt_mul.hip:1247-1256): drain the device. */
void ntt_gpu_emergency_reset(void)
{
    if (g.init) (void)hipDeviceSynchronize();
}

int ntt_engine_read_twiddles(uint64_t *out, size_t n_words)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    size_t need = (size_t)NTT_NPRIMES * 2 * NTT_TW_SIZE * 2;
    if (n_words < need) return fail(NTT_ERR_ARG, "twiddle buffer too small (%zu < %zu)", n_words, need);
    HIP_TRY(hipMemcpy(out, g.tw, need * 8, hipMemcpyDeviceToHost));
    return NTT_OK;
}

}  // extern "C"
