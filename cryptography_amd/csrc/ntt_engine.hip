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
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/ntt_engine.h"
#include "modarith.h"
#include "ntt_kernels.h"
#include "engine_internal.h"

using namespace nttc;

namespace ntte {

Engine g;
std::recursive_mutex g_lock;
unsigned long g_generation = 0;

Workspace g_ws;

/* Scratch of the no-allocation multiply (ntt_mul_device_noalloc): the residue
 * rows of one product whose transform has at most 2^NTT_NOALLOC_MAX_LOG
 * coefficients (2 operands x 3 primes x 2^19 words, plus room for the row
 * pads a tuned pass split may add: 24.8 MB), then its carry cells
 * (generate masks and block cells of up to 2^17 limbs: 17 KB of 64 KiB).
 * One copy per device, held from module load to process end.  Host code uses
 * the address hipGetSymbolAddress returns (g.noalloc_scratch), never the
 * array's name. */
#define NOALLOC_ROW_WORDS (((size_t)1 << NTT_NOALLOC_MAX_LOG) + ((size_t)1 << (NTT_NOALLOC_MAX_LOG - 12)) * NTT_PAD_WORDS)
#define NOALLOC_RES_WORDS ((size_t)2 * NTT_NPRIMES * NOALLOC_ROW_WORDS)
#define NOALLOC_CARRY_BYTES ((size_t)64 << 10)
static __device__ __align__(256) uint64_t scratch[NOALLOC_RES_WORDS + NOALLOC_CARRY_BYTES / 8];

/* Record a timing event named for the launch that precedes it. */
hipError_t mark(const char *name)
{
    if (!g.timing || g.n_marks >= NTT_MAX_MARKS) return hipSuccess;
    g.mark_name[g.n_marks] = name;
    return hipEventRecord(g.ev[g.n_marks++], g.stream);
}

/* The message of the calling thread's last error: callers on several threads
 * (serving submitters, the scheduler thread, bigint_mul from OpenMP regions)
 * each read their own. */
static thread_local char t_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_err, sizeof t_err, fmt, ap);
    va_end(ap);
    return code;
}

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
    /* queued work may still read the old buffer */
    if (*buf && g.stream) (void)hipStreamSynchronize(g.stream);
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
    /* this device's copy of the no-allocation scratch array */
    HIP_TRY(hipGetSymbolAddress((void **)&g.noalloc_scratch, HIP_SYMBOL(scratch)));
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
    HIP_TRY(hipMalloc((void **)&g.small_tw, sizeof(TwPair) * NTT_NPRIMES * (FT_TILE / 2)));
    HIP_TRY(launch_small_twiddles(g.tw, g.small_tw, g.stream));
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
    g_generation++;
    /* a tuning file named by the environment (as ntt_engine_autotune_save writes it) */
    if (const char *tf = getenv("NTT_AMD_TUNING")) {
        if (*tf) {
            const int rc = ntt_engine_autotune_load(tf);
            if (rc) return rc;
        }
    }
    return NTT_OK;
}

void engine_teardown()
{
    if (!g.init) return;
    (void)hipStreamSynchronize(g.stream);
    (void)hipFree(g.tw);
    (void)hipFree(g.small_tw);
    (void)hipFree(g.flag);
    for (auto &row : g.xtab)
        for (uint64_t *t : row)
            if (t) (void)hipFree(t);
    for (auto &row : g.btab)
        for (TwPair *t : row)
            if (t) (void)hipFree(t);
    g_ws.release();
    if (g.d_desc) (void)hipFree(g.d_desc);
    if (g.d_a) (void)hipFree(g.d_a);
    if (g.d_b) (void)hipFree(g.d_b);
    if (g.d_c) (void)hipFree(g.d_c);
    for (int i = 0; i < 2; i++) {
        if (g.pin[i]) (void)hipHostFree(g.pin[i]);
        if (g.pin_ev[i]) (void)hipEventDestroy(g.pin_ev[i]);
    }
    for (int i = 0; i < NTT_MAX_MARKS; i++) (void)hipEventDestroy(g.ev[i]);
    (void)hipStreamDestroy(g.stream);
    int thr = g.threshold, tb = g.tile_batch, xt = g.use_xtab;
    uint8_t ml[NTT_MAX_LOG_LEN + 1], sn[NTT_MAX_LOG_LEN + 1], sp[NTT_MAX_LOG_LEN + 1][8];
    memcpy(ml, g.max_logs, sizeof ml);
    memcpy(sn, g.split_n, sizeof sn);
    memcpy(sp, g.split, sizeof sp);
    g.~Engine();
    new (&g) Engine();
    g.threshold = thr;
    g.tile_batch = tb;
    g.use_xtab = xt;
    memcpy(g.max_logs, ml, sizeof ml);
    memcpy(g.split_n, sn, sizeof sn);
    memcpy(g.split, sp, sizeof sp);
}

/* Copies ordered after everything queued on the engine stream (the stream is
 * non-blocking, so a plain hipMemcpy would not wait for its kernels). */
int copy_h2d(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return NTT_OK;
}
int copy_d2h(void *dst, const void *src, size_t bytes)
{
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g.stream));
    HIP_TRY(hipStreamSynchronize(g.stream));
    return NTT_OK;
}

int log_len(size_t na, size_t nb)
{
    size_t need = 4 * (na + nb);
    int l = 1;
    while (((size_t)1 << l) < need) l++;
    return l;
}

/* Pass sizes of a length-2^logL transform: as few passes as possible with at
 * most 2^cap points each (cap from the autotuner, else FT_MAX_LOGS),
 * balanced, the largest last (the last pass is contiguous, the earlier ones
 * read columns). */
PassPlan plan_passes(int logL)
{
    PassPlan pp{};
    if (logL >= 1 && logL <= NTT_MAX_LOG_LEN && g.split_n[logL]) {
        pp.npass = g.split_n[logL];
        for (int i = 0; i < pp.npass; i++) pp.logS[i] = g.split[logL][i];
        return pp;
    }
    int cap = g.max_logs[logL] ? g.max_logs[logL] : FT_MAX_LOGS;
    pp.npass = (logL + cap - 1) / cap;
    if (pp.npass < 1) pp.npass = 1;
    int base = logL / pp.npass, extra = logL % pp.npass;
    for (int i = 0; i < pp.npass; i++) pp.logS[i] = base + (i >= pp.npass - extra ? 1 : 0);
    return pp;
}

uint64_t plan_signature(int logL)
{
    const PassPlan pp = plan_passes(logL);
    uint64_t sig = (uint64_t)pp.npass;
    for (int i = 0; i < pp.npass; i++) sig |= (uint64_t)pp.logS[i] << (4 + 4 * i);
    return sig;
}

const uint64_t *xtab_for(int logS, int logT, hipStream_t st)
{
    const int logN = logS + logT;
    if (!g.use_xtab || logS < 3 || logS > FT_MAX_LOGS || logT < 0 || logN < 8 || logN > NTT_XTAB_MAX_LOG)
        return nullptr;
    uint64_t *&t = g.xtab[logS][logN];
    if (t) return t;
    if (g.no_alloc) return nullptr;      /* the passes use the chains instead */
    /* made on the engine stream, which must not be recording a graph */
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone, cg = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    if (hipStreamIsCapturing(g.stream, &cg) != hipSuccess || cg != hipStreamCaptureStatusNone) return nullptr;
    uint64_t *p = nullptr;
    if (hipMalloc((void **)&p, (size_t)NTT_XTAB_DIRS * NTT_NPRIMES * ((size_t)8 << logN)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;      /* the passes fall back to the chains */
    }
    XtabScale sc{};
    for (int i = 0; i < NTT_NPRIMES; i++) {   /* 2^64 / N mod p */
        const uint64_t q = prime(i), r = (uint64_t)(((u128)1 << 64) % q);
        sc.s[i] = h_mulmod(r, h_inv(((uint64_t)1 << logN) % q, q), q);
        sc.sp[i] = h_shoup(sc.s[i], q);
    }
    /* complete before any stream's pass reads it */
    if (launch_xtab_gen(g.tw, p, logS, logT, sc, g.stream) != hipSuccess ||
        hipStreamSynchronize(g.stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(p);
        return nullptr;
    }
    t = p;
    return t;
}

/* Chain ratio tables (launch_btab_gen): the first forward pass of 8-point
 * last-round groups (kind 8) and the inverse passes that apply their cross
 * twiddles at the load without a table of them (kind 9) step their chains by
 * Shoup products with B from here instead of Montgomery products with B
 * looked up per thread.  2 x 3 x T entries of 16 bytes (24 MB for the first
 * pass at L = 2^27); nullptr where the passes compute B themselves. */
const TwPair *btab_for(int logS, int logT, hipStream_t st)
{
    const int logN = logS + logT;
    if (!g.use_xtab || logS < 3 || logS > FT_MAX_LOGS || logT < 0 || logN > NTT_MAX_LOG_LEN) return nullptr;
    TwPair *&t = g.btab[logS][logN];
    if (t) return t;
    if (g.no_alloc) return nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone, cg = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return nullptr;
    if (hipStreamIsCapturing(g.stream, &cg) != hipSuccess || cg != hipStreamCaptureStatusNone) return nullptr;
    TwPair *p = nullptr;
    if (hipMalloc((void **)&p, (size_t)NTT_BTAB_DIRS * NTT_NPRIMES * sizeof(TwPair) << logT) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (launch_btab_gen(g.tw, p, logS, logT, g.stream) != hipSuccess || hipStreamSynchronize(g.stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(p);
        return nullptr;
    }
    t = p;
    return t;
}

void xtab_prepare(int logL)
{
    if (logL <= 0) return;
    const PassPlan pp = plan_passes(logL);
    int done = 0;
    for (int i = 0; i < pp.npass - 1; i++) {
        (void)xtab_for((int)pp.logS[i], logL - done - (int)pp.logS[i], g.stream);
        (void)btab_for((int)pp.logS[i], logL - done - (int)pp.logS[i], g.stream);
        done += (int)pp.logS[i];
    }
}

/* Residues of a class row: nprod * L rounded up to whole tiles. */
size_t row_residues(int logL, uint32_t nprod)
{
    size_t len = ((size_t)nprod) << logL;
    return (len + FT_TILE - 1) / FT_TILE * FT_TILE;
}

/* Row padding of a length-2^logL class: a pad after every row of the first
 * pass (stride 2^(logL - S1) residues) when that stride is 4096 or more. */
uint32_t pad_log_for(int logL)
{
    const PassPlan pp = plan_passes(logL);
    if (pp.npass < 2) return NTT_PAD_NONE;
    int logT1 = logL - pp.logS[0];
    return logT1 >= 12 ? (uint32_t)logT1 : NTT_PAD_NONE;
}

/* Row stride in words: the residues plus their pads. */
size_t row_len(int logL, uint32_t nprod)
{
    size_t n = row_residues(logL, nprod);
    return res_at(n, pad_log_for(logL));
}

__global__ void k_fill_zero(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}


int ws_reserve(Workspace &ws, size_t res_words, size_t scratch_bytes)
{
    int rc = grow_dev(&ws.res, &ws.res_words, res_words);
    if (rc) return rc;
    if (ws.scratch_bytes < scratch_bytes) {
        if (ws.scratch && g.stream) (void)hipStreamSynchronize(g.stream);
        if (ws.scratch) (void)hipFree(ws.scratch);
        ws.scratch = nullptr;
        ws.scratch_bytes = 0;
        hipError_t e = hipMalloc((void **)&ws.scratch, scratch_bytes);
        if (e != hipSuccess)
            return fail(NTT_ERR_NOMEM, "hipMalloc(%zu) failed: %s", scratch_bytes, hipGetErrorString(e));
        ws.scratch_bytes = scratch_bytes;
    }
    return NTT_OK;
}

/* Carry geometry of a class whose largest product has max_nout limbs. */
/* Carry blocks per product for a class whose largest product spans max_nout
 * limbs, and the scratch bytes one product needs (generate masks + block cells). */
uint32_t carry_blocks(uint32_t max_nout) { return (max_nout + CB_LIMBS - 1) / CB_LIMBS; }
size_t carry_scratch_per_prod(uint32_t blk) { return (size_t)blk * 4 * CW_ROUNDS * 8 + blk; }

/* Workspace a class needs: residues and carry scratch. */
void class_needs(int logL, uint32_t nprod, uint32_t max_nout, int square, size_t &res_words, size_t &scratch)
{
    uint32_t blk = carry_blocks(max_nout);
    res_words = (size_t)(square ? 1 : 2) * NTT_NPRIMES * row_len(logL, nprod);
    scratch = (size_t)nprod * carry_scratch_per_prod(blk) + 64;
}

/* Enqueue one size class: nprod products, transform length 2^logL, all on
 * `st`.  d_desc == nullptr means the single product `one`. */
bool spectrum_cacheable(int logL) { return logL >= 12 && plan_passes(logL).npass >= 2; }
size_t spectrum_rows_len(int logL) { return row_len(logL, 1); }

int enqueue_ntt_class(Workspace &ws, const MulDesc *d_desc, const MulDesc &one, uint32_t nprod, int logL,
                      uint32_t max_nout, int square, hipStream_t st, int residues_in, uint32_t b_period,
                      uint32_t max_nab, const uint64_t *b_res, int fwd_only)
{
    const size_t L = (size_t)1 << logL;
    const size_t len = row_len(logL, nprod);
    const int nops = square ? 1 : 2;
    const uint32_t pad_log = pad_log_for(logL);
    if ((b_res || fwd_only) && (!spectrum_cacheable(logL) || residues_in || (b_res && (square || fwd_only)) ||
                                (fwd_only && (!square || nprod != 1))))
        return fail(NTT_ERR_ARG, "resident spectrum not usable for a class of length 2^%d (internal)", logL);
    size_t need_res, need_scr;
    class_needs(logL, nprod, max_nout, square || b_res, need_res, need_scr);
    if (ws.res_words < need_res || (!fwd_only && ws.scratch_bytes < need_scr))
        return fail(NTT_ERR_ARG, "workspace too small for the class (internal)");
    const PassPlan pp = plan_passes(logL);
    const size_t ntiles = row_residues(logL, nprod) / FT_TILE;

    PassArgs pa{};
    pa.res = ws.res;
    pa.tw = g.tw;
    pa.small_tw = g.small_tw;
    pa.desc = d_desc;
    pa.one = one;
    pa.len = len;
    pa.pad_log = pad_log;
    pa.nprod = nprod;
    pa.logL = logL;
    pa.square = square;
    pa.ntiles = (uint32_t)ntiles;
    /* Several tiles per block once the grid has blocks to spare (about 4096
     * blocks per prime row fill the chip several times over).  The forward
     * table pass (kind 6), the wave-local fused pass (kind 2 at S = 2^6, 2^9)
     * and the load-time inverse passes (kinds 5, 7, 9) start each next tile's
     * transfer during the current tile's last round, and take tiles nblk
     * apart, so neighbouring column tiles still run side by side: measured at
     * L = 2^27 (ms, 1 / 2 / 4 / 8 tiles per block) kind 6 with consecutive
     * tiles 3.75 / 3.62 (two), strided 3.24-3.25 at two and 3.20-3.22 at
     * eight, kind 2 3.11 / 3.08 / 3.04 / 3.02-3.05, kind 7 1.75 / 1.71 /
     * 1.71 / 1.67, kind 9 1.92 / 1.87 / 1.87 / 1.88.  The first pass is
     * slower with more than one (3.18 against 3.21-3.41 ms). */
    auto tiles_per_block = [&](size_t nt, int kind) {
        const size_t cap = kind == 2   ? std::max(g.tile_batch, 8)
                           : kind == 6 ? std::max(g.tile_batch, 8)
                           : kind == 7 ? std::max(g.tile_batch, 8)
                           : (kind == 5 || kind == 9) ? std::max(g.tile_batch, 4)
                                                      : g.tile_batch;
        return (uint32_t)std::min<size_t>(std::max<size_t>(nt * NTT_NPRIMES / 4096, 1), cap);
    };
    pa.tpb = tiles_per_block(ntiles, 0);
    for (int i = 0; i < NTT_NPRIMES; i++) {
        uint64_t p = prime(i);
        uint64_t r = (uint64_t)(((u128)1 << 64) % p);
        pa.scale[i] = h_mulmod(r, h_inv(L % p, p), p);
        pa.scale_p[i] = h_shoup(pa.scale[i], p);
        pa.scale_m[i] = h_mulmod(pa.scale[i], r, p);
    }
    /* geometry of pass i: S = 2^logS columns of stride T inside blocks of S*T */
    auto geom = [&](int i, uint32_t done) {
        uint32_t logS = pp.logS[i];
        pa.logT = logL - done - logS;
        uint32_t room = 12 - logS;                   /* log2(FT_TILE / S) */
        pa.logC = pa.logT < room ? pa.logT : room;
        pa.logG = room - pa.logC;
    };
    /* residues_in: the rows already hold the residues (polynomial products),
     * so no pass reads the operands' limbs */
    /* b repeating with period P < nprod (products i and i + P share b) is
     * transformed for the first P products only; the fused pass reads tile
     * t mod (P L / 4096) of those rows.  Needs whole tiles per product
     * (L >= 4096) and separate passes. */
    const bool share = !b_res && b_period && b_period < nprod && nprod % b_period == 0 && !square &&
                       !residues_in && pp.npass >= 2 && logL >= 12;
    const int fwd_ops = (share || b_res) ? 1 : nops;
    uint32_t done = 0;
    for (int i = 0; i < pp.npass - 1; i++) {
        geom(i, done);
        pa.first = i == 0 && !residues_in;
        /* operands of at most L/2 coefficients leave the first pass's
         * columns zero from row S/2 on */
        pa.half_zero = pa.first && max_nab && (uint64_t)8 * max_nab <= L;
        int kind = pa.first ? 3 : 0;
        pa.xtab = nullptr;
        pa.btab = nullptr;
        if (kind == 0 && (pa.xtab = xtab_for((int)pp.logS[i], (int)pa.logT, st))) kind = 6;
        if (kind == 3 && (pp.logS[i] == 6 || pp.logS[i] == 9) &&
            (pa.btab = btab_for((int)pp.logS[i], (int)pa.logT, st)))
            kind = 8;      /* dir 0 of the table */
        pa.tpb = tiles_per_block(ntiles, kind);
        HIP_TRY(launch_fast_pass(pp.logS[i], kind, pa, ntiles, fwd_ops, st));
        HIP_TRY(mark("k_fast_pass<fwd>"));
        if (share) {   /* b of the first b_period products */
            PassArgs pb = pa;
            pb.op0 = 1;
            pb.nprod = b_period;
            pb.ntiles = (uint32_t)(b_period * (L / FT_TILE));
            pb.tpb = tiles_per_block(pb.ntiles, kind);
            HIP_TRY(launch_fast_pass(pp.logS[i], kind, pb, pb.ntiles, 1, st));
            HIP_TRY(mark("k_fast_pass<fwd b>"));
        }
        done += pp.logS[i];
    }
    if (fwd_only) return NTT_OK;         /* the rows hold a's partial spectrum */
    /* the last inverse pass (pass 0's) applies the 1/L scale when it loads
     * its tile with the cross twiddles (kinds 5 and 9: its chain's first
     * twiddle carries the scale; kind 7: its table does); otherwise the
     * middle pass does */
    geom(0, 0);
    const bool scale_late = pp.npass >= 2 && pp.logS[0] >= 3 && pa.logC + pp.logS[0] == 12;
    geom(pp.npass - 1, done);
    pa.xtab = nullptr;
    pa.btab = nullptr;
    pa.half_zero = 0;
    pa.tpb = tiles_per_block(ntiles, 0);
    pa.first = pp.npass == 1 && !residues_in;
    pa.b_tiles = share ? (uint32_t)(b_period * (L / FT_TILE)) : b_res ? (uint32_t)(L / FT_TILE) : 0;
    pa.bres = b_res;
    pa.blen = b_res ? spectrum_rows_len(logL) : 0;
    pa.scale_late = scale_late;
    /* the wave-local fused pass (S = 2^6, 2^9) takes several tiles per block */
    pa.tpb = (!pa.first && (pp.logS[pp.npass - 1] == 6 || pp.logS[pp.npass - 1] == 9)) ? tiles_per_block(ntiles, 2) : 1;
    HIP_TRY(launch_fast_pass(pp.logS[pp.npass - 1], pa.first ? 4 : 2, pa, ntiles, nops, st));
    pa.b_tiles = 0;
    pa.bres = nullptr;
    pa.scale_late = 0;
    HIP_TRY(mark("k_fast_pass<mid>"));
    for (int i = pp.npass - 2; i >= 0; i--) {
        done -= pp.logS[i];
        geom(i, done);
        pa.first = 0;
        int kind = (pp.logS[i] >= 3 && pa.logC + pp.logS[i] == 12) ? 5 : 1;
        pa.xtab = nullptr;
        pa.iscale = i == 0 && scale_late;
        pa.btab = nullptr;
        if (kind == 5) {
            const uint64_t *xt = xtab_for((int)pp.logS[i], (int)pa.logT, st);
            const TwPair *bt = xt ? nullptr : btab_for((int)pp.logS[i], (int)pa.logT, st);
            if (xt) {   /* inverse tables follow the forward one; pass 0's block is
                           the whole transform, so its scaled table is the 1/L one */
                kind = 7;
                pa.xtab = xt + ((size_t)(pa.iscale ? 2 : 1) * NTT_NPRIMES << (pp.logS[i] + pa.logT));
            } else if (bt) {   /* dir 1: B = w_N^-r */
                kind = 9;
                pa.btab = bt + ((size_t)NTT_NPRIMES << pa.logT);
            }
        }
        pa.tpb = tiles_per_block(ntiles, kind);
        HIP_TRY(launch_fast_pass(pp.logS[i], kind, pa, ntiles, nops, st));
        pa.iscale = 0;
        HIP_TRY(mark("k_fast_pass<inv>"));
    }
    g.ntt_mark = g.n_marks - 1;
    if (residues_in) return NTT_OK;      /* the caller reconstructs from the residues */
    GarnerArgs ga{};
    ga.res = ws.res;
    ga.len = len;
    ga.pad_log = pad_log;
    ga.logL = logL;
    ga.desc = d_desc;
    ga.one = one;
    ga.blk_per_prod = carry_blocks(max_nout);
    ga.gen = ws.scratch;
    ga.blk = ws.scratch + (size_t)nprod * ga.blk_per_prod * 4 * CW_ROUNDS * 8;
    ga.flag = ws.flag ? ws.flag : g.flag;
    hipLaunchKernelGGL(k_garner_pack, dim3(ga.blk_per_prod, nprod), dim3(256), 0, st, ga);
    HIP_TRY(mark("k_garner_pack"));
    hipLaunchKernelGGL(k_carry_scan, dim3(nprod), dim3(1024), 0, st, ga.blk, ga.blk_per_prod, (uint8_t)0);
    HIP_TRY(mark("k_carry_scan"));
    hipLaunchKernelGGL(k_carry_apply, dim3(ga.blk_per_prod, nprod), dim3(256), 0, st, ga);
    HIP_TRY(mark("k_carry_apply"));
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

int enqueue_schoolbook(const MulDesc *d_desc, const MulDesc &one, uint32_t nprod, hipStream_t st, int *flag)
{
    hipLaunchKernelGGL(k_schoolbook, dim3((nprod + 3) / 4), dim3(256), 0, st, d_desc, one, nprod,
                       flag ? flag : g.flag);
    HIP_TRY(mark("k_schoolbook"));
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

bool use_schoolbook(size_t na, size_t nb, size_t n_out)
{
    (void)n_out;   /* the kernel's LDS holds na + nb column sums */
    return (int64_t)na <= g.threshold && (int64_t)nb <= g.threshold && na + nb <= SB_MAX_OUT;
}

int check_desc(const MulDesc &d)
{
    if ((d.na && !d.a) || (d.nb && !d.b) || (d.n_out && !d.c) || (d.nadd && !d.add))
        return fail(NTT_ERR_ARG, "null pointer in a product descriptor");
    /* n_out may be below na + nb when the caller knows the value is shorter
     * (operands with leading zero limbs); anything that does not fit is an
     * integrity error, never a silent truncation */
    if (d.n_out < d.nadd || (d.n_out == 0 && d.na && d.nb))
        return fail(NTT_ERR_ARG, "n_out=%u too small for the %u-limb addend", d.n_out, d.nadd);
    if (log_len(d.na, d.nb) > NTT_MAX_LOG_LEN)
        return fail(NTT_ERR_TOO_LARGE, "product of %u x %u limbs needs 2^%d coefficients (max 2^%d)", d.na, d.nb,
                    log_len(d.na, d.nb), NTT_MAX_LOG_LEN);
    return NTT_OK;
}

/* Enqueue c = a*b (or a^2) on the engine stream.  Device pointers. */
int enqueue_mul(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, int square,
                int force_ntt)
{
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((na && !d_a) || (nb && !d_b) || ((na + nb) && !d_c)) return fail(NTT_ERR_ARG, "null operand pointer");
    if (na > 0xffffffffu || nb > 0xffffffffu) return fail(NTT_ERR_TOO_LARGE, "operand too large");
    const size_t nout = na + nb;
    if (log_len(na, nb) > NTT_MAX_LOG_LEN)
        return fail(NTT_ERR_TOO_LARGE, "product of %zu x %zu limbs needs 2^%d coefficients (max 2^%d)", na, nb,
                    log_len(na, nb), NTT_MAX_LOG_LEN);
    g.dispatches++;
    g.n_marks = 0;
    HIP_TRY(mark("start"));
    if (na == 0 || nb == 0) {
        if (nout)
            hipLaunchKernelGGL(k_fill_zero, dim3((nout + 255) / 256), dim3(256), 0, g.stream, d_c, nout);
        HIP_TRY(hipGetLastError());
        return NTT_OK;
    }
    MulDesc one{};
    one.a = d_a;
    one.b = d_b;
    one.c = d_c;
    one.na = (uint32_t)na;
    one.nb = (uint32_t)nb;
    one.n_out = (uint32_t)nout;
    if (!force_ntt && use_schoolbook(na, nb, nout)) return enqueue_schoolbook(nullptr, one, 1, g.stream, nullptr);
    const int logL = log_len(na, nb);
    size_t rw, sb;
    class_needs(logL, 1, (uint32_t)nout, square, rw, sb);
    int rc = ws_reserve(g_ws, rw, sb);
    if (rc) return rc;
    rc = enqueue_ntt_class(g_ws, nullptr, one, 1, logL, (uint32_t)nout, square, g.stream, 0, 0,
                           (uint32_t)std::max(na, nb));
    g.timing_pending = g.timing;
    return rc;
}

/* c = a*b on the engine stream with the static scratch as workspace: no
 * device or host allocation (products up to 2^NTT_NOALLOC_MAX_LOG
 * coefficients).  The caller holds g_lock. */
int enqueue_mul_noalloc(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb)
{
    if (!g.init || !g.noalloc_scratch) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((na && !d_a) || (nb && !d_b) || ((na + nb) && !d_c)) return fail(NTT_ERR_ARG, "null operand pointer");
    if (na > 0xffffffffu || nb > 0xffffffffu || log_len(na, nb) > NTT_NOALLOC_MAX_LOG)
        return fail(NTT_ERR_TOO_LARGE,
                    "product of %zu x %zu limbs needs a transform of 2^%d coefficients; the no-allocation "
                    "entry point takes at most 2^%d (na + nb <= %zu limbs)",
                    na, nb, log_len(na, nb), NTT_NOALLOC_MAX_LOG, (size_t)1 << (NTT_NOALLOC_MAX_LOG - 2));
    const size_t nout = na + nb;
    if (na == 0 || nb == 0 || use_schoolbook(na, nb, nout))
        return enqueue_mul(d_c, d_a, na, d_b, nb, 0);       /* no workspace on these paths */
    const int logL = log_len(na, nb);
    size_t rw, sb;
    class_needs(logL, 1, (uint32_t)nout, 0, rw, sb);
    if (rw > NOALLOC_RES_WORDS || sb > NOALLOC_CARRY_BYTES)
        return fail(NTT_ERR_TOO_LARGE, "product of %zu x %zu limbs exceeds the no-allocation scratch", na, nb);
    Workspace ws;
    ws.res = g.noalloc_scratch;
    ws.res_words = NOALLOC_RES_WORDS;
    ws.scratch = (uint8_t *)(g.noalloc_scratch + NOALLOC_RES_WORDS);
    ws.scratch_bytes = NOALLOC_CARRY_BYTES;
    MulDesc one{};
    one.a = d_a;
    one.b = d_b;
    one.c = d_c;
    one.na = (uint32_t)na;
    one.nb = (uint32_t)nb;
    one.n_out = (uint32_t)nout;
    g.dispatches++;
    g.n_marks = 0;
    HIP_TRY(mark("start"));
    g.no_alloc = 1;
    const int rc = enqueue_ntt_class(ws, nullptr, one, 1, logL, (uint32_t)nout, 0, g.stream, 0, 0,
                                     (uint32_t)std::max(na, nb));
    g.no_alloc = 0;
    /* ws does not own the scratch: nothing to release */
    ws.res = nullptr;
    ws.scratch = nullptr;
    g.timing_pending = g.timing;
    return rc;
}

/* ---- batches ------------------------------------------------------------- */

int plan_batch(const ntt_mul_desc *h, size_t n, BatchPlan &bp)
{
    static_assert(sizeof(ntt_mul_desc) == sizeof(MulDesc), "descriptor layouts differ");
    /* class key: -1 empty operand, 0 schoolbook, else log2 of the transform
     * length (<= NTT_MAX_LOG_LEN); a counting sort keeps the batch order
     * within a class and costs O(n) */
    constexpr int NK = NTT_MAX_LOG_LEN + 2;
    std::vector<uint8_t> key(n);
    /* large batches (the product tree's lower levels) are keyed and
     * scattered by chunks in parallel; the order within a class is the
     * batch order either way */
    const unsigned nch = std::max<unsigned>(1, std::min<unsigned>(host_threads(), (unsigned)(n >> 15)));
    std::vector<size_t> ccnt((size_t)nch * NK, 0);
    std::vector<uint32_t> cmax((size_t)nch * NK, 0), cmab((size_t)nch * NK, 0);
    std::vector<size_t> cbad(nch, SIZE_MAX);
    auto classify = [&](size_t lo, size_t hi, unsigned w) {
        size_t *cnt = &ccnt[(size_t)w * NK];
        uint32_t *mx = &cmax[(size_t)w * NK], *mab = &cmab[(size_t)w * NK];
        for (size_t i = lo; i < hi; i++) {
            MulDesc d;
            memcpy(&d, &h[i], sizeof d);
            if ((d.na && !d.a) || (d.nb && !d.b) || (d.n_out && !d.c) || (d.nadd && !d.add) || d.n_out < d.nadd ||
                (d.n_out == 0 && d.na && d.nb) || log_len(d.na, d.nb) > NTT_MAX_LOG_LEN) {
                cbad[w] = i;
                return;
            }
            int k = (d.na == 0 || d.nb == 0) ? -1 : use_schoolbook(d.na, d.nb, d.n_out) ? 0 : log_len(d.na, d.nb);
            key[i] = (uint8_t)(k + 1);
            cnt[k + 1]++;
            mx[k + 1] = std::max(mx[k + 1], std::max(d.n_out, d.na + d.nb));
            mab[k + 1] = std::max(mab[k + 1], std::max(d.na, d.nb));
        }
    };
    if (nch > 1) {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nch; w++) th.emplace_back(classify, n * w / nch, n * (w + 1) / nch, w);
        for (auto &x : th) x.join();
    } else {
        classify(0, n, 0);
    }
    for (unsigned w = 0; w < nch; w++)
        if (cbad[w] != SIZE_MAX) {
            MulDesc d;
            memcpy(&d, &h[cbad[w]], sizeof d);
            return check_desc(d);   /* the same test, with its message */
        }
    /* first slot of key k for chunk w: all chunks' keys below k, then the
     * earlier chunks' key k */
    size_t total[NK] = {};
    uint32_t maxo[NK] = {}, maxab[NK] = {};
    for (unsigned w = 0; w < nch; w++)
        for (int k = 0; k < NK; k++) {
            total[k] += ccnt[(size_t)w * NK + k];
            maxo[k] = std::max(maxo[k], cmax[(size_t)w * NK + k]);
            maxab[k] = std::max(maxab[k], cmab[(size_t)w * NK + k]);
        }
    size_t first[NK + 1] = {};
    for (int k = 0; k < NK; k++) first[k + 1] = first[k] + total[k];
    std::vector<size_t> slot((size_t)nch * NK);
    for (int k = 0; k < NK; k++) {
        size_t at = first[k];
        for (unsigned w = 0; w < nch; w++) {
            slot[(size_t)w * NK + k] = at;
            at += ccnt[(size_t)w * NK + k];
        }
    }
    bp.sorted.resize(n);
    bp.classes.clear();
    for (int k = 0; k < NK; k++)
        if (total[k]) {
            bp.classes.push_back({k - 1, (uint32_t)first[k], (uint32_t)total[k], maxo[k]});
            bp.classes.back().max_nab = maxab[k];
        }
    auto scatter = [&](size_t lo, size_t hi, unsigned w) {
        size_t *sl = &slot[(size_t)w * NK];
        for (size_t i = lo; i < hi; i++) memcpy(&bp.sorted[sl[key[i]]++], &h[i], sizeof(MulDesc));
    };
    if (nch > 1) {
        std::vector<std::thread> th;
        for (unsigned w = 0; w < nch; w++) th.emplace_back(scatter, n * w / nch, n * (w + 1) / nch, w);
        for (auto &x : th) x.join();
    } else {
        scatter(0, n, 0);
    }
    bp.res_words = bp.scratch = 0;
    for (auto &c : bp.classes) {
        if (c.logL <= 0) continue;
        /* b repeating with period 1 (one b for all) or count/2 (pairs i and
         * i + count/2, as the product tree lays out B = B_L B_R and
         * P = P_L B_R + P_R) */
        const MulDesc *cd = bp.sorted.data() + c.first;
        auto period_ok = [&](uint32_t per) {
            /* cd[i] against cd[i - per]: the same as against cd[i mod per] */
            std::atomic<bool> ok{true};
            par_chunks(c.count - per, [&](size_t lo, size_t hi, unsigned) {
                for (size_t i = per + lo; i < per + hi && ok.load(std::memory_order_relaxed); i++)
                    if (cd[i].b != cd[i - per].b || cd[i].nb != cd[i - per].nb) {
                        ok = false;
                        return;
                    }
            });
            return ok.load();
        };
        c.b_period = c.count;
        if (c.count > 1 && period_ok(1)) c.b_period = 1;
        else if (c.count > 1 && c.count % 2 == 0 && period_ok(c.count / 2)) c.b_period = c.count / 2;
        size_t rw, sb;
        class_needs(c.logL, c.count, c.max_nout, 0, rw, sb);
        bp.res_words = std::max(bp.res_words, rw);
        bp.scratch = std::max(bp.scratch, sb);
    }
    return NTT_OK;
}

__global__ void k_zero_products(const MulDesc *desc, uint32_t n)
{
    /* products with an empty operand: c = addend (or 0) */
    const MulDesc d = desc[blockIdx.y];
    for (size_t j = (size_t)blockIdx.x * 256 + threadIdx.x; j < d.n_out; j += (size_t)gridDim.x * 256)
        d.c[j] = (d.add && j < d.nadd) ? d.add[j] : 0;
    (void)n;
}

/* Enqueue every class of a planned batch whose descriptors are at d_desc. */
int enqueue_batch(Workspace &ws, const BatchPlan &bp, const MulDesc *d_desc, hipStream_t st)
{
    for (auto &c : bp.classes) {
        const MulDesc *dd = d_desc + c.first;
        int rc;
        if (c.logL < 0) {
            hipLaunchKernelGGL(k_zero_products, dim3(16, c.count), dim3(256), 0, st, dd, c.count);
            HIP_TRY(hipGetLastError());
            rc = NTT_OK;
        } else if (c.logL == 0) {
            rc = enqueue_schoolbook(dd, MulDesc{}, c.count, st, ws.flag);
        } else {
            rc = enqueue_ntt_class(ws, dd, MulDesc{}, c.count, c.logL, c.max_nout, 0, st, 0, c.b_period, c.max_nab);
        }
        if (rc) return rc;
    }
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
    int rcf = copy_d2h(&flag, g.flag, sizeof(int));
    if (rcf) return rcf;
    if (flag) {
        HIP_TRY(hipMemset(g.flag, 0, sizeof(int)));
        return fail(NTT_ERR_INTEGRITY, "CRT integrity check failed: a product coefficient exceeded its bound");
    }
    return NTT_OK;
}

/* ---- large host transfers through pinned staging ---------------------------
 * A pageable copy moves 18-30 GB/s over this box's link; a pinned one about
 * 57 GB/s.  Transfers above PIN_MIN go in PIN_CHUNK pieces through two pinned
 * buffers: the host side copies one piece (on up to 16 threads, which also
 * spreads the page faults of a fresh destination) while the DMA of the other
 * runs. */
#define PIN_CHUNK ((size_t)64 << 20)
#define PIN_MIN ((size_t)16 << 20)

static void par_memcpy(void *dst, const void *src, size_t n)
{
    par_chunks(n, [&](size_t lo, size_t hi, unsigned) {
        memcpy((char *)dst + lo, (const char *)src + lo, hi - lo);
    }, (size_t)1 << 22);
}

static int pin_ready()
{
    if (g.pin_bytes) return NTT_OK;
    for (int i = 0; i < 2; i++) {
        if (hipHostMalloc(&g.pin[i], PIN_CHUNK, hipHostMallocDefault) != hipSuccess)
            return fail(NTT_ERR_NOMEM, "pinned staging buffer");
        HIP_TRY(hipEventCreateWithFlags(&g.pin_ev[i], hipEventDisableTiming));
    }
    g.pin_bytes = PIN_CHUNK;
    return NTT_OK;
}

/* host -> device, ordered on the engine stream; returns once the host
 * buffer may be reused */
static int h2d_staged(void *dst, const void *src, size_t bytes)
{
    if (bytes < PIN_MIN) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g.stream));
        return NTT_OK;
    }
    int rc = pin_ready();
    if (rc) return rc;
    size_t k = 0;
    for (size_t off = 0; off < bytes; off += PIN_CHUNK, k++) {
        const size_t len = std::min(PIN_CHUNK, bytes - off);
        const int b = (int)(k & 1);
        HIP_TRY(hipEventSynchronize(g.pin_ev[b]));   /* its last DMA (this call's or an earlier one's) is done */
        par_memcpy(g.pin[b], (const char *)src + off, len);
        HIP_TRY(hipMemcpyAsync((char *)dst + off, g.pin[b], len, hipMemcpyHostToDevice, g.stream));
        HIP_TRY(hipEventRecord(g.pin_ev[b], g.stream));
    }
    return NTT_OK;
}

/* device -> host after everything queued on the engine stream; waits */
static int d2h_staged(void *dst, const void *src, size_t bytes)
{
    if (bytes < PIN_MIN) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        return NTT_OK;
    }
    int rc = pin_ready();
    if (rc) return rc;
    const size_t nchunk = (bytes + PIN_CHUNK - 1) / PIN_CHUNK;
    for (size_t k = 0; k <= nchunk; k++) {
        if (k < nchunk) {   /* start piece k (its buffer was emptied at step k - 1) */
            const size_t off = k * PIN_CHUNK, len = std::min(PIN_CHUNK, bytes - off);
            HIP_TRY(hipMemcpyAsync(g.pin[k & 1], (const char *)src + off, len, hipMemcpyDeviceToHost, g.stream));
            HIP_TRY(hipEventRecord(g.pin_ev[k & 1], g.stream));
        }
        if (k >= 1) {       /* empty piece k - 1 while piece k moves */
            const size_t j = k - 1, off = j * PIN_CHUNK, len = std::min(PIN_CHUNK, bytes - off);
            HIP_TRY(hipEventSynchronize(g.pin_ev[j & 1]));
            par_memcpy((char *)dst + off, g.pin[j & 1], len);
        }
    }
    return NTT_OK;
}

int host_mul(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb, int square, int force_ntt)
{
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if ((na && !a) || (nb && !b) || ((na + nb) && !c)) return fail(NTT_ERR_ARG, "null operand pointer");
    const size_t nout = na + nb;
    if (nout == 0) return NTT_OK;
    int rc = grow_dev(&g.d_a, &g.d_a_words, na ? na : 1);
    if (!rc && !square) rc = grow_dev(&g.d_b, &g.d_b_words, nb ? nb : 1);
    if (!rc) rc = grow_dev(&g.d_c, &g.d_c_words, nout);
    if (rc) return rc;
    if (na) rc = h2d_staged(g.d_a, a, na * 8);
    if (!rc && !square && nb) rc = h2d_staged(g.d_b, b, nb * 8);
    if (!rc) rc = enqueue_mul(g.d_c, g.d_a, na, square ? g.d_a : g.d_b, nb, square, force_ntt);
    if (!rc) rc = d2h_staged(c, g.d_c, nout * 8);
    if (rc) {
        (void)hipStreamSynchronize(g.stream);
        return rc;
    }
    return sync_check();
}

}  // namespace ntte

using namespace ntte;

extern "C" {

const char *ntt_engine_last_error(void) { return t_err; }

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
    return host_mul(c, a, na, b, nb, 0, 0);
}

int ntt_mul_limbs_ntt(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return host_mul(c, a, na, b, nb, 0, 1);
}

int ntt_sqr_limbs(uint64_t *c, const uint64_t *a, size_t na)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    return host_mul(c, a, na, a, na, 1, 0);
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

int ntt_mul_device_noalloc(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int rc = enqueue_mul_noalloc(d_c, d_a, na, d_b, nb);
    return rc ? rc : sync_check();
}

int ntt_sqr_device(uint64_t *d_c, const uint64_t *d_a, size_t na)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int rc = enqueue_mul(d_c, d_a, na, d_a, na, 1);
    return rc ? rc : sync_check();
}

int ntt_mul_batch_device_async(const ntt_mul_desc *descs, size_t n)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (n == 0) return NTT_OK;
    if (!descs) return fail(NTT_ERR_ARG, "null descriptor array");
    BatchPlan bp;
    int rc = plan_batch(descs, n, bp);
    if (rc) return rc;
    /* the descriptor buffer may still be read by an earlier batch */
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (g.d_desc_n < n) {
        if (g.d_desc) (void)hipFree(g.d_desc);   /* the stream was synchronized above */
        g.d_desc = nullptr;
        g.d_desc_n = 0;
        HIP_TRY(hipMalloc((void **)&g.d_desc, n * sizeof(MulDesc)));
        g.d_desc_n = n;
    }
    rc = copy_h2d(g.d_desc, bp.sorted.data(), n * sizeof(MulDesc));
    if (rc) return rc;
    rc = ws_reserve(g_ws, bp.res_words, bp.scratch);
    if (rc) return rc;
    g.dispatches += n;
    g.n_marks = 0;
    HIP_TRY(mark("start"));
    rc = enqueue_batch(g_ws, bp, g.d_desc, g.stream);
    g.timing_pending = g.timing;
    return rc;
}

int ntt_mul_batch_device(const ntt_mul_desc *descs, size_t n)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    int rc = ntt_mul_batch_device_async(descs, n);
    return rc ? rc : sync_check();
}

struct ntt_batch_graph {
    BatchPlan bp;
    Workspace ws;
    MulDesc *d_desc = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t n = 0;
    unsigned long gen = 0;      /* engine generation the graph was captured in */
};

void ntt_batch_graph_destroy(ntt_batch_graph *gr)
{
    if (!gr) return;
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (g.init) (void)hipStreamSynchronize(g.stream);
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    if (gr->d_desc) (void)hipFree(gr->d_desc);
    gr->ws.release();
    delete gr;
}

int ntt_batch_graph_create(const ntt_mul_desc *descs, size_t n, ntt_batch_graph **out)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!out) return fail(NTT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (n == 0 || !descs) return fail(NTT_ERR_ARG, "empty batch");
    ntt_batch_graph *gr = new ntt_batch_graph();
    gr->n = n;
    gr->gen = g_generation;
    int rc = plan_batch(descs, n, gr->bp);
    if (!rc) rc = ws_reserve(gr->ws, gr->bp.res_words ? gr->bp.res_words : 1, gr->bp.scratch ? gr->bp.scratch : 64);
    if (!rc && hipMalloc((void **)&gr->d_desc, n * sizeof(MulDesc)) != hipSuccess)
        rc = fail(NTT_ERR_NOMEM, "descriptor allocation failed");
    if (!rc) rc = copy_h2d(gr->d_desc, gr->bp.sorted.data(), n * sizeof(MulDesc));
    if (!rc && hipStreamSynchronize(g.stream) != hipSuccess) rc = fail(NTT_ERR_HIP, "stream sync failed");
    if (rc) {
        ntt_batch_graph_destroy(gr);
        return rc;
    }
    for (auto &c : gr->bp.classes) xtab_prepare(c.logL);   /* no allocation under capture */
    int saved_timing = g.timing;
    g.timing = 0;   /* no event nodes in the graph */
    hipError_t e = hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        g.timing = saved_timing;
        ntt_batch_graph_destroy(gr);
        return fail(NTT_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    rc = enqueue_batch(gr->ws, gr->bp, gr->d_desc, g.stream);
    e = hipStreamEndCapture(g.stream, &gr->graph);
    g.timing = saved_timing;
    if (!rc && e != hipSuccess) rc = fail(NTT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (!rc) {
        e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fail(NTT_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    if (rc) {
        ntt_batch_graph_destroy(gr);
        return rc;
    }
    *out = gr;
    return NTT_OK;
}

int ntt_batch_graph_launch(ntt_batch_graph *gr)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (!gr || !gr->exec) return fail(NTT_ERR_ARG, "null graph");
    if (gr->gen != g_generation)
        return fail(NTT_ERR_NOT_INIT, "batch graph captured before the engine was torn down and re-initialised");
    g.dispatches += gr->n;
    HIP_TRY(hipGraphLaunch(gr->exec, g.stream));
    return NTT_OK;
}

int ntt_engine_get_threshold(void) { return g.threshold; }

int ntt_engine_set_pass_cap(int log_len, int cap)
{
    if (log_len < 1 || log_len > NTT_MAX_LOG_LEN || (cap != 0 && (cap < 4 || cap > FT_MAX_LOGS)))
        return fail(NTT_ERR_ARG, "pass cap %d for 2^%d out of range", cap, log_len);
    g.max_logs[log_len] = (uint8_t)cap;
    g.split_n[log_len] = 0;       /* the latest setting decides the plan */
    return NTT_OK;
}

int ntt_engine_get_pass_cap(int log_len)
{
    if (log_len < 1 || log_len > NTT_MAX_LOG_LEN) return 0;
    return g.max_logs[log_len] ? g.max_logs[log_len] : FT_MAX_LOGS;
}

int ntt_engine_set_pass_split(int log_len, const int *log_s, int npass)
{
    if (log_len < 1 || log_len > NTT_MAX_LOG_LEN) return fail(NTT_ERR_ARG, "pass split for 2^%d: no such length", log_len);
    if (npass == 0) {
        g.split_n[log_len] = 0;
        return NTT_OK;
    }
    if (npass < 1 || npass > 8 || !log_s) return fail(NTT_ERR_ARG, "pass split for 2^%d: %d passes", log_len, npass);
    int sum = 0;
    for (int i = 0; i < npass; i++) {
        /* passes of 2^3..2^10 points (a single pass may be shorter) */
        if (log_s[i] < (npass == 1 ? 1 : 3) || log_s[i] > FT_MAX_LOGS)
            return fail(NTT_ERR_ARG, "pass split for 2^%d: a pass of 2^%d points", log_len, log_s[i]);
        sum += log_s[i];
    }
    if (sum != log_len) return fail(NTT_ERR_ARG, "pass split for 2^%d sums to 2^%d", log_len, sum);
    for (int i = 0; i < npass; i++) g.split[log_len][i] = (uint8_t)log_s[i];
    g.split_n[log_len] = (uint8_t)npass;
    return NTT_OK;
}

int ntt_engine_get_plan(int log_len, int *log_s, int max)
{
    if (log_len < 1 || log_len > NTT_MAX_LOG_LEN) return 0;
    const PassPlan pp = plan_passes(log_len);
    for (int i = 0; i < pp.npass && i < max; i++) log_s[i] = (int)pp.logS[i];
    return pp.npass;
}
int ntt_engine_set_tile_batch(int tiles)
{
    if (tiles < 1 || tiles > 64) return fail(NTT_ERR_ARG, "tile batch %d out of range [1, 64]", tiles);
    g.tile_batch = tiles;
    return NTT_OK;
}
int ntt_engine_get_tile_batch(void) { return g.tile_batch; }
int ntt_engine_set_xtab(int on)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    const int old = g.use_xtab;
    g.use_xtab = on ? 1 : 0;
    return old;
}
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
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    return copy_h2d(dst, src, bytes);
}
int ntt_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    return copy_d2h(dst, src, bytes);
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
    return copy_d2h(out, g.tw, need * 8);
}

}  // extern "C"
