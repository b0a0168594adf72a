/*
 * engine_internal.h — engine state and executor functions shared by the
 * engine's translation units (ntt_engine.hip, tree.hip).  Not part of the ABI.
 */
#ifndef CRYPTOGRAPHY_AMD_ENGINE_INTERNAL_H
#define CRYPTOGRAPHY_AMD_ENGINE_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/ntt_engine.h"
#include "ntt_kernels.h"

#define NTT_MAX_MARKS 40
#include "modarith.h"

namespace ntte {

struct Engine {
    int init = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    TwPair *tw = nullptr;
    TwPair *small_tw = nullptr;     /* w_4096^e per prime: stage twiddles */
    int *flag = nullptr;
    MulDesc *d_desc = nullptr;      /* descriptor array of the last batch call */
    size_t d_desc_n = 0;
    /* host staging for the host-buffer API */
    uint64_t *d_a = nullptr, *d_b = nullptr, *d_c = nullptr;
    size_t d_a_words = 0, d_b_words = 0, d_c_words = 0;
    /* two pinned host buffers for chunked transfers of large operands
     * (allocated on first use), each with the event of its last copy */
    void *pin[2] = {nullptr, nullptr};
    size_t pin_bytes = 0;
    hipEvent_t pin_ev[2] = {};
    int threshold = 64;
    /* largest sub-transform (log2) per pass, per transform length 2^l;
     * 0 = default (FT_MAX_LOGS).  Set by the autotuner. */
    uint8_t max_logs[NTT_MAX_LOG_LEN + 1] = {};
    /* explicit pass splits (autotuned): split_n[l] passes of 2^split[l][i]
     * points for transforms of length 2^l, first pass first; 0 = by max_logs */
    uint8_t split_n[NTT_MAX_LOG_LEN + 1] = {};
    uint8_t split[NTT_MAX_LOG_LEN + 1][8] = {};
    /* most tiles one block of a forward / inverse pass works through */
    int tile_batch = 1;
    /* cross-twiddle tables of passes whose S x T block has at most
     * 2^NTT_XTAB_MAX_LOG residues, by (log2 S, log2 S*T), made on first use;
     * use_xtab = 0 keeps every pass on the in-round Montgomery chains */
    uint64_t *xtab[FT_MAX_LOGS + 1][NTT_XTAB_MAX_LOG + 1] = {};
    /* chain ratio tables (B per column with its Shoup partner) of passes of
     * 2^logS rows and 2^logN residues per block, made on first use; they
     * follow use_xtab */
    TwPair *btab[FT_MAX_LOGS + 1][NTT_MAX_LOG_LEN + 1] = {};
    int use_xtab = 1;
    /* set while a call that must not allocate runs: xtab_for then hands out
     * only the tables that already exist (the passes fall back to chains) */
    int no_alloc = 0;
    /* device address of this device's copy of the no-allocation scratch
     * array (hipGetSymbolAddress at init; the array's name is not it) */
    uint64_t *noalloc_scratch = nullptr;
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
};

extern Engine g;
extern std::recursive_mutex g_lock;
/* Incremented by every successful engine initialisation.  Handles that
 * captured the engine's tables (batch graphs, product trees) record it and
 * refuse to run after a teardown and re-initialisation freed those tables. */
extern unsigned long g_generation;

/* Device workspace of one executor (the engine's own, or a graph's). */
struct Workspace {
    uint64_t *res = nullptr;
    size_t res_words = 0;
    uint8_t *scratch = nullptr;
    size_t scratch_bytes = 0;
    int *flag = nullptr;          /* integrity flag for work on this workspace (null: the engine's) */
    void release()
    {
        if (res) (void)hipFree(res);
        if (scratch) (void)hipFree(scratch);
        res = nullptr;
        scratch = nullptr;
        res_words = scratch_bytes = 0;
    }
};

extern Workspace g_ws;

struct PassPlan {
    int npass;
    uint32_t logS[8];
};
/* npass and every pass size of the plan in one word (a spectrum made with
 * one plan is only read back under the same plan) */
uint64_t plan_signature(int logL);

/* Allocator whose resize() leaves trivially constructible elements
 * uninitialized: descriptor arrays of millions of entries are written in
 * full right after (in parallel), so zero-filling them first on one thread
 * only costs time. */
template <class T>
struct uninit_alloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = uninit_alloc<U>;
    };
    uninit_alloc() = default;
    template <class U>
    uninit_alloc(const uninit_alloc<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept
    {
        ::new ((void *)p) U;
    }
    template <class U, class... Args>
    void construct(U *p, Args &&...args)
    {
        ::new ((void *)p) U(std::forward<Args>(args)...);
    }
};

struct BatchPlan {
    struct Class {
        int logL;            /* 0: schoolbook */
        uint32_t first, count, max_nout;
        /* products i and i + b_period have the same b (pointer and length):
         * 1 when all share one, count when none repeats */
        uint32_t b_period = 0;
        uint32_t max_nab = 0;   /* longest operand (limbs) */
    };
    std::vector<MulDesc, uninit_alloc<MulDesc>> sorted;   /* descriptors grouped by class */
    std::vector<Class> classes;
    size_t res_words = 0, scratch = 0;
};

int fail(int code, const char *fmt, ...);
#define HIP_TRY(x)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(NTT_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

hipError_t mark(const char *name);
int grow_dev(uint64_t **buf, size_t *have, size_t words);
int ws_reserve(Workspace &ws, size_t res_words, size_t scratch_bytes);
int log_len(size_t na, size_t nb);
/* Row stride (words, with pads) of a class of nprod products of length 2^logL. */
size_t row_len(int logL, uint32_t nprod);
/* Residues per row (nprod * L rounded up to 4096) and the row pad exponent. */
size_t row_residues(int logL, uint32_t nprod);
uint32_t pad_log_for(int logL);
/* Workspace of a size class, and its transforms (+ Garner and carry unless
 * residues_in: then the rows hold residues on entry and on exit). */
void class_needs(int logL, uint32_t nprod, uint32_t max_nout, int square, size_t &res_words, size_t &scratch);
/* Cross-twiddle table pair (forward, inverse) of an S x T pass, or nullptr
 * (tables off, block too large, or not yet made while a stream captures). */
const uint64_t *xtab_for(int logS, int logT, hipStream_t st);
const TwPair *btab_for(int logS, int logT, hipStream_t st);
/* Make the tables the passes of a length-2^logL transform use (before a
 * graph capture, which cannot allocate). */
void xtab_prepare(int logL);
/* b_res: b's rows after all its forward passes but the last (a resident
 * spectrum, spectrum_rows_len words per prime), shared by every product of the
 * class; only a is transformed and the workspace holds one operand's rows.
 * fwd_only: run only those forward passes, of operand a (square = 1), into
 * the workspace rows: that makes such a spectrum.  Both need
 * spectrum_cacheable(logL). */
int enqueue_ntt_class(Workspace &ws, const MulDesc *d_desc, const MulDesc &one, uint32_t nprod, int logL,
                      uint32_t max_nout, int square, hipStream_t st, int residues_in = 0, uint32_t b_period = 0,
                      uint32_t max_nab = 0, const uint64_t *b_res = nullptr, int fwd_only = 0);
/* Whether a length-2^logL transform keeps an operand's partial spectrum in
 * whole tiles between passes (two or more passes, L >= 4096). */
bool spectrum_cacheable(int logL);
/* Words per prime of such a spectrum (one product's padded row). */
size_t spectrum_rows_len(int logL);
int plan_batch(const ntt_mul_desc *h, size_t n, BatchPlan &bp);

/* Host threads for planning work (descriptor sorting, tree construction):
 * at most 16, the share of a GPU's host cores this library assumes. */
inline unsigned host_threads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hc ? hc : 1u));
}
/* f(begin, end, worker) over contiguous chunks of [0, n), one per thread
 * (inline for n below two chunks of min_chunk).  Returns the chunk count. */
template <class F>
unsigned par_chunks(size_t n, F &&f, size_t min_chunk = (size_t)1 << 15)
{
    const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>(host_threads(), n / min_chunk));
    if (nt <= 1) {
        f((size_t)0, n, 0u);
        return 1;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (unsigned w = 0; w < nt; w++) th.emplace_back([&f, n, nt, w] { f(n * w / nt, n * (w + 1) / nt, w); });
    for (auto &x : th) x.join();
    return nt;
}
int enqueue_batch(Workspace &ws, const BatchPlan &bp, const MulDesc *d_desc, hipStream_t st);
/* Whether a product of na x nb limbs runs on the schoolbook kernel. */
bool use_schoolbook(size_t na, size_t nb, size_t n_out);
int sync_check();
/* force_ntt: the CRT-NTT path whatever the operand sizes (no schoolbook). */
int enqueue_mul(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb, int square,
                int force_ntt = 0);
int copy_h2d(void *dst, const void *src, size_t bytes);
int copy_d2h(void *dst, const void *src, size_t bytes);

}  // namespace ntte

#endif
