#pragma once
/*
 * ntt_engine.h — C ABI of the MI355X (gfx950) 3-prime CRT-NTT multiply engine.
 *
 * Status-returning entry points.  The reference's fail-stop entry points
 * (multiply.h: bigint_mul, ntt_bigint_mul, ntt_mul_init, ...) are thin
 * wrappers over these that print the error and exit(1) as the reference does.
 *
 * Operands and products are little-endian u64 limb arrays (a u64 limb array
 * is also the 32-bit-limb array the coefficient split reads).  A product of
 * na and nb limbs always has na + nb limbs (untrimmed).
 *
 * Threading: one engine per process on one GPU.  Every entry point takes the
 * engine lock, so concurrent callers serialize, as the reference's
 * ntt_gpu_lock does (ntt/lib/arith/multiply.c:22-27,86-88).
 */
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NTT_OK = 0,
    NTT_ERR_NOT_INIT = 1,     /* engine could not be initialised (no GPU, HIP error at init) */
    NTT_ERR_TOO_LARGE = 2,    /* transform would exceed 2^30 coefficients (cf. ntt_mul.hip:1013-1022) */
    NTT_ERR_HIP = 3,          /* a HIP runtime call failed (cf. HIP_CHECK, ntt_mul.hip:101-109) */
    NTT_ERR_ARG = 4,          /* invalid argument (null pointer with nonzero size, ...) */
    NTT_ERR_INTEGRITY = 5,    /* CRT integrity check failed: a corrupted product */
    NTT_ERR_NOMEM = 6         /* device or host allocation failed */
} ntt_status;

/* Message of the last error on the calling thread (never NULL). */
const char *ntt_engine_last_error(void);

/* Initialise on HIP device `device` (idempotent).  Builds the twiddle and
 * Shoup tables on the device.  No counterpart needed by callers of the
 * reference API: ntt_mul_init calls it.  When the environment variable
 * NTT_AMD_TUNING names a tuning file (ntt_engine_autotune_save's format), it
 * is loaded here; a file that cannot be read or parsed fails the init. */
int ntt_engine_init(int device);
/* Release every device buffer and the stream (cf. ntt_mul_teardown_impl). */
void ntt_engine_teardown(void);
int ntt_engine_is_init(void);

/* log2 of the transform length used for a product of na and nb limbs
 * (16-bit coefficients: least power of two >= 4(na+nb)). */
int ntt_engine_log_len(size_t na, size_t nb);

/* Host-buffer multiply / square: c[0..na+nb) = a*b.  Synchronous. */
int ntt_mul_limbs(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb);
int ntt_sqr_limbs(uint64_t *c, const uint64_t *a, size_t na);
/* The same product on the CRT-NTT path whatever the operand sizes (the
 * status-returning form of ntt_bigint_mul, multiply.h:46): the threshold is
 * neither read nor changed, so concurrent callers of bigint_mul keep theirs. */
int ntt_mul_limbs_ntt(uint64_t *c, const uint64_t *a, size_t na, const uint64_t *b, size_t nb);

/* Device-resident multiply / square on the engine stream.  d_c, d_a, d_b are
 * device pointers; d_c must not alias d_a or d_b.  Synchronous: returns after
 * the product is complete and its CRT integrity check has been read. */
int ntt_mul_device(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb);
int ntt_sqr_device(uint64_t *d_c, const uint64_t *d_a, size_t na);
/* Same, enqueued only (no host wait, no integrity read); pair with
 * ntt_engine_sync(), which waits and reports a failed integrity check. */
int ntt_mul_device_async(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb);
int ntt_sqr_device_async(uint64_t *d_c, const uint64_t *d_a, size_t na);
int ntt_engine_sync(void);

/* No counterpart in the reference.  For callers that cannot allocate device
 * or host memory (code under hipGraph stream capture, tight serving loops,
 * harnesses that forbid hipMalloc): ntt_mul_device with the same arguments,
 * results and synchronous integrity check, whose workspace is a static
 * per-device scratch array of 24.8 MB instead of engine allocations.
 * It calls no hipMalloc, hipFree or allocating library routine; the engine
 * must have been initialised (ntt_engine_init allocates its tables there).
 * Limit: transforms of at most 2^19 coefficients (4 (na + nb) <= 2^19, e.g.
 * two 2^22-bit operands); a larger product returns NTT_ERR_TOO_LARGE.  Every
 * call on a GPU shares that GPU's array and leaves its contents behind: calls
 * are serialised by the engine lock and stream, and anything else that
 * enqueued work on the same array concurrently would race. */
int ntt_mul_device_noalloc(uint64_t *d_c, const uint64_t *d_a, size_t na, const uint64_t *d_b, size_t nb);

/* ---- batches ---------------------------------------------------------------
 * One product of a batch: c[0..n_out) = a*b + add (add may be NULL, nadd 0).
 * All pointers are device pointers; n_out >= na + nb and >= nadd, and the
 * result must fit n_out limbs (else NTT_ERR_INTEGRITY at the next sync).
 * No counterpart in the reference (every reference multiply is one call):
 * this is what the level-batched product tree and the serving loop use. */
typedef struct {
    const uint64_t *a, *b;
    uint64_t *c;
    const uint64_t *add;
    uint32_t na, nb, nadd, n_out;
} ntt_mul_desc;

/* Products are grouped by algorithm and transform length (schoolbook class,
 * one class per NTT length); each class runs as one launch per kernel stage. */
int ntt_mul_batch_device(const ntt_mul_desc *descs, size_t n);
int ntt_mul_batch_device_async(const ntt_mul_desc *descs, size_t n);

/* The same batch captured once into a hipGraph with its own workspace, then
 * replayed with one hipGraphLaunch per call (pointers and sizes are fixed at
 * capture; operand contents may change between launches). */
typedef struct ntt_batch_graph ntt_batch_graph;
int ntt_batch_graph_create(const ntt_mul_desc *descs, size_t n, ntt_batch_graph **out);
int ntt_batch_graph_launch(ntt_batch_graph *g);   /* async; pair with ntt_engine_sync() */
void ntt_batch_graph_destroy(ntt_batch_graph *g);

/* Schoolbook / NTT crossover in limbs (both operands at or below it use the
 * GPU schoolbook kernel).  Cf. multiply.h:23-31. */
int ntt_engine_get_threshold(void);
void ntt_engine_set_threshold(int limbs);

/* Largest sub-transform (log2, 4..10; 0 = default 10) of each pass for
 * transforms of length 2^log_len, split into near-equal passes.  Setting it
 * clears an explicit split of that length (below). */
int ntt_engine_set_pass_cap(int log_len, int cap);
int ntt_engine_get_pass_cap(int log_len);
/* An explicit pass split for transforms of length 2^log_len: npass passes
 * of 2^log_s[0], 2^log_s[1], ... points (first pass first; each 3..10, or one
 * pass of log_len; the sum log_len, npass <= 8).  It takes precedence over the pass cap; npass = 0
 * clears it.  Set by the autotuner.  A change applies to later calls: batch
 * graphs and product trees keep the plan they were captured with, and a
 * server rebuilds a resident spectrum made under another plan. */
int ntt_engine_set_pass_split(int log_len, const int *log_s, int npass);
/* The plan in effect for length 2^log_len: writes up to max pass sizes
 * (log2) to log_s and returns the number of passes (0 for a bad length). */
int ntt_engine_get_plan(int log_len, int *log_s, int max);
/* Most tiles (4096 residues each) one block of a pass transforms in turn
 * (default 1, range 1..64; applied to large transforms only, small ones use
 * 1).  Some passes take more whatever the setting: the forward table pass,
 * the 64- and 512-point fused pass and the inverse table pass 8, the other
 * inverse passes that apply their cross twiddles at the load 4. */
int ntt_engine_set_tile_batch(int tiles);
int ntt_engine_get_tile_batch(void);
/* Cross twiddles of passes whose column block has at most 2^20 residues:
 * from precomputed per-block tables (1, the default; one Montgomery product
 * per residue) or from Montgomery chains inside the pass (0).  With 1, the
 * chains of the other passes (the first forward pass, the inverse passes of
 * larger blocks) take their ratio from a per-column table and step by Shoup
 * products; with 0 they compute it.  Both give the same products.  Returns
 * the previous setting. */
int ntt_engine_set_xtab(int on);

/* Autotuner: times the schoolbook/NTT crossover on batches of equal-size
 * products, the pass splits of lengths 2^min_log..2^max_log (11..28: every
 * split into the fewest passes of at most 2^10 points, and up to 2^20 one
 * pass more, each pass of 2^3 points or more) and, when max_log >= 24, the
 * tile batch (1, 2 or 4) on the longest length, and sets the fastest.  Some
 * seconds.  save/load keep the result in a small JSON file ({"threshold": T,
 * "tile_batch": B, "pass_cap": {"<log_len>": cap, ...}, "pass_split":
 * {"<log_len>": [s0, s1, ...], ...}, "split_us": {"<log_len>": [[[s0, ...],
 * us], ...]}, "crossover_us": [[n, schoolbook, ntt, karatsuba_lower_bound],
 * ...]}); the loader reads threshold, tile_batch, pass_cap and pass_split (a
 * file without tile_batch, pass_cap or pass_split leaves the current value)
 * and skips other keys.  split_us records the microseconds per product of
 * every split tried.
 * crossover_us records, per operand size n, microseconds per product of a
 * batch of 256 n x n products on each algorithm and, as a lower bound for one
 * Karatsuba level, three products of ceil(n/2) + 1 limbs on the faster one. */
int ntt_engine_autotune(int min_log, int max_log);
int ntt_engine_autotune_save(const char *path);
int ntt_engine_autotune_load(const char *path);

/* Number of device multiplies issued since init (cf. ntt_mul.hip:595). */
unsigned long ntt_engine_dispatch_count(void);

/* Timing of the NTT stages (forward passes, pointwise, inverse passes) of the
 * last device multiply, measured with HIP events on the engine stream;
 * enabled by ntt_engine_set_timing(1).  Milliseconds; 0 when disabled. */
void ntt_engine_set_timing(int on);
double ntt_engine_last_ntt_ms(void);
double ntt_engine_last_total_ms(void);
/* Per-launch durations of the last NTT multiply (timing enabled): up to max
 * entries; names[32*i] holds kernel i's name, ms[i] its duration.  Returns
 * the count. */
int ntt_engine_kernel_times(char *names, double *ms, int max);

/* Harness helpers: device memory owned by the caller. */
void *ntt_device_alloc(size_t bytes);
void ntt_device_free(void *p);
int ntt_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ntt_memcpy_d2h(void *dst, const void *src, size_t bytes);

/* Copy of the on-device twiddle table (3 primes x {w^j, w^(j 2^15)} x 2^15
 * Shoup pairs of u64, i.e. 3*2*2^15*2 words) for verification. */
int ntt_engine_read_twiddles(uint64_t *out, size_t n_words);

#ifdef __cplusplus
}
#endif
