#pragma once
/*
 * ntt_serve.h — continuous-batching multiply service on one GPU (no
 * counterpart in the reference, whose multiplies are serialized one at a time
 * behind a mutex: ntt/lib/arith/multiply.c:22-27,86-88).
 *
 * A server owns a scheduler thread, an HBM arena of hbm_budget bytes and up
 * to max_inflight executor slots (one HIP stream each).  Callers on any
 * thread submit requests (device-resident operands and result buffers, as in
 * ntt_mul_desc) and poll or wait for them.  The scheduler admits queued
 * requests FIFO into batches whenever a slot is free and the batch's working
 * memory fits what the batches in flight leave of the arena, so requests that
 * arrive while others run are admitted as slots and memory free up, and their
 * work overlaps on the device.
 *
 * Resident operands: an operand pinned with ntt_serve_pin (pointer and limb
 * count, contents unchanged until unpinned) keeps its forward transforms on
 * the device, one per transform length it is used at, made on first use and
 * counted against the arena.  Requests whose a or b is that operand skip its
 * transform.
 */
#include <stddef.h>
#include <stdint.h>

#include "ntt_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ntt_server ntt_server;

typedef struct {
    size_t hbm_budget;        /* bytes of device memory for requests' working memory and resident spectra */
    int max_inflight;         /* batches on the device at once, each on its own stream (1..16) */
    uint32_t max_batch;       /* most requests per batch (0: 4096) */
} ntt_serve_config;

typedef struct {
    double makespan_s;        /* first submit to last completion */
    size_t batches;           /* batches launched */
    size_t max_inflight;      /* most batches on the device at once */
    size_t completed;         /* requests completed */
    size_t spectrum_hits;     /* requests that used a resident spectrum */
    size_t spectra_made;      /* resident spectra transformed */
} ntt_serve_stats;

/* Open a server on the initialised engine.  The arena is allocated here. */
int ntt_serve_open(const ntt_serve_config *cfg, ntt_server **out);
/* Queue one request; thread-safe, returns at once with its ticket.  A
 * request whose working memory exceeds the arena is rejected
 * (NTT_ERR_TOO_LARGE), as are malformed descriptors (NTT_ERR_ARG). */
int ntt_serve_submit(ntt_server *s, const ntt_mul_desc *req, uint64_t *ticket);
/* 1 when the request is complete (its status in *status, its completion time
 * in seconds since open in *t_done; either may be NULL; the ticket is then
 * retired), 0 while it is queued or running, or a negative ntt_status for an
 * unknown ticket. */
int ntt_serve_poll(ntt_server *s, uint64_t ticket, int *status, double *t_done);
/* Block until the request completes (then as ntt_serve_poll). */
int ntt_serve_wait(ntt_server *s, uint64_t ticket, int *status, double *t_done);
/* Keep the forward transforms of operand (d_op, n_limbs) resident; its
 * contents must not change until ntt_serve_unpin.  Unpin fails with
 * NTT_ERR_ARG while requests that use the operand are queued or running. */
int ntt_serve_pin(ntt_server *s, const uint64_t *d_op, uint32_t n_limbs);
int ntt_serve_unpin(ntt_server *s, const uint64_t *d_op, uint32_t n_limbs);
int ntt_serve_get_stats(ntt_server *s, ntt_serve_stats *stats);
/* Finish every queued request, then release the server. */
void ntt_serve_close(ntt_server *s);

/* Replay an arrival stream through a server: request i (device pointers, as
 * in a batch descriptor) is submitted arrival[i] seconds after the start
 * (non-decreasing) by the calling thread; t_done[i] receives its completion
 * time in seconds since the start. */
int ntt_serve_replay(const ntt_mul_desc *reqs, const double *arrival, size_t n, size_t hbm_budget, int slots,
                     double *t_done, ntt_serve_stats *stats);

#ifdef __cplusplus
}
#endif
