#pragma once
/*
 * ntt_serve.h — continuous-batching multiply service (no counterpart in the
 * reference, whose multiplies are serialized one at a time behind a mutex).
 */
#include <stddef.h>

#include "ntt_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double makespan_s;        /* first arrival to last completion */
    size_t batches;           /* batches launched */
    size_t max_inflight;      /* most batches on the device at once */
} ntt_serve_stats;

/* Replay an arrival stream: request i (device pointers, as in a batch
 * descriptor) arrives arrival[i] seconds after the start (non-decreasing).
 * Arrived requests are admitted FIFO into batches on `slots` executor slots
 * (1..16), each with its own stream and a 1/slots share of hbm_budget bytes
 * of working memory; a batch launches as soon as a slot is free and its
 * requests' working memory fits what the busy slots leave of the budget.
 * t_done[i] receives request i's completion time (seconds since the start),
 * observed when its batch's completion event fires. */
int ntt_serve_replay(const ntt_mul_desc *reqs, const double *arrival, size_t n, size_t hbm_budget, int slots,
                     double *t_done, ntt_serve_stats *stats);

#ifdef __cplusplus
}
#endif
