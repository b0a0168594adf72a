/*
 * autotune.hip — measures and sets the engine's tunable parameters:
 *   - the schoolbook / NTT crossover (limbs per operand), timed on batches of
 *     equal-size products with each algorithm;
 *   - the pass split of every transform length (the largest sub-transform per
 *     pass, 2^8 .. 2^10), timed on one product per length;
 *   - the tile shape of the long passes (tiles per block), timed on the
 *     longest length.
 * Results can be saved to and loaded from a small JSON file.
 * (The reference fixes its crossover at 64 limbs, ntt/lib/arith/multiply.h:23-25,
 * and picks kernels by log_n, ntt/lib/arith/ntt_mul.hip:630-645.)
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <vector>

#include "engine_internal.h"

using namespace ntte;

namespace {

double wall()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

__global__ void k_fill_pattern(uint64_t *p, size_t n, uint64_t seed)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        p[i] = z ^ (z >> 31);
    }
}

/* Seconds per product for `count` products of n x n limbs (median of reps). */
int time_batch(uint64_t *buf, size_t n, int count, int reps, double *sec)
{
    std::vector<ntt_mul_desc> d(count);
    for (int i = 0; i < count; i++) {
        uint64_t *a = buf + (size_t)i * 4 * n;
        d[i] = ntt_mul_desc{a, a + n, a + 2 * n, nullptr, (uint32_t)n, (uint32_t)n, 0, (uint32_t)(2 * n)};
    }
    std::vector<double> t;
    for (int r = 0; r < reps + 1; r++) {
        double t0 = wall();
        int rc = ntt_mul_batch_device(d.data(), d.size());
        if (rc) return rc;
        if (r) t.push_back(wall() - t0);
    }
    std::sort(t.begin(), t.end());
    *sec = t[t.size() / 2] / count;
    return NTT_OK;
}

}  // namespace

extern "C" {

int ntt_engine_autotune_save(const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) return fail(NTT_ERR_ARG, "cannot write %s", path);
    fprintf(f, "{\"threshold\": %d, \"tile_batch\": %d, \"pass_cap\": {", g.threshold, g.tile_batch);
    int first = 1;
    for (int l = 1; l <= NTT_MAX_LOG_LEN; l++)
        if (g.max_logs[l]) {
            fprintf(f, "%s\"%d\": %d", first ? "" : ", ", l, g.max_logs[l]);
            first = 0;
        }
    fprintf(f, "}}\n");
    fclose(f);
    return NTT_OK;
}

/* Minimal reader for the file written above. */
int ntt_engine_autotune_load(const char *path)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    FILE *f = fopen(path, "r");
    if (!f) return fail(NTT_ERR_ARG, "cannot read %s", path);
    char buf[4096];
    size_t n = fread(buf, 1, sizeof buf - 1, f);
    fclose(f);
    buf[n] = 0;
    const char *p = strstr(buf, "\"threshold\":");
    if (!p) return fail(NTT_ERR_ARG, "%s: no threshold", path);
    int thr = atoi(p + 12);
    int tb = g.tile_batch;
    const char *pt = strstr(buf, "\"tile_batch\":");
    if (pt) tb = atoi(pt + 13);
    if (tb < 1 || tb > 64) return fail(NTT_ERR_ARG, "%s: tile_batch %d out of range", path, tb);
    uint8_t caps[NTT_MAX_LOG_LEN + 1] = {};
    p = strstr(buf, "\"pass_cap\":");
    if (p) {
        p = strchr(p, '{');
        while (p && *p && *p != '}') {
            const char *q = strchr(p, '"');
            if (!q) break;
            int l = atoi(q + 1);
            const char *c = strchr(q + 1, ':');
            if (!c) break;
            int cap = atoi(c + 1);
            if (l >= 1 && l <= NTT_MAX_LOG_LEN && cap >= 4 && cap <= FT_MAX_LOGS) caps[l] = (uint8_t)cap;
            p = strchr(c, ',');
            if (!p) break;
        }
    }
    if (thr < 0 || thr > SB_MAX_OUT / 2) return fail(NTT_ERR_ARG, "%s: threshold %d out of range", path, thr);
    g.threshold = thr;
    g.tile_batch = tb;
    memcpy(g.max_logs, caps, sizeof caps);
    return NTT_OK;
}

/* Measure and set the crossover, the pass caps for transform lengths
 * 2^min_log .. 2^max_log and (max_log >= 24) the tile batch.  Takes a few
 * seconds. */
int ntt_engine_autotune(int min_log, int max_log)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (min_log < 11) min_log = 11;
    if (max_log > 28) max_log = 28;
    /* operand buffer: big enough for the largest single product and the batches */
    size_t words = std::max((size_t)1 << (max_log - 1), (size_t)256 * 4 * 128);
    uint64_t *buf = nullptr;
    if (hipMalloc((void **)&buf, words * 8) != hipSuccess) return fail(NTT_ERR_NOMEM, "autotune buffer");
    hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, g.stream, buf, words,
                       (uint64_t)12345);
    int rc = NTT_OK;
    /* 1. schoolbook / NTT crossover */
    const int saved = g.threshold;
    int best_thr = 0;
    for (int n : {8, 16, 24, 32, 48, 64, 96, 128}) {
        double t_sb = 0, t_ntt = 0;
        g.threshold = n;
        rc = time_batch(buf, n, 256, 3, &t_sb);
        if (rc) break;
        g.threshold = 0;
        rc = time_batch(buf, n, 256, 3, &t_ntt);
        if (rc) break;
        if (t_sb <= t_ntt) best_thr = n;
    }
    g.threshold = rc ? saved : best_thr;
    /* 2. pass caps: one product of L/8 x L/8 limbs per length */
    for (int l = min_log; l <= max_log && !rc; l++) {
        size_t n = (size_t)1 << (l - 3);
        if (4 * n > words) break;
        double best = 1e30;
        int best_cap = FT_MAX_LOGS;
        int last_npass = -1;
        for (int cap = FT_MAX_LOGS; cap >= 8; cap--) {
            int npass = (l + cap - 1) / cap;
            if (npass == last_npass) continue;   /* same plan as a larger cap */
            last_npass = npass;
            g.max_logs[l] = (uint8_t)cap;
            std::vector<double> t;
            for (int r = 0; r < 4 && !rc; r++) {
                double t0 = wall();
                rc = enqueue_mul(buf + 2 * n, buf, n, buf + n, n, 0);
                if (!rc) rc = sync_check();
                if (r) t.push_back(wall() - t0);
            }
            if (rc) break;
            std::sort(t.begin(), t.end());
            if (t[t.size() / 2] < best) {
                best = t[t.size() / 2];
                best_cap = cap;
            }
        }
        g.max_logs[l] = (uint8_t)best_cap;
    }
    /* 3. tile shape: tiles per block of the forward / inverse passes, timed on
     * the longest transform.  A block takes more than one tile only when a
     * pass has over 4096 blocks' worth of tiles (lengths of 2^24 and up). */
    if (!rc && max_log >= 24) {
        size_t n = std::min((size_t)1 << (max_log - 3), words / 4);
        double best = 1e30;
        const int saved_tb = g.tile_batch;
        int best_tb = 1;
        for (int tb : {1, 2, 4}) {
            g.tile_batch = tb;
            std::vector<double> t;
            for (int r = 0; r < 4 && !rc; r++) {
                double t0 = wall();
                rc = enqueue_mul(buf + 2 * n, buf, n, buf + n, n, 0);
                if (!rc) rc = sync_check();
                if (r) t.push_back(wall() - t0);
            }
            if (rc) break;
            std::sort(t.begin(), t.end());
            if (t[t.size() / 2] < best) {
                best = t[t.size() / 2];
                best_tb = tb;
            }
        }
        g.tile_batch = rc ? saved_tb : best_tb;
    }
    (void)hipFree(buf);
    return rc;
}

}  // extern "C"
