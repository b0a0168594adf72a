/*
 * autotune.hip — measures and sets the engine's tunable parameters:
 *   - the schoolbook / NTT crossover (limbs per operand), timed on batches of
 *     equal-size products with each algorithm;
 *   - the pass split of every transform length: every split into the fewest
 *     passes of at most 2^10 points (up to 2^20, also one pass more), each
 *     pass 2^3 points or more, timed on products enqueued back to back;
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
#include <functional>
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

/* Crossover measurements of the last autotune run, per operand size n:
 * seconds per product of a batch of 256 n x n products on the schoolbook
 * kernel and on the NTT, and a lower bound for one Karatsuba level over the
 * faster of the two: a batch of 3 x 256 products of ceil(n/2)+1 limbs (the
 * three half products a0 b0, a1 b1, (a0 + a1)(b0 + b1) with its carry limb),
 * per product of size n, leaving out the additions and the recombination. */
struct CrossSample {
    int n;
    double t_sb, t_ntt, t_kara_lb;
};
std::vector<CrossSample> g_cross;

/* Split measurements of the last autotune run: per length, every pass split
 * tried and its seconds per product. */
struct SplitSamples {
    int logL;
    std::vector<std::pair<std::vector<int>, double>> cand;
};
std::vector<SplitSamples> g_tune;

/* ---- a small JSON reader for the tuning file --------------------------------
 * Values: numbers, strings, objects and arrays; the keys the engine reads
 * are looked up in the top-level object, everything else is skipped. */
struct JsonIn {
    const char *p;
    bool ok = true;
    void ws()
    {
        while (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r') p++;
    }
    bool eat(char c)
    {
        ws();
        if (*p != c) return false;
        p++;
        return true;
    }
    bool str(char *out, size_t cap)
    {
        ws();
        if (*p != '"') return ok = false;
        p++;
        size_t n = 0;
        while (*p && *p != '"') {
            if (n + 1 < cap) out[n++] = *p;
            p++;
        }
        if (*p != '"') return ok = false;
        p++;
        if (cap) out[n] = 0;
        return true;
    }
    bool num(double &v)
    {
        ws();
        char *end = nullptr;
        v = strtod(p, &end);
        if (end == p) return ok = false;
        p = end;
        return true;
    }
    void skip()   /* any value */
    {
        ws();
        if (*p == '{') {
            p++;
            if (eat('}')) return;
            do {
                char k[64];
                if (!str(k, sizeof k) || !eat(':')) { ok = false; return; }
                skip();
            } while (ok && eat(','));
            if (!eat('}')) ok = false;
        } else if (*p == '[') {
            p++;
            if (eat(']')) return;
            do skip();
            while (ok && eat(','));
            if (!eat(']')) ok = false;
        } else if (*p == '"') {
            char k[256];
            str(k, sizeof k);
        } else {
            double v;
            num(v);
        }
    }
};

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
    fprintf(f, "}, \"pass_split\": {");
    first = 1;
    for (int l = 1; l <= NTT_MAX_LOG_LEN; l++)
        if (g.split_n[l]) {
            fprintf(f, "%s\"%d\": [", first ? "" : ", ", l);
            for (int i = 0; i < g.split_n[l]; i++) fprintf(f, "%s%d", i ? ", " : "", g.split[l][i]);
            fprintf(f, "]");
            first = 0;
        }
    fprintf(f, "}");
    if (!g_tune.empty()) {
        /* microseconds per product of every split tried, for the record */
        fprintf(f, ", \"split_us\": {");
        for (size_t i = 0; i < g_tune.size(); i++) {
            fprintf(f, "%s\"%d\": [", i ? ", " : "", g_tune[i].logL);
            for (size_t k = 0; k < g_tune[i].cand.size(); k++) {
                const auto &c = g_tune[i].cand[k];
                fprintf(f, "%s[[", k ? ", " : "");
                for (size_t j = 0; j < c.first.size(); j++) fprintf(f, "%s%d", j ? ", " : "", c.first[j]);
                fprintf(f, "], %.3f]", c.second * 1e6);
            }
            fprintf(f, "]");
        }
        fprintf(f, "}");
    }
    if (!g_cross.empty()) {
        /* microseconds per product, for the record (the loader skips it) */
        fprintf(f, ", \"crossover_us\": [");
        for (size_t i = 0; i < g_cross.size(); i++)
            fprintf(f, "%s[%d, %.4f, %.4f, %.4f]", i ? ", " : "", g_cross[i].n, g_cross[i].t_sb * 1e6,
                    g_cross[i].t_ntt * 1e6, g_cross[i].t_kara_lb * 1e6);
        fprintf(f, "]");
    }
    fprintf(f, "}\n");
    fclose(f);
    return NTT_OK;
}

/* Reader for the file written above: {"threshold": T, "tile_batch": B,
 * "pass_cap": {"<log_len>": cap, ...}, "pass_split": {"<log_len>": [s0, s1,
 * ...], ...}, ...}; other keys are skipped, a missing tile_batch, pass_cap or
 * pass_split leaves the current value, and anything malformed or out of range
 * is an error that changes nothing. */
int ntt_engine_autotune_load(const char *path)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    FILE *f = fopen(path, "r");
    if (!f) return fail(NTT_ERR_ARG, "cannot read %s", path);
    std::vector<char> buf(1 << 16);
    size_t n = fread(buf.data(), 1, buf.size() - 1, f);
    fclose(f);
    buf[n] = 0;
    JsonIn in{buf.data()};
    int thr = -1, tb = g.tile_batch;
    uint8_t caps[NTT_MAX_LOG_LEN + 1], sn[NTT_MAX_LOG_LEN + 1], sp[NTT_MAX_LOG_LEN + 1][8];
    memcpy(caps, g.max_logs, sizeof caps);
    memcpy(sn, g.split_n, sizeof sn);
    memcpy(sp, g.split, sizeof sp);
    bool have_caps = false, have_split = false;
    if (!in.eat('{')) return fail(NTT_ERR_ARG, "%s: not a JSON object", path);
    if (!in.eat('}')) {
        do {
            char key[64];
            if (!in.str(key, sizeof key) || !in.eat(':')) break;
            double v;
            if (!strcmp(key, "threshold")) {
                if (in.num(v)) thr = (int)v;
            } else if (!strcmp(key, "tile_batch")) {
                if (in.num(v)) tb = (int)v;
            } else if (!strcmp(key, "pass_split")) {
                have_split = true;
                memset(sn, 0, sizeof sn);
                if (!in.eat('{')) { in.ok = false; break; }
                if (!in.eat('}')) {
                    do {
                        char lk_[16];
                        if (!in.str(lk_, sizeof lk_) || !in.eat(':') || !in.eat('[')) { in.ok = false; break; }
                        const int l = atoi(lk_);
                        if (l < 1 || l > NTT_MAX_LOG_LEN) return fail(NTT_ERR_ARG, "%s: pass split for 2^%d", path, l);
                        int np = 0, sum = 0, lo = FT_MAX_LOGS;
                        do {
                            if (!in.num(v)) break;
                            const int ls = (int)v;
                            if (np >= 8 || ls < 1 || ls > FT_MAX_LOGS || ls != v)
                                return fail(NTT_ERR_ARG, "%s: pass split for 2^%d out of range", path, l);
                            sp[l][np++] = (uint8_t)ls;
                            sum += ls;
                            lo = std::min(lo, ls);
                        } while (in.ok && in.eat(','));
                        if (!in.ok || !in.eat(']')) { in.ok = false; break; }
                        if (np == 0 || sum != l)
                            return fail(NTT_ERR_ARG, "%s: pass split for 2^%d sums to 2^%d", path, l, sum);
                        if (np > 1 && lo < 3)       /* passes of 2^3..2^10 points, as ntt_engine_set_pass_split */
                            return fail(NTT_ERR_ARG, "%s: pass split for 2^%d has a pass of 2^%d points", path, l, lo);
                        sn[l] = (uint8_t)np;
                    } while (in.ok && in.eat(','));
                    if (in.ok && !in.eat('}')) in.ok = false;
                }
            } else if (!strcmp(key, "pass_cap")) {
                have_caps = true;
                memset(caps, 0, sizeof caps);
                if (!in.eat('{')) { in.ok = false; break; }
                if (!in.eat('}')) {
                    do {
                        char lk_[16];
                        if (!in.str(lk_, sizeof lk_) || !in.eat(':') || !in.num(v)) { in.ok = false; break; }
                        const int l = atoi(lk_), cap = (int)v;
                        if (l < 1 || l > NTT_MAX_LOG_LEN || cap < 4 || cap > FT_MAX_LOGS)
                            return fail(NTT_ERR_ARG, "%s: pass cap %d for 2^%d out of range", path, cap, l);
                        caps[l] = (uint8_t)cap;
                    } while (in.ok && in.eat(','));
                    if (in.ok && !in.eat('}')) in.ok = false;
                }
            } else {
                in.skip();
            }
        } while (in.ok && in.eat(','));
        if (in.ok && !in.eat('}')) in.ok = false;
    }
    if (!in.ok) return fail(NTT_ERR_ARG, "%s: malformed tuning file", path);
    if (thr < 0) return fail(NTT_ERR_ARG, "%s: no threshold", path);
    if (thr > SB_MAX_OUT / 2) return fail(NTT_ERR_ARG, "%s: threshold %d out of range", path, thr);
    if (tb < 1 || tb > 64) return fail(NTT_ERR_ARG, "%s: tile_batch %d out of range", path, tb);
    g.threshold = thr;
    g.tile_batch = tb;
    memcpy(g.max_logs, caps, sizeof caps);
    /* the latest setting decides a length's plan: new caps clear the splits
     * of the lengths they list, then the file's splits apply */
    if (have_caps)
        for (int l = 1; l <= NTT_MAX_LOG_LEN; l++)
            if (caps[l]) sn[l] = have_split ? sn[l] : 0;
    memcpy(g.split_n, sn, sizeof sn);
    memcpy(g.split, sp, sizeof sp);
    return NTT_OK;
}

/* Measure and set the crossover, the pass splits for transform lengths
 * 2^min_log .. 2^max_log and (max_log >= 24) the tile batch.  Takes some
 * seconds. */
int ntt_engine_autotune(int min_log, int max_log)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (min_log < 11) min_log = 11;
    if (max_log > 28) max_log = 28;
    /* operand buffer: big enough for the largest single product and the batches */
    size_t words = std::max((size_t)1 << (max_log - 1), (size_t)768 * 4 * 128);
    uint64_t *buf = nullptr;
    if (hipMalloc((void **)&buf, words * 8) != hipSuccess) return fail(NTT_ERR_NOMEM, "autotune buffer");
    hipLaunchKernelGGL(k_fill_pattern, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, g.stream, buf, words,
                       (uint64_t)12345);
    int rc = NTT_OK;
    /* 1. schoolbook / NTT crossover */
    const int saved = g.threshold;
    int best_thr = 0;
    g_cross.clear();
    for (int n : {8, 16, 24, 32, 48, 64, 96, 128}) {
        double t_sb = 0, t_ntt = 0, t_h_sb = 0, t_h_ntt = 0;
        const int h = (n + 1) / 2 + 1;
        g.threshold = n;
        rc = time_batch(buf, n, 256, 3, &t_sb);
        if (!rc) rc = time_batch(buf, h, 768, 3, &t_h_sb);
        if (rc) break;
        g.threshold = 0;
        rc = time_batch(buf, n, 256, 3, &t_ntt);
        if (!rc) rc = time_batch(buf, h, 768, 3, &t_h_ntt);
        if (rc) break;
        if (t_sb <= t_ntt) best_thr = n;
        /* per product of size n: three half products each */
        g_cross.push_back({n, t_sb, t_ntt, 3 * std::min(t_h_sb, t_h_ntt)});
    }
    g.threshold = rc ? saved : best_thr;
    /* 2. pass splits: every split of length 2^l into the fewest passes of at
     * most 2^10 points (and, up to 2^20, one pass more), each of 2^3 or more,
     * timed on products of L/8 x L/8 limbs enqueued back to back */
    g_tune.clear();
    for (int l = min_log; l <= max_log && !rc; l++) {
        size_t n = (size_t)1 << (l - 3);
        if (4 * n > words) break;
        const int np0 = (l + FT_MAX_LOGS - 1) / FT_MAX_LOGS;
        std::vector<std::vector<int>> cands;
        for (int np = np0; np <= (l <= 20 ? np0 + 1 : np0); np++) {
            std::vector<int> v(np, 3);
            /* compositions of l into np parts in [3, 10], in lexicographic order */
            std::function<void(int, int)> rec = [&](int i, int left) {
                if (i == np - 1) {
                    if (left >= 3 && left <= FT_MAX_LOGS) {
                        v[i] = left;
                        cands.push_back(v);
                    }
                    return;
                }
                for (int x = 3; x <= FT_MAX_LOGS && x <= left - 3 * (np - 1 - i); x++) {
                    v[i] = x;
                    rec(i + 1, left - x);
                }
            };
            rec(0, l);
        }
        if (cands.size() < 2) continue;
        const int k = (int)std::min<size_t>(64, std::max<size_t>(1, (size_t)1 << std::max(0, 24 - l)));
        const uint8_t saved_n = g.split_n[l];
        uint8_t saved_s[8];
        memcpy(saved_s, g.split[l], 8);
        double best = 1e30;
        int best_i = -1;
        SplitSamples rec_l{l, {}};
        for (size_t ci = 0; ci < cands.size() && !rc; ci++) {
            g.split_n[l] = (uint8_t)cands[ci].size();
            for (size_t j = 0; j < cands[ci].size(); j++) g.split[l][j] = (uint8_t)cands[ci][j];
            std::vector<double> t;
            for (int r = 0; r < 4 && !rc; r++) {
                double t0 = wall();
                for (int q = 0; q < k && !rc; q++) rc = enqueue_mul(buf + 2 * n, buf, n, buf + n, n, 0);
                if (!rc) rc = sync_check();
                if (r) t.push_back((wall() - t0) / k);
            }
            if (rc) break;
            std::sort(t.begin(), t.end());
            rec_l.cand.push_back({cands[ci], t[t.size() / 2]});
            if (t[t.size() / 2] < best) {
                best = t[t.size() / 2];
                best_i = (int)ci;
            }
        }
        if (rc || best_i < 0) {
            g.split_n[l] = saved_n;
            memcpy(g.split[l], saved_s, 8);
            break;
        }
        g.split_n[l] = (uint8_t)cands[best_i].size();
        for (size_t j = 0; j < cands[best_i].size(); j++) g.split[l][j] = (uint8_t)cands[best_i][j];
        g_tune.push_back(rec_l);
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
