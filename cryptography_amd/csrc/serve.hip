/*
 * serve.hip — continuous-batching multiply service on one GPU.
 *
 * Requests (device-resident operands and result buffers) arrive over time.
 * The scheduler keeps a FIFO of arrived requests and a few executor slots,
 * each with its own HIP stream, workspace, descriptor buffer and integrity
 * flag.  Whenever a slot is free, every queued request whose working memory
 * fits the HBM budget (counting what the busy slots hold) is admitted into
 * one batch and launched on that slot's stream; the batch runs while later
 * arrivals queue up for the next free slot, so several batches are on the
 * device at once.  A request's working memory is its residue rows
 * (2 operands x 3 primes x L x 8 bytes) plus carry scratch; operands and
 * results are the caller's.
 *
 * No counterpart in the reference (its multiplies are serialized behind one
 * mutex, ntt/lib/arith/multiply.c:22-27,86-88).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <vector>

#include "../../include/ntt_serve.h"
#include "engine_internal.h"

using namespace ntte;

namespace {

double wall()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Working memory of one request (bytes). */
size_t req_mem(const ntt_mul_desc &d)
{
    size_t span = std::max<size_t>((size_t)d.na + d.nb, d.n_out);
    if (d.na == 0 || d.nb == 0) return 64;
    if (d.na <= (uint32_t)g.threshold && d.nb <= (uint32_t)g.threshold && (size_t)d.na + d.nb <= SB_MAX_OUT)
        return 64;
    size_t rowL = row_len(log_len(d.na, d.nb), 1);
    return 2 * NTT_NPRIMES * rowL * 8 + (span / CB_LIMBS + 1) * (4 * CW_ROUNDS * 8 + 1);
}

struct Slot {
    hipStream_t st = nullptr;
    Workspace ws;
    MulDesc *d_desc = nullptr;
    MulDesc *h_desc = nullptr;   /* pinned staging */
    size_t desc_cap = 0;
    int *flag = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    size_t mem = 0;
    std::vector<size_t> reqs;
};

struct Server {
    size_t budget = 0, in_use = 0;
    std::vector<Slot> slots;
    size_t max_batch = 4096;
    int init(size_t budget_bytes, int nslots, size_t ws_words_per_slot)
    {
        budget = budget_bytes;
        slots.resize(nslots);
        for (auto &s : slots) {
            HIP_TRY(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
            HIP_TRY(hipMalloc((void **)&s.flag, sizeof(int)));
            HIP_TRY(hipMemsetAsync(s.flag, 0, sizeof(int), s.st));
            s.desc_cap = max_batch;
            HIP_TRY(hipMalloc((void **)&s.d_desc, s.desc_cap * sizeof(MulDesc)));
            HIP_TRY(hipHostMalloc((void **)&s.h_desc, s.desc_cap * sizeof(MulDesc), hipHostMallocDefault));
            s.ws.flag = s.flag;
            int rc = ws_reserve(s.ws, ws_words_per_slot, ws_words_per_slot / 8 + 4096);
            if (rc) return rc;
        }
        HIP_TRY(hipDeviceSynchronize());
        return NTT_OK;
    }
    void release()
    {
        for (auto &s : slots) {
            if (s.st) (void)hipStreamSynchronize(s.st);
            s.ws.release();
            if (s.d_desc) (void)hipFree(s.d_desc);
            if (s.h_desc) (void)hipHostFree(s.h_desc);
            if (s.flag) (void)hipFree(s.flag);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.st) (void)hipStreamDestroy(s.st);
        }
        slots.clear();
    }
};

}  // namespace

extern "C" int ntt_serve_replay(const ntt_mul_desc *reqs, const double *arrival, size_t n, size_t hbm_budget,
                                int nslots, double *t_done, ntt_serve_stats *stats)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (!reqs || !arrival || !t_done || n == 0 || nslots < 1 || nslots > 16)
        return fail(NTT_ERR_ARG, "bad serve arguments");
    /* a slot's workspace is words * 8 bytes: the same bound admits a request
     * here and in the loop below, so every request that passes can run */
    const size_t words = hbm_budget / nslots / 8;
    for (size_t i = 0; i < n; i++)
        if (req_mem(reqs[i]) > words * 8)
            return fail(NTT_ERR_TOO_LARGE, "request %zu needs %zu bytes, more than a slot's share of the budget",
                        i, req_mem(reqs[i]));
    Server sv;
    int rc = sv.init(hbm_budget, nslots, words);
    if (rc) {
        sv.release();
        return rc;
    }
    ntt_serve_stats st{};
    std::deque<size_t> queue;
    size_t next = 0, done = 0, batches = 0, max_inflight = 0;
    const double t0 = wall();
    while (!rc && done < n) {
        double now = wall() - t0;
        while (next < n && arrival[next] <= now) queue.push_back(next++);
        bool progress = false;
        /* retire finished batches */
        for (auto &s : sv.slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = fail(NTT_ERR_HIP, "batch failed: %s", hipGetErrorString(q));
                break;
            }
            int flag = 0;
            if (hipMemcpyAsync(&flag, s.flag, sizeof(int), hipMemcpyDeviceToHost, s.st) != hipSuccess ||
                hipStreamSynchronize(s.st) != hipSuccess) {
                rc = fail(NTT_ERR_HIP, "flag read failed");
                break;
            }
            if (flag) {
                rc = fail(NTT_ERR_INTEGRITY, "CRT integrity check failed in a served batch");
                break;
            }
            double tnow = wall() - t0;
            for (size_t r : s.reqs) t_done[r] = tnow;
            done += s.reqs.size();
            sv.in_use -= s.mem;
            s.reqs.clear();
            s.mem = 0;
            s.busy = false;
            progress = true;
        }
        if (rc) break;
        /* admit the queue head into a free slot */
        int inflight = 0;
        for (auto &s : sv.slots) inflight += s.busy;
        for (auto &s : sv.slots) {
            if (s.busy || queue.empty()) continue;
            std::vector<ntt_mul_desc> batch;
            size_t mem = 0;
            while (!queue.empty() && batch.size() < sv.max_batch) {
                size_t r = queue.front();
                size_t m = req_mem(reqs[r]);
                if (sv.in_use + mem + m > sv.budget || mem + m > words * 8) break;
                batch.push_back(reqs[r]);
                s.reqs.push_back(r);
                mem += m;
                queue.pop_front();
            }
            if (batch.empty()) break;    /* wait for memory */
            BatchPlan bp;
            rc = plan_batch(batch.data(), batch.size(), bp);
            if (!rc && (bp.res_words > s.ws.res_words || bp.scratch > s.ws.scratch_bytes))
                rc = fail(NTT_ERR_NOMEM, "batch exceeds its slot workspace");
            if (rc) break;
            memcpy(s.h_desc, bp.sorted.data(), bp.sorted.size() * sizeof(MulDesc));
            if (hipMemcpyAsync(s.d_desc, s.h_desc, bp.sorted.size() * sizeof(MulDesc), hipMemcpyHostToDevice,
                               s.st) != hipSuccess) {
                rc = fail(NTT_ERR_HIP, "descriptor upload");
                break;
            }
            rc = enqueue_batch(s.ws, bp, s.d_desc, s.st);
            if (!rc && hipEventRecord(s.done, s.st) != hipSuccess) rc = fail(NTT_ERR_HIP, "event record");
            if (rc) break;
            s.busy = true;
            s.mem = mem;
            sv.in_use += mem;
            batches++;
            g.dispatches += batch.size();
            inflight++;
            max_inflight = std::max<size_t>(max_inflight, inflight);
            progress = true;
        }
        if (!progress) {
            /* nothing to launch or retire: wait a little */
            timespec ts{0, 20000};
            nanosleep(&ts, nullptr);
        }
    }
    st.makespan_s = wall() - t0;
    st.batches = batches;
    st.max_inflight = max_inflight;
    sv.release();
    if (stats) *stats = st;
    return rc;
}
