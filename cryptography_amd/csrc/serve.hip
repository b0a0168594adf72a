/*
 * serve.hip — continuous-batching multiply service on one GPU.
 *
 * A server (include/ntt_serve.h) runs one scheduler thread.  Submitters on
 * any thread append requests to a queue and wake it.  The scheduler retires
 * finished batches (an event per executor slot), then, while a slot is free,
 * forms the next batch from the head of the queue: requests in arrival order
 * as long as the batch's working memory (residue rows and carry cells of its
 * largest size class, as plan_batch sizes them) fits a free block of the
 * server's HBM arena.  Each batch runs on its slot's stream, so batches
 * admitted while others run overlap on the device.  Requests whose operand
 * is pinned form batches of their own per (operand, transform length): the
 * pinned operand's partial spectrum is made once into the arena and read by
 * the fused pass of every such batch, so only the other operand is
 * transformed.
 *
 * No counterpart in the reference (its multiplies are serialized behind one
 * mutex, ntt/lib/arith/multiply.c:22-27,86-88).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
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

constexpr size_t GRAIN = 256;
size_t round_up(size_t n) { return (n + GRAIN - 1) / GRAIN * GRAIN; }

/* First-fit allocator over one device allocation. */
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;   /* offset -> length, coalesced */
    int init(size_t bytes)
    {
        size = bytes / GRAIN * GRAIN;
        if (hipMalloc((void **)&base, size) != hipSuccess) {
            (void)hipGetLastError();
            base = nullptr;
            return fail(NTT_ERR_NOMEM, "serve arena of %zu bytes could not be allocated", size);
        }
        free_[0] = size;
        return NTT_OK;
    }
    size_t largest() const
    {
        size_t m = 0;
        for (auto &f : free_) m = std::max(m, f.second);
        return m;
    }
    uint8_t *alloc(size_t n)
    {
        n = round_up(std::max<size_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                return base + off;
            }
        return nullptr;
    }
    void release(uint8_t *p, size_t n)
    {
        if (!p) return;
        n = round_up(std::max<size_t>(n, 1));
        size_t off = (size_t)(p - base);
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + n == nx->first) {   /* merge with the next block */
            n += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {           /* and with the previous one */
                pv->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    void destroy()
    {
        if (base) (void)hipFree(base);
        base = nullptr;
        free_.clear();
    }
};

struct Req {
    ntt_mul_desc d;           /* a pinned operand, if any, is b */
    uint64_t ticket = 0;
    int logL = 0;             /* 0: no transform (schoolbook or empty operand) */
    int pin = -1;             /* pinned operand used as b's spectrum */
    int status = NTT_OK;
    bool done = false;
    double t_done = 0;
};

struct Spectrum {
    uint8_t *mem = nullptr;
    size_t bytes = 0;
    hipEvent_t ready = nullptr;   /* recorded after the transform that made it */
};

uint64_t spec_key(int logL) { return ((uint64_t)logL << 40) | plan_signature(logL); }

struct Pinned {
    const uint64_t *ptr = nullptr;
    uint32_t n = 0;
    bool live = false;
    int users = 0;            /* queued or running requests that use it */
    /* by transform length and the plan it was made under (spec_key): a
     * spectrum made before the length's pass plan changed is not read again */
    std::map<uint64_t, Spectrum> spec;
};

struct Slot {
    hipStream_t st = nullptr;
    MulDesc *d_desc = nullptr;
    MulDesc *h_desc = nullptr;    /* pinned staging */
    size_t cap = 0;
    int *d_flag = nullptr;
    int *h_flag = nullptr;        /* pinned: the integrity flag copied back before `done` */
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Req *> reqs;
    uint8_t *ws_mem = nullptr;
    size_t ws_bytes = 0;
};

/* Working memory of a batch, as plan_batch sizes it: the size classes run one
 * after the other on one workspace, so it holds the largest class's residue
 * rows and, after them, the largest class's carry cells. */
struct Need {
    std::map<int, std::pair<uint32_t, uint32_t>> cls;   /* logL -> (count, max n_out) */
    size_t res_bytes = 0, scr_bytes = 64;
    size_t bytes() const { return round_up(res_bytes) + round_up(scr_bytes); }
    void grow(const Req &r, int one_op, size_t &rb, size_t &sb) const
    {
        rb = res_bytes;
        sb = scr_bytes;
        if (r.logL <= 0) return;
        auto it = cls.find(r.logL);
        const uint32_t cnt = (it == cls.end() ? 0 : it->second.first) + 1;
        const uint32_t mx = std::max(it == cls.end() ? 0u : it->second.second,
                                     std::max(r.d.n_out, r.d.na + r.d.nb));
        size_t rw, s;
        class_needs(r.logL, cnt, mx, one_op, rw, s);
        rb = std::max(rb, rw * 8);
        sb = std::max(sb, s);
    }
    size_t with(const Req &r, int one_op) const
    {
        size_t rb, sb;
        grow(r, one_op, rb, sb);
        return round_up(rb) + round_up(sb);
    }
    void add(const Req &r, int one_op)
    {
        grow(r, one_op, res_bytes, scr_bytes);
        if (r.logL <= 0) return;
        auto &c = cls[r.logL];
        c.first++;
        c.second = std::max(c.second, std::max(r.d.n_out, r.d.na + r.d.nb));
    }
};

}  // namespace

struct ntt_server {
    ntt_serve_config cfg{};
    Arena arena;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;                   /* FIFO of queued requests */
    std::unordered_map<uint64_t, Req *> live;    /* until retired by poll / wait */
    std::vector<Pinned> pins;
    uint64_t next_ticket = 1;
    bool stop = false;
    int fatal = NTT_OK;
    double t_open = 0, t_first = -1, t_last = 0;
    ntt_serve_stats st{};
    std::thread worker;

    void complete(Req *r, int status, double t)
    {
        r->status = status;
        r->done = true;
        r->t_done = t;
        if (r->pin >= 0) pins[r->pin].users--;
        st.completed++;
        t_last = std::max(t_last, t);
    }
    bool any_busy() const
    {
        for (auto &s : slots)
            if (s.busy) return true;
        return false;
    }
    void retire();
    bool admit(Slot &s);
    int launch(Slot &s, std::vector<Req *> &batch, size_t need);
    void run();
};

/* Finished batches: their requests complete with the slot's integrity flag. */
void ntt_server::retire()
{
    for (auto &s : slots) {
        if (!s.busy) continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipErrorNotReady) continue;
        int status = NTT_OK;
        if (q != hipSuccess) {
            status = fail(NTT_ERR_HIP, "served batch failed: %s", hipGetErrorString(q));
            fatal = status;
        } else if (*s.h_flag) {
            status = fail(NTT_ERR_INTEGRITY, "CRT integrity check failed in a served batch");
        }
        const double t = wall() - t_open;
        for (Req *r : s.reqs) complete(r, status, t);
        s.reqs.clear();
        arena.release(s.ws_mem, s.ws_bytes);
        s.ws_mem = nullptr;
        s.ws_bytes = 0;
        s.busy = false;
        cv_done.notify_all();
    }
}

/* Form and launch one batch on the free slot s from the head of the queue.
 * Returns whether a batch was launched. */
bool ntt_server::admit(Slot &s)
{
    if (pending.empty()) return false;
    Req *head = pending.front();
    const int pin = head->pin, pin_log = head->logL;
    const bool one_op = pin >= 0;
    /* memory the pinned operand's spectrum still needs (made with this batch) */
    size_t spec_need = 0;
    if (one_op && !pins[pin].spec.count(spec_key(pin_log)))
        spec_need = round_up(NTT_NPRIMES * spectrum_rows_len(pin_log) * 8);
    const size_t room = arena.largest();
    std::vector<Req *> batch;
    Need need;
    const uint32_t max_batch = cfg.max_batch ? cfg.max_batch : 4096;
    for (Req *r : pending) {
        if (batch.size() >= max_batch) break;
        if (one_op ? (r->pin != pin || r->logL != pin_log) : r->pin >= 0) break;   /* keep arrival order */
        /* a spectrum still to make comes out of the same block (conservative) */
        if (need.with(*r, one_op) + spec_need > room) break;
        need.add(*r, one_op);
        batch.push_back(r);
    }
    if (batch.empty()) {
        if (!any_busy()) {
            /* nothing in flight and the head still does not fit: a pinned
             * request runs on the ordinary path; otherwise the arena holds
             * resident spectra nobody needs now, or is too small */
            if (one_op) {
                for (Req *r : pending)
                    if (r->pin == pin && r->logL == pin_log) {
                        pins[pin].users--;
                        r->pin = -1;
                    }
                return true;   /* retry at once */
            }
            for (auto &p : pins)
                if (!p.users)
                    for (auto &sp : p.spec) {
                        arena.release(sp.second.mem, sp.second.bytes);
                        if (sp.second.ready) (void)hipEventDestroy(sp.second.ready);
                    }
            for (auto &p : pins)
                if (!p.users) p.spec.clear();
            if (need.with(*head, 0) > arena.largest()) {
                pending.pop_front();
                complete(head, fail(NTT_ERR_NOMEM, "request does not fit the serve arena"), wall() - t_open);
                cv_done.notify_all();
            }
            return true;
        }
        return false;      /* wait for memory */
    }
    pending.erase(pending.begin(), pending.begin() + (long)batch.size());
    const int rc = launch(s, batch, need.bytes());
    if (rc) {
        const double t = wall() - t_open;
        for (Req *r : batch) complete(r, rc, t);
        cv_done.notify_all();
        if (rc == NTT_ERR_HIP) fatal = rc;
    }
    return true;
}

int ntt_server::launch(Slot &s, std::vector<Req *> &batch, size_t need)
{
    const int pin = batch[0]->pin, logL = batch[0]->logL;
    const size_t n = batch.size();
    if (n > s.cap) {
        if (s.d_desc) (void)hipFree(s.d_desc);
        if (s.h_desc) (void)hipHostFree(s.h_desc);
        s.d_desc = nullptr;
        s.h_desc = nullptr;
        s.cap = 0;
        HIP_TRY(hipMalloc((void **)&s.d_desc, n * sizeof(MulDesc)));
        HIP_TRY(hipHostMalloc((void **)&s.h_desc, n * sizeof(MulDesc), hipHostMallocDefault));
        s.cap = n;
    }
    Spectrum *sp = nullptr;
    if (pin >= 0) {
        const uint64_t key = spec_key(logL);
        auto it = pins[pin].spec.find(key);
        if (it == pins[pin].spec.end()) {
            Spectrum fresh;
            fresh.bytes = NTT_NPRIMES * spectrum_rows_len(logL) * 8;
            fresh.mem = arena.alloc(fresh.bytes);
            if (!fresh.mem) return fail(NTT_ERR_NOMEM, "no arena block for a resident spectrum");
            HIP_TRY(hipEventCreateWithFlags(&fresh.ready, hipEventDisableTiming));
            it = pins[pin].spec.emplace(key, fresh).first;
            /* the pinned operand's forward passes, into the spectrum's rows */
            Workspace wsS;
            wsS.res = (uint64_t *)fresh.mem;
            wsS.res_words = fresh.bytes / 8;
            MulDesc one{};
            one.a = pins[pin].ptr;
            one.na = pins[pin].n;
            int rc;
            {
                std::lock_guard<std::recursive_mutex> lk(g_lock);
                rc = enqueue_ntt_class(wsS, nullptr, one, 1, logL, 0, 1, s.st, 0, 0, pins[pin].n, nullptr, 1);
            }
            wsS.res = nullptr;
            if (rc) return rc;
            HIP_TRY(hipEventRecord(fresh.ready, s.st));
            st.spectra_made++;
        } else {
            HIP_TRY(hipStreamWaitEvent(s.st, it->second.ready, 0));
        }
        sp = &it->second;
        st.spectrum_hits += n;
    }
    s.ws_mem = arena.alloc(need);
    if (!s.ws_mem) return fail(NTT_ERR_NOMEM, "no arena block for a batch");
    s.ws_bytes = need;
    Workspace ws;
    ws.flag = s.d_flag;
    /* residue rows first, then the carry cells (8-byte aligned) */
    ws.res = (uint64_t *)s.ws_mem;
    int rc = NTT_OK;
    HIP_TRY(hipMemsetAsync(s.d_flag, 0, sizeof(int), s.st));
    if (pin < 0) {
        std::vector<ntt_mul_desc> descs(n);
        for (size_t i = 0; i < n; i++) descs[i] = batch[i]->d;
        BatchPlan bp;
        rc = plan_batch(descs.data(), n, bp);
        if (!rc) {
            ws.res_words = bp.res_words;
            ws.scratch = s.ws_mem + round_up(bp.res_words * 8);
            ws.scratch_bytes = bp.scratch;
            if (round_up(bp.res_words * 8) + round_up(bp.scratch) > need)
                rc = fail(NTT_ERR_ARG, "batch plan exceeds its admitted memory (internal)");
        }
        if (!rc) {
            memcpy(s.h_desc, bp.sorted.data(), n * sizeof(MulDesc));
            HIP_TRY(hipMemcpyAsync(s.d_desc, s.h_desc, n * sizeof(MulDesc), hipMemcpyHostToDevice, s.st));
            std::lock_guard<std::recursive_mutex> lk(g_lock);
            rc = enqueue_batch(ws, bp, s.d_desc, s.st);
        }
    } else {
        uint32_t max_nout = 0, max_nab = 0;
        for (size_t i = 0; i < n; i++) {
            memcpy(&s.h_desc[i], &batch[i]->d, sizeof(MulDesc));
            max_nout = std::max(max_nout, std::max(batch[i]->d.n_out, batch[i]->d.na + batch[i]->d.nb));
            max_nab = std::max(max_nab, std::max(batch[i]->d.na, batch[i]->d.nb));
        }
        size_t rw, sb;
        class_needs(logL, (uint32_t)n, max_nout, 1, rw, sb);
        ws.res_words = rw;
        ws.scratch = s.ws_mem + round_up(rw * 8);
        ws.scratch_bytes = sb;
        HIP_TRY(hipMemcpyAsync(s.d_desc, s.h_desc, n * sizeof(MulDesc), hipMemcpyHostToDevice, s.st));
        std::lock_guard<std::recursive_mutex> lk(g_lock);
        rc = enqueue_ntt_class(ws, s.d_desc, MulDesc{}, (uint32_t)n, logL, max_nout, 0, s.st, 0, 0, max_nab,
                               (const uint64_t *)sp->mem, 0);
    }
    ws.res = nullptr;
    ws.scratch = nullptr;
    if (rc) {
        (void)hipStreamSynchronize(s.st);
        arena.release(s.ws_mem, s.ws_bytes);
        s.ws_mem = nullptr;
        return rc;
    }
    HIP_TRY(hipMemcpyAsync(s.h_flag, s.d_flag, sizeof(int), hipMemcpyDeviceToHost, s.st));
    HIP_TRY(hipEventRecord(s.done, s.st));
    s.reqs = batch;
    s.busy = true;
    st.batches++;
    {
        std::lock_guard<std::recursive_mutex> lk(g_lock);
        g.dispatches += n;
    }
    size_t inflight = 0;
    for (auto &x : slots) inflight += x.busy;
    st.max_inflight = std::max(st.max_inflight, inflight);
    return NTT_OK;
}

void ntt_server::run()
{
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        retire();
        bool progress = false;
        for (auto &s : slots) {
            if (s.busy || pending.empty()) continue;
            if (fatal) {       /* the device failed: nothing more runs */
                const double t = wall() - t_open;
                for (Req *r : pending) complete(r, fatal, t);
                pending.clear();
                cv_done.notify_all();
                break;
            }
            if (!admit(s)) break;
            progress = true;
        }
        if (stop && pending.empty() && !any_busy()) break;
        if (progress) continue;
        if (any_busy())
            cv_work.wait_for(lk, std::chrono::microseconds(20));   /* poll the slots' events */
        else
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
    }
}

static int validate(const ntt_mul_desc &d)
{
    MulDesc m;
    memcpy(&m, &d, sizeof m);
    if ((m.na && !m.a) || (m.nb && !m.b) || (m.n_out && !m.c) || (m.nadd && !m.add))
        return fail(NTT_ERR_ARG, "null pointer in a request");
    if (m.n_out < m.nadd || (m.n_out == 0 && m.na && m.nb))
        return fail(NTT_ERR_ARG, "n_out=%u too small for the request", m.n_out);
    if (log_len(m.na, m.nb) > NTT_MAX_LOG_LEN)
        return fail(NTT_ERR_TOO_LARGE, "request of %u x %u limbs needs 2^%d coefficients (max 2^%d)", m.na, m.nb,
                    log_len(m.na, m.nb), NTT_MAX_LOG_LEN);
    return NTT_OK;
}

extern "C" {

int ntt_serve_open(const ntt_serve_config *cfg, ntt_server **out)
{
    if (!out || !cfg) return fail(NTT_ERR_ARG, "null argument");
    *out = nullptr;
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 16 || cfg->hbm_budget < (1u << 20))
        return fail(NTT_ERR_ARG, "bad serve configuration (1..16 slots, at least 1 MiB of budget)");
    ntt_server *s = new ntt_server();
    s->cfg = *cfg;
    int rc = s->arena.init(cfg->hbm_budget);
    s->slots.resize(cfg->max_inflight);
    for (auto &sl : s->slots) {
        if (rc) break;
        if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void **)&sl.d_flag, sizeof(int)) != hipSuccess ||
            hipHostMalloc((void **)&sl.h_flag, sizeof(int), hipHostMallocDefault) != hipSuccess)
            rc = fail(NTT_ERR_HIP, "serve slot set-up failed");
        else
            *sl.h_flag = 0;
    }
    if (rc) {
        s->stop = true;
        ntt_serve_close(s);
        return rc;
    }
    s->t_open = wall();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return NTT_OK;
}

int ntt_serve_submit(ntt_server *s, const ntt_mul_desc *req, uint64_t *ticket)
{
    if (!s || !req || !ticket) return fail(NTT_ERR_ARG, "null argument");
    int rc = validate(*req);
    if (rc) return rc;
    Req *r = new Req();
    r->d = *req;
    const MulDesc &d = *(const MulDesc *)&r->d;
    const bool sb = d.na == 0 || d.nb == 0 || use_schoolbook(d.na, d.nb, d.n_out);
    r->logL = sb ? 0 : log_len(d.na, d.nb);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        delete r;
        return fail(NTT_ERR_ARG, "server is closing");
    }
    if (r->logL && spectrum_cacheable(r->logL)) {
        for (size_t i = 0; i < s->pins.size(); i++) {
            const Pinned &p = s->pins[i];
            if (!p.live) continue;
            if (r->d.b == p.ptr && r->d.nb == p.n) {
                r->pin = (int)i;
            } else if (r->d.a == p.ptr && r->d.na == p.n) {   /* a*b = b*a: the pinned one becomes b */
                std::swap(r->d.a, r->d.b);
                std::swap(r->d.na, r->d.nb);
                r->pin = (int)i;
            }
            if (r->pin >= 0) break;
        }
    }
    Need alone;
    if (alone.with(*r, 0) > s->arena.size) {    /* it could never run, even alone */
        const size_t need = alone.with(*r, 0);
        delete r;
        return fail(NTT_ERR_TOO_LARGE, "request needs %zu bytes of working memory, more than the serve budget",
                    need);
    }
    if (r->pin >= 0) s->pins[r->pin].users++;
    r->ticket = s->next_ticket++;
    s->live[r->ticket] = r;
    s->pending.push_back(r);
    if (s->t_first < 0) s->t_first = wall() - s->t_open;
    *ticket = r->ticket;
    s->cv_work.notify_one();
    return NTT_OK;
}

static int take_done(ntt_server *s, uint64_t ticket, int *status, double *t_done, bool wait)
{
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return -fail(NTT_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Req *r = it->second;
    if (wait) s->cv_done.wait(lk, [&] { return r->done; });
    if (!r->done) return 0;
    if (status) *status = r->status;
    if (t_done) *t_done = r->t_done;
    s->live.erase(it);
    delete r;
    return 1;
}

int ntt_serve_poll(ntt_server *s, uint64_t ticket, int *status, double *t_done)
{
    if (!s) return -fail(NTT_ERR_ARG, "null server");
    return take_done(s, ticket, status, t_done, false);
}

int ntt_serve_wait(ntt_server *s, uint64_t ticket, int *status, double *t_done)
{
    if (!s) return -fail(NTT_ERR_ARG, "null server");
    return take_done(s, ticket, status, t_done, true);
}

int ntt_serve_pin(ntt_server *s, const uint64_t *d_op, uint32_t n_limbs)
{
    if (!s || !d_op || !n_limbs) return fail(NTT_ERR_ARG, "bad operand to pin");
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &p : s->pins)
        if (p.live && p.ptr == d_op && p.n == n_limbs) return NTT_OK;
    Pinned p;
    p.ptr = d_op;
    p.n = n_limbs;
    p.live = true;
    s->pins.push_back(p);
    return NTT_OK;
}

int ntt_serve_unpin(ntt_server *s, const uint64_t *d_op, uint32_t n_limbs)
{
    if (!s) return fail(NTT_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &p : s->pins) {
        if (!p.live || p.ptr != d_op || p.n != n_limbs) continue;
        if (p.users) return fail(NTT_ERR_ARG, "operand is used by %d queued or running requests", p.users);
        for (auto &sp : p.spec) {
            s->arena.release(sp.second.mem, sp.second.bytes);
            if (sp.second.ready) (void)hipEventDestroy(sp.second.ready);
        }
        p.spec.clear();
        p.live = false;
        return NTT_OK;
    }
    return fail(NTT_ERR_ARG, "operand is not pinned");
}

int ntt_serve_get_stats(ntt_server *s, ntt_serve_stats *stats)
{
    if (!s || !stats) return fail(NTT_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->st;
    stats->makespan_s = s->t_first < 0 ? 0 : std::max(0.0, s->t_last - s->t_first);
    return NTT_OK;
}

void ntt_serve_close(ntt_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (auto &sl : s->slots) {
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.d_desc) (void)hipFree(sl.d_desc);
        if (sl.h_desc) (void)hipHostFree(sl.h_desc);
        if (sl.d_flag) (void)hipFree(sl.d_flag);
        if (sl.h_flag) (void)hipHostFree(sl.h_flag);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    for (auto &p : s->pins)
        for (auto &sp : p.spec)
            if (sp.second.ready) (void)hipEventDestroy(sp.second.ready);
    s->arena.destroy();
    for (auto &kv : s->live) delete kv.second;
    delete s;
}

int ntt_serve_replay(const ntt_mul_desc *reqs, const double *arrival, size_t n, size_t hbm_budget, int nslots,
                     double *t_done, ntt_serve_stats *stats)
{
    if (!reqs || !arrival || !t_done || n == 0) return fail(NTT_ERR_ARG, "bad serve arguments");
    ntt_serve_config cfg{};
    cfg.hbm_budget = hbm_budget;
    cfg.max_inflight = nslots;
    ntt_server *s = nullptr;
    int rc = ntt_serve_open(&cfg, &s);
    if (rc) return rc;
    const double t0 = wall();
    const double off = t0 - s->t_open;      /* server clock -> replay clock */
    std::vector<uint64_t> tk(n);
    size_t submitted = 0;
    for (; submitted < n && !rc; submitted++) {
        for (;;) {    /* until the arrival time: short sleeps, then a spin */
            const double left = arrival[submitted] - (wall() - t0);
            if (left <= 0) break;
            if (left > 200e-6) {
                timespec ts{0, (long)((left - 100e-6) * 1e9)};
                nanosleep(&ts, nullptr);
            }
        }
        rc = ntt_serve_submit(s, &reqs[submitted], &tk[submitted]);
    }
    int first_bad = NTT_OK;
    for (size_t i = 0; i < submitted - (rc ? 1 : 0); i++) {
        int status = NTT_OK;
        double td = 0;
        if (ntt_serve_wait(s, tk[i], &status, &td) != 1) status = NTT_ERR_ARG;
        t_done[i] = td - off;
        if (status && !first_bad) first_bad = status;
    }
    if (stats) (void)ntt_serve_get_stats(s, stats);
    ntt_serve_close(s);
    return rc ? rc : first_bad;
}

}  // extern "C"
