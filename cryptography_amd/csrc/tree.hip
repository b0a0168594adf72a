/*
 * tree.hip — level-batched binary-splitting product tree for sum 1/k!.
 *
 * Same tree as the reference (ntt/app/compute_e/binary_split.c:74-133): the
 * range [a, b) splits at m = a + (b-a)/2; a node computes B = B_L * B_R and
 * P = P_L * B_R + P_R.  Instead of a depth-first recursion with one
 * serialized multiply per call, every node of one height is evaluated in one
 * batched launch sequence (both products of a node, the addend P_R fused into
 * the carry pass), and the whole tree is captured into one hipGraph.
 *
 * Nodes whose range has at most LEAF_TERMS terms are evaluated directly by a
 * leaf kernel (one thread per node): B = prod k, P by Horner
 * (P = 1; P = P*j + 1 for j = a+1..b-1), which equals the reference's
 * recursion on that range because P and B depend only on [a, b).
 *
 * Every value lives in one device arena at a fixed offset with a fixed
 * capacity taken from an upper bound on its bits (lgamma), so the complete
 * launch sequence is known before any work runs.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "engine_internal.h"

using namespace ntte;

#define LEAF_TERMS 16
#define LEAF_LIMBS 8   /* 16 terms of < 32 bits: at most 512 bits */

struct LeafDesc {
    uint64_t a, b;           /* range [a, b) */
    uint64_t offP, offB;     /* arena offsets (limbs) */
    uint32_t capP, capB;     /* capacities (limbs) */
};

/* x[0..n) *= v; returns the carry limb */
__device__ __forceinline__ uint64_t mul_small(uint64_t *x, int n, uint64_t v)
{
    uint64_t carry = 0;
    for (int i = 0; i < n; i++) {
        uint64_t lo = x[i] * v, hi = __umul64hi(x[i], v);
        uint64_t s = lo + carry;
        carry = hi + (s < lo);
        x[i] = s;
    }
    return carry;
}

__global__ void __launch_bounds__(256) k_leaves(const LeafDesc *leaves, uint32_t n, uint64_t *arena)
{
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const LeafDesc d = leaves[i];
    uint64_t P[LEAF_LIMBS], B[LEAF_LIMBS];
    for (int k = 0; k < LEAF_LIMBS; k++) P[k] = B[k] = 0;
    if (d.b <= d.a) {          /* empty range: P = 0, B = 1 (binary_split.c:84-90) */
        B[0] = 1;
    } else {
        P[0] = 1;
        B[0] = d.a;
        for (uint64_t j = d.a + 1; j < d.b; j++) {
            mul_small(B, LEAF_LIMBS, j);
            mul_small(P, LEAF_LIMBS, j);
            /* + 1 */
            for (int k = 0; k < LEAF_LIMBS; k++) {
                if (++P[k] != 0) break;
            }
        }
    }
    for (uint32_t k = 0; k < d.capP; k++) arena[d.offP + k] = k < LEAF_LIMBS ? P[k] : 0;
    for (uint32_t k = 0; k < d.capB; k++) arena[d.offB + k] = k < LEAF_LIMBS ? B[k] : 0;
}

struct TreeNode {
    long a, b;
    uint64_t offP = 0, offB = 0;
    uint32_t capP = 0, capB = 0;
};

/* Upper bound on the limbs of B = prod_{k=a}^{b-1} k (and of P < 2B). */
/* ln((x-1)!) = lgamma(x): Stirling's series for x >= 16 (the dropped terms
 * are below 1/(1680 x^7) < 1e-11), the library function below. */
static double ln_gamma(double x)
{
    if (x < 16.0) return lgamma(x);
    const double r = 1.0 / x, r2 = r * r;
    return (x - 0.5) * log(x) - x + 0.91893853320467274 + r * (1.0 / 12 - r2 * (1.0 / 360 - r2 * (1.0 / 1260)));
}

/* Limbs that hold B = a (a+1) .. (b-1): log2 B = (ln (b-1)! - ln (a-1)!) / ln 2,
 * with a relative margin for rounding and two spare limbs. */
static uint32_t cap_limbs_B(long a, long b)
{
    if (b <= a) return 1;
    double bits = (ln_gamma((double)b) - ln_gamma((double)a)) / log(2.0);
    return (uint32_t)(bits / 64.0 + 1e-9 * bits) + 2;
}

/* Height of a node with s terms: 0 for a leaf, else one more than its larger
 * child's (the right one, ceil(s/2) terms). */
static int node_height(long s)
{
    int h = 0;
    for (; s > LEAF_TERMS; s -= s / 2) h++;
    return h;
}

struct ntt_split_tree {
    long a = 0, b = 0;
    TreeNode root;
    uint64_t *arena = nullptr;
    size_t arena_words = 0;
    LeafDesc *d_leaves = nullptr;
    uint32_t n_leaves = 0;
    MulDesc *d_desc = nullptr;
    std::vector<BatchPlan> levels;
    std::vector<size_t> level_off;   /* offset of each level's descriptors in d_desc */
    Workspace ws;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t n_products = 0;
    int per_product = 0;
    unsigned long gen = 0;      /* engine generation the tree was planned in */
};

static void tree_free(ntt_split_tree *t)
{
    if (!t) return;
    if (t->exec) (void)hipGraphExecDestroy(t->exec);
    if (t->graph) (void)hipGraphDestroy(t->graph);
    if (t->arena) (void)hipFree(t->arena);
    if (t->d_leaves) (void)hipFree(t->d_leaves);
    if (t->d_desc) (void)hipFree(t->d_desc);
    t->ws.release();
    delete t;
}

/* Enqueue leaves and every level on `st` (also used under graph capture).
 * per_product: launch every product on its own (the unbatched device path,
 * same arena and order) instead of one batch per level. */
static int tree_enqueue(ntt_split_tree *t, hipStream_t st, int per_product)
{
    if (t->n_leaves) {
        hipLaunchKernelGGL(k_leaves, dim3((t->n_leaves + 255) / 256), dim3(256), 0, st, t->d_leaves,
                           t->n_leaves, t->arena);
        HIP_TRY(hipGetLastError());
    }
    for (size_t h = 0; h < t->levels.size(); h++) {
        if (t->levels[h].sorted.empty()) continue;
        if (per_product) {
            for (const MulDesc &d : t->levels[h].sorted) {
                /* one launch sequence per product, bounded by its n_out */
                ntt_mul_desc one{d.a, d.b, d.c, d.add, d.na, d.nb, d.nadd, d.n_out};
                int rc = ntt_mul_batch_device_async(&one, 1);
                if (rc) return rc;
            }
            continue;
        }
        int rc = enqueue_batch(t->ws, t->levels[h], t->d_desc + t->level_off[h], st);
        if (rc) return rc;
    }
    return NTT_OK;
}

extern "C" {

void ntt_split_tree_destroy(ntt_split_tree *t)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (t && g.init) (void)hipStreamSynchronize(g.stream);
    tree_free(t);
}

int ntt_split_tree_create(long a, long b, int use_graph, ntt_split_tree **out)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!out) return fail(NTT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (a < 1 || (b > a && b - 1 > 0xffffffffL)) return fail(NTT_ERR_ARG, "range [%ld, %ld) unsupported", a, b);
    ntt_split_tree *t = new ntt_split_tree();
    t->a = a;
    t->b = b;
    t->per_product = use_graph == 2;
    t->gen = g_generation;
    /* The tree in one array, depth by depth, left to right: node i covers
     * [lo[i], lo[i] + sz[i]); an internal node's children are kid[i] and
     * kid[i] + 1.  Then three passes over all nodes in parallel chunks:
     * (A) heights, capacities, and per chunk its arena words, leaves and
     * products per height; (B) arena offsets and the leaf descriptors;
     * (C) the product descriptors, which read the children's offsets. */
    std::vector<long, uninit_alloc<long>> lo(1, a);
    std::vector<uint32_t, uninit_alloc<uint32_t>> sz(1, (uint32_t)std::max<long>(b - a, 0)), kid(1, 0);
    {
        /* one depth at a time: count the internal nodes per chunk, then each
         * chunk writes its nodes' children from its own offset.  Leaves hold
         * at least LEAF_TERMS / 2 terms (or the whole range), so there are at
         * most 2 (b - a) / (LEAF_TERMS / 2) + 2 nodes. */
        const size_t cap = 4 * (size_t)std::max<long>(b - a, 0) / LEAF_TERMS + 2;
        lo.reserve(cap);
        sz.reserve(cap);
        kid.reserve(cap);
        std::vector<size_t> cnt(host_threads() + 1, 0);
        for (size_t d0 = 0, d1 = 1; d0 < d1;) {
            const size_t n = d1 - d0;
            const unsigned nc = par_chunks(n, [&](size_t i0, size_t i1, unsigned w) {
                size_t c = 0;
                for (size_t i = d0 + i0; i < d0 + i1; i++) c += sz[i] > LEAF_TERMS;
                cnt[w] = c;
            });
            size_t m = 0;
            for (unsigned w = 0; w < nc; w++) {
                const size_t c = cnt[w];
                cnt[w] = m;
                m += c;
            }
            lo.resize(d1 + 2 * m);
            sz.resize(d1 + 2 * m);
            kid.resize(d1 + 2 * m);
            par_chunks(n, [&](size_t i0, size_t i1, unsigned w) {
                size_t at = d1 + 2 * cnt[w];
                for (size_t i = d0 + i0; i < d0 + i1; i++) {
                    const uint32_t s_ = sz[i];
                    if (s_ <= LEAF_TERMS) {
                        kid[i] = 0;
                        continue;
                    }
                    kid[i] = (uint32_t)at;
                    lo[at] = lo[i];                    /* [a, m), m = a + (b-a)/2 */
                    sz[at] = s_ / 2;
                    lo[at + 1] = lo[i] + s_ / 2;
                    sz[at + 1] = s_ - s_ / 2;
                    at += 2;
                }
            });
            d0 = d1;
            d1 += 2 * m;
        }
    }
    const size_t N = lo.size();
    std::vector<uint8_t, uninit_alloc<uint8_t>> height(N);
    std::vector<uint32_t, uninit_alloc<uint32_t>> capP(N), capB(N);
    std::vector<uint64_t, uninit_alloc<uint64_t>> offP(N), offB(N);
    const unsigned NT = host_threads();
    constexpr int HMAX = 40;                       /* heights < 33 for 2^32 terms */
    std::vector<uint64_t> c_words(NT, 0), c_leaves(NT, 0);
    std::vector<uint64_t> c_prod((size_t)NT * (HMAX + 1), 0);
    const unsigned nch = par_chunks(N, [&](size_t i0, size_t i1, unsigned w) {
        uint64_t words = 0, nl = 0;
        uint64_t *np = &c_prod[(size_t)w * (HMAX + 1)];
        for (size_t i = i0; i < i1; i++) {
            const long l = lo[i], r = l + sz[i];
            const int h = node_height(sz[i]);
            height[i] = (uint8_t)h;
            if (h == 0) {
                /* leaf: B < prod 2^bitlen(k) over at most LEAF_TERMS terms,
                 * P < 2B, both within LEAF_LIMBS */
                uint32_t bits = 0;
                for (long k = l; k < r; k++) bits += 64 - __builtin_clzl((unsigned long)k);
                capB[i] = std::max<uint32_t>(std::min<uint32_t>(bits / 64 + 1, LEAF_LIMBS), 1);
                capP[i] = std::max<uint32_t>(std::min<uint32_t>((bits + 1) / 64 + 1, LEAF_LIMBS), 1);
                nl++;
            } else {
                capB[i] = cap_limbs_B(l, r);
                capP[i] = capB[i] + 1;
                np[h] += 1;
            }
            words += capP[i] + capB[i];
        }
        c_words[w] = words;
        c_leaves[w] = nl;
    });
    /* chunk starts: arena words and leaves in node order; products per
     * height, chunk after chunk */
    int H = 0;
    uint64_t off = 0, nleaf = 0;
    std::vector<size_t> per_n(HMAX + 1, 0);
    for (unsigned w = 0; w < nch; w++) {
        const uint64_t ww = c_words[w], lw = c_leaves[w];
        c_words[w] = off;
        off += ww;
        c_leaves[w] = nleaf;
        nleaf += lw;
        for (int h = 1; h <= HMAX; h++) {
            const uint64_t v = c_prod[(size_t)w * (HMAX + 1) + h];
            c_prod[(size_t)w * (HMAX + 1) + h] = per_n[h];
            per_n[h] += v;
            if (v) H = std::max(H, h);
        }
    }
    std::vector<LeafDesc, uninit_alloc<LeafDesc>> leaves(nleaf);
    par_chunks(N, [&](size_t i0, size_t i1, unsigned w) {
        uint64_t o = c_words[w], li = c_leaves[w];
        for (size_t i = i0; i < i1; i++) {
            offP[i] = o;
            o += capP[i];
            offB[i] = o;
            o += capB[i];
            if (height[i] == 0)
                leaves[li++] = {(uint64_t)lo[i], (uint64_t)(lo[i] + sz[i]), offP[i], offB[i], capP[i], capB[i]};
        }
    });
    t->root = TreeNode{a, b, offP[0], offB[0], capP[0], capB[0]};
    t->arena_words = off;
    int rc = NTT_OK;
    if (hipMalloc((void **)&t->arena, off * 8) != hipSuccess) rc = fail(NTT_ERR_NOMEM, "tree arena (%zu limbs)", off);
    t->n_leaves = (uint32_t)leaves.size();
    if (!rc && hipMalloc((void **)&t->d_leaves, leaves.size() * sizeof(LeafDesc)) != hipSuccess)
        rc = fail(NTT_ERR_NOMEM, "leaf descriptors");
    if (!rc) rc = copy_h2d(t->d_leaves, leaves.data(), leaves.size() * sizeof(LeafDesc));
    /* one batch per height: the B = B_L*B_R products of its nodes in node
     * order, then their P = P_L*B_R + P_R products in the same order, so
     * products i and i + n/2 share B_R (transformed once, see plan_batch) */
    std::vector<std::vector<ntt_mul_desc, uninit_alloc<ntt_mul_desc>>> per(H + 1);
    for (int h = 1; h <= H; h++) per[h].resize(2 * per_n[h]);
    uint64_t *A = t->arena;
    par_chunks(N, [&](size_t i0, size_t i1, unsigned w) {
        uint64_t *at = &c_prod[(size_t)w * (HMAX + 1)];
        for (size_t i = i0; i < i1; i++) {
            const int h = height[i];
            if (!h) continue;
            const size_t l = kid[i], r = l + 1, k = at[h]++;
            per[h][k] = ntt_mul_desc{A + offB[l], A + offB[r], A + offB[i], nullptr, capB[l], capB[r], 0, capB[i]};
            per[h][per_n[h] + k] = ntt_mul_desc{A + offP[l], A + offB[r], A + offP[i], A + offP[r], capP[l], capB[r],
                                                capP[r], capP[i]};
        }
    });
    t->levels.resize(H + 1);
    t->level_off.resize(H + 1);
    size_t ndesc = 0, res_words = 1, scratch = 64;
    for (int h = 1; h <= H && !rc; h++) {
        t->level_off[h] = ndesc;
        rc = plan_batch(per[h].data(), per[h].size(), t->levels[h]);
        ndesc += per[h].size();
        res_words = std::max(res_words, t->levels[h].res_words);
        scratch = std::max(scratch, t->levels[h].scratch);
    }
    t->n_products = ndesc;
    if (!rc && ndesc) {
        if (hipMalloc((void **)&t->d_desc, ndesc * sizeof(MulDesc)) != hipSuccess)
            rc = fail(NTT_ERR_NOMEM, "tree descriptors");
        for (int h = 1; h <= H && !rc; h++)
            if (!t->levels[h].sorted.empty())
                rc = copy_h2d(t->d_desc + t->level_off[h], t->levels[h].sorted.data(),
                              t->levels[h].sorted.size() * sizeof(MulDesc));
    }
    if (!rc) rc = ws_reserve(t->ws, res_words, scratch);
    if (!rc && use_graph == 1) {
        for (int h = 1; h <= H; h++)      /* tables first: nothing allocates under capture */
            for (auto &c : t->levels[h].classes) xtab_prepare(c.logL);
        if (hipStreamSynchronize(g.stream) != hipSuccess) rc = fail(NTT_ERR_HIP, "stream sync");
        int saved = g.timing;
        g.timing = 0;
        hipError_t e = rc ? hipSuccess : hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal);
        if (!rc && e != hipSuccess) rc = fail(NTT_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(e));
        if (!rc) {
            rc = tree_enqueue(t, g.stream, 0);
            e = hipStreamEndCapture(g.stream, &t->graph);
            if (!rc && e != hipSuccess) rc = fail(NTT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
        }
        g.timing = saved;
        if (!rc && hipGraphInstantiate(&t->exec, t->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = fail(NTT_ERR_HIP, "hipGraphInstantiate failed");
    }
    if (rc) {
        tree_free(t);
        return rc;
    }
    *out = t;
    return NTT_OK;
}

/* Evaluate the tree (graph replay, or eager launches) and wait. */
int ntt_split_tree_run(ntt_split_tree *t)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!g.init) return fail(NTT_ERR_NOT_INIT, "engine not initialised");
    if (!t) return fail(NTT_ERR_ARG, "null tree");
    if (t->gen != g_generation)
        return fail(NTT_ERR_NOT_INIT, "product tree made before the engine was torn down and re-initialised");
    if (t->exec) {
        g.dispatches += t->n_products;
        HIP_TRY(hipGraphLaunch(t->exec, g.stream));
    } else {
        if (!t->per_product) g.dispatches += t->n_products;
        int rc = tree_enqueue(t, g.stream, t->per_product);
        if (rc) return rc;
    }
    return sync_check();
}

int ntt_split_tree_info(ntt_split_tree *t, size_t *n_products, int *n_levels, size_t *n_leaves, size_t *arena_bytes)
{
    if (!t) return fail(NTT_ERR_ARG, "null tree");
    if (n_products) *n_products = t->n_products;
    if (n_levels) *n_levels = (int)t->levels.size() - 1;
    if (n_leaves) *n_leaves = t->n_leaves;
    if (arena_bytes) *arena_bytes = t->arena_words * 8;
    return NTT_OK;
}

/* Root values on the device (capacities, not trimmed lengths). */
int ntt_split_tree_device_result(ntt_split_tree *t, const uint64_t **dP, size_t *capP, const uint64_t **dB,
                                 size_t *capB)
{
    if (!t) return fail(NTT_ERR_ARG, "null tree");
    const TreeNode &r = t->root;
    if (dP) *dP = t->arena + r.offP;
    if (capP) *capP = r.capP;
    if (dB) *dB = t->arena + r.offB;
    if (capB) *capB = r.capB;
    return NTT_OK;
}

/* Copy the root P and B to host buffers; lengths returned trimmed. */
int ntt_split_tree_result(ntt_split_tree *t, uint64_t *P, size_t capP, size_t *nP, uint64_t *B, size_t capB,
                          size_t *nB)
{
    std::lock_guard<std::recursive_mutex> lk(g_lock);
    if (!t) return fail(NTT_ERR_ARG, "null tree");
    const TreeNode &r = t->root;
    if (capP < r.capP || capB < r.capB)
        return fail(NTT_ERR_ARG, "result buffers too small (need %u and %u limbs)", r.capP, r.capB);
    int rc = copy_d2h(P, t->arena + r.offP, (size_t)r.capP * 8);
    if (!rc) rc = copy_d2h(B, t->arena + r.offB, (size_t)r.capB * 8);
    if (rc) return rc;
    size_t np = r.capP, nb = r.capB;
    while (np && P[np - 1] == 0) np--;
    while (nb && B[nb - 1] == 0) nb--;
    if (nP) *nP = np;
    if (nB) *nB = nb;
    return NTT_OK;
}

int ntt_split_tree_root_caps(ntt_split_tree *t, size_t *capP, size_t *capB)
{
    if (!t) return fail(NTT_ERR_ARG, "null tree");
    if (capP) *capP = t->root.capP;
    if (capB) *capB = t->root.capB;
    return NTT_OK;
}

}  // extern "C"
