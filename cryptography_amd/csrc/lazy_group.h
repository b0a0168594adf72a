/*
 * lazy_group.h — radix-2^RL butterfly groups with lazy reduction, planned at
 * compile time.  Host and device (the host build is checked against exact
 * modular arithmetic in tests/host_stub/modarith_check.cpp).
 */
#ifndef CRYPTOGRAPHY_AMD_LAZY_GROUP_H
#define CRYPTOGRAPHY_AMD_LAZY_GROUP_H

#include "modarith.h"
#include "ntt_kernels.h"

namespace nttc {

/* ---- lazy radix-2^RL groups with compile-time bounds ----------------------
 * A thread's 2^RL values run through RL butterfly stages in registers.  They
 * are kept below b*p for bounds b known at compile time, up to LZ_MAX = 16
 * (16p < 2^64 as p < 2^60):
 *   - forward (DIF) butterfly (x, y) -> (x + y, (x - y + k p) * w), k the
 *     bound of y: the sum's bound is the sum of both bounds, the Shoup
 *     product is below 2p for any input (below 4p in a loose step, which
 *     uses the truncated quotient of shoup_lazy4: one multiply, a carry and
 *     a register move fewer);
 *   - inverse (DIT) butterfly (x, y) -> (x + t, x - t + 2p), t = y*w below
 *     2p for any y: both bounds are x's plus 2 (in a loose step t is below
 *     4p, the difference adds 4p, and both bounds are x's plus 4);
 *   - a butterfly with twiddle 1 keeps (x + y, x - y + k p): both below the
 *     sum of the bounds.
 * Where a butterfly would pass LZ_MAX its larger input is reduced first
 * (reduce_any: below 2p in one step), and at the end every value above the
 * bound its consumer accepts is reduced.  Harvey's butterflies reduce one
 * value (a conditional subtraction) per butterfly and two per twiddle-1
 * butterfly: a radix-8 forward table round reduces 4 values here instead
 * of 12.  The plan of a group is computed by the compiler (lz_plan) and
 * every reduction and multiple of p in the code is a compile-time choice. */
#define LZ_MAX 16

struct LzPlan {
    unsigned char k[3][8];     /* step j, x-side index i: the multiple of p in the difference */
    bool pre[3][8];            /* reduce value i before step j */
    bool post[8];              /* reduce value i after the last step */
    bool loose[3];             /* step j's products use shoup_lazy4 (below 4p) */
    unsigned char out[8];      /* bound of value i at the end */
    int cost;                  /* reductions * 5 - loose products * 4 (issue cycles, roughly) */
    bool ok;
};

/* Stage order: step j is stage S = j (DIF) or S = RL-1-j (DIT), half-size
 * R >> (S+1).  TW1: the compile-time rounds' butterflies with twiddle 1 (x-side
 * index i with i mod half = 0); table rounds have none.  Steps before J0 are
 * skipped (done otherwise by the caller); values enter below BIN*p and leave
 * below BOUT*p. */
__host__ __device__ constexpr LzPlan lz_plan_mask(int RL, bool dit, bool tw1, int bin, int bout, int j0,
                                                  unsigned loosemask)
{
    LzPlan P{};
    int b[8] = {};
    const int R = 1 << RL;
    for (int i = 0; i < R; i++) b[i] = bin;
    P.ok = bin <= LZ_MAX && bout >= 2 && RL >= 1 && RL <= 3;
    for (int j = 0; j < 3; j++) P.loose[j] = j < RL && ((loosemask >> j) & 1);
    for (int j = j0; j < RL; j++) {
        const int S = dit ? RL - 1 - j : j;
        const int half = R >> (S + 1);
        for (int i = 0; i < R; i++) {
            if (i & half) continue;
            const int y = i + half;
            const bool one = tw1 && (i & (half - 1)) == 0;
            if (dit && !one) {
                const int sb = P.loose[j] ? 4 : 2;     /* t = y*w below sb*p */
                if (b[i] + sb > LZ_MAX) {
                    P.pre[j][i] = true;
                    P.cost += 5;
                    b[i] = 2;
                }
                P.k[j][i] = (unsigned char)sb;
                b[i] = b[y] = b[i] + sb;
                if (P.loose[j]) P.cost -= 4;
            } else {
                for (int t = 0; t < 2; t++) {
                    if (b[i] + b[y] > LZ_MAX) {
                        const int r = b[i] >= b[y] ? i : y;
                        P.pre[j][r] = true;
                        P.cost += 5;
                        b[r] = 2;
                    }
                }
                P.k[j][i] = (unsigned char)b[y];
                const int sum = b[i] + b[y];
                b[i] = sum;
                b[y] = one ? sum : P.loose[j] ? 4 : 2;
                if (!one && P.loose[j]) P.cost -= 4;
            }
            P.ok = P.ok && b[i] <= LZ_MAX && b[y] <= LZ_MAX;
        }
    }
    for (int i = 0; i < R; i++) {
        if (b[i] > bout) {
            P.post[i] = true;
            P.cost += 5;
            b[i] = 2;
        }
        P.out[i] = (unsigned char)b[i];
    }
    return P;
}

/* The plan of least cost over the choices of loose steps (those whose
 * products use shoup_lazy4, below 4p instead of 2p).  A loose forward
 * step's products feed later butterflies of the group, which take values up
 * to 16p; where the group's outputs must leave below 2p (LDS and HBM), loose
 * products in the last step would each cost a reduction, so the cost decides.
 * A loose inverse product t enters both outputs, x + t and x - t + 4p, so it
 * raises both bounds by 4 instead of 2.  Of plans of equal cost the one
 * with the lower output bound (the next inverse round's input bound) wins. */
__host__ __device__ constexpr LzPlan lz_plan(int RL, bool dit, bool tw1, int bin, int bout, int j0)
{
    LzPlan best = lz_plan_mask(RL, dit, tw1, bin, bout, j0, 0);
    int bmax = 0;
    for (int i = 0; i < (1 << RL); i++) bmax = best.out[i] > bmax ? best.out[i] : bmax;
    for (unsigned m = 1; m < (1u << RL); m++) {
        if (m & ((1u << j0) - 1)) continue;
        const LzPlan c = lz_plan_mask(RL, dit, tw1, bin, bout, j0, m);
        int cmax = 0;
        for (int i = 0; i < (1 << RL); i++) cmax = c.out[i] > cmax ? c.out[i] : cmax;
        if (c.ok && (c.cost < best.cost || (c.cost == best.cost && cmax < bmax))) {
            best = c;
            bmax = cmax;
        }
    }
    return best;
}

template <int RL, bool DIT, bool TW1, int BIN, int BOUT, int J0>
struct LzRound {
    static constexpr LzPlan plan = lz_plan(RL, DIT, TW1, BIN, BOUT, J0);
    static_assert(plan.ok, "a lazy group's values must stay below 16p");
};

/* Step J, butterfly of x-side index I (then I + 1, .., then step J + 1);
 * J == RL: the final reductions.  TW supplies the twiddle of (stage, I). */
template <int PI, int RL, bool DIT, bool TW1, int BIN, int BOUT, int J0, int J, int I, class TW>
__host__ __device__ __forceinline__ void lz_step(uint64_t *x, const TW &tw)
{
    using RD = LzRound<RL, DIT, TW1, BIN, BOUT, J0>;
    constexpr int R = 1 << RL;
    if constexpr (J < RL) {
        if constexpr (I == R) {
            lz_step<PI, RL, DIT, TW1, BIN, BOUT, J0, J + 1, 0>(x, tw);
        } else {
            constexpr int S = DIT ? RL - 1 - J : J;
            constexpr int half = R >> (S + 1);
            if constexpr ((I & half) == 0) {
                constexpr int Y = I + half;
                if constexpr (RD::plan.pre[J][I]) x[I] = reduce_any<PI>(x[I]);
                if constexpr (RD::plan.pre[J][Y]) x[Y] = reduce_any<PI>(x[Y]);
                constexpr uint64_t KP = (uint64_t)RD::plan.k[J][I] * prime(PI);
                const uint64_t a = x[I], b = x[Y];
                if constexpr (TW1 && (I & (half - 1)) == 0) {
                    x[I] = a + b;
                    x[Y] = a - b + KP;
                } else if constexpr (!DIT) {
                    const TwPair w = tw.template get<S, I>();
                    x[I] = a + b;
                    if constexpr (RD::plan.loose[J]) x[Y] = shoup_lazy4<PI>(a - b + KP, w.w, w.wp);
                    else x[Y] = shoup_lazy<PI>(a - b + KP, w.w, w.wp);
                } else {
                    const TwPair w = tw.template get<S, I>();
                    uint64_t t;
                    if constexpr (RD::plan.loose[J]) t = shoup_lazy4<PI>(b, w.w, w.wp);
                    else t = shoup_lazy<PI>(b, w.w, w.wp);
                    if constexpr (TW::NEG) {     /* w = -(the twiddle): sum and difference trade places */
                        x[I] = a - t + KP;
                        x[Y] = a + t;
                    } else {
                        x[I] = a + t;
                        x[Y] = a - t + KP;
                    }
                }
            }
            lz_step<PI, RL, DIT, TW1, BIN, BOUT, J0, J, I + 1>(x, tw);
        }
    } else if constexpr (I < R) {
        if constexpr (RD::plan.post[I]) x[I] = reduce_any<PI>(x[I]);
        lz_step<PI, RL, DIT, TW1, BIN, BOUT, J0, J, I + 1>(x, tw);
    }
}

/* The largest bound a plan of 2^RL values leaves. */
__host__ __device__ constexpr int lz_out_max(const LzPlan &P, int RL)
{
    int m = 0;
    for (int i = 0; i < (1 << RL); i++) m = P.out[i] > m ? P.out[i] : m;
    return m;
}

/* Inverse rounds pass their values on unreduced (below 16p).  The bound of
 * the values after the inverse rounds of a sub-transform up to top
 * half-size 2^LH (radix 8 from the top, the smallest round with compile-time
 * twiddles), the smallest round's inputs below BIN0*p; LH < 0: BIN0. */
__host__ __device__ constexpr int dit_bound(int LH, int bin0)
{
    if (LH < 0) return bin0;
    const int RL = LH + 1 >= 3 ? 3 : LH + 1;
    const int prev = dit_bound(LH - RL, bin0);
    return lz_out_max(lz_plan(RL, true, LH - (RL - 1) == 0, prev, LZ_MAX, 0), RL);
}

/* All steps of one group: inputs below BIN*p, outputs below BOUT*p. */
template <int PI, int RL, bool DIT, bool TW1, int BIN, int BOUT, int J0 = 0, class TW>
__host__ __device__ __forceinline__ void lz_group(uint64_t *x, const TW &tw)
{
    lz_step<PI, RL, DIT, TW1, BIN, BOUT, J0, J0, 0>(x, tw);
}

/* The smallest stages with compile-time twiddles: a round whose groups are
 * 2^RL consecutive points (group spacing d = 1) uses only w_(2hs)^e with
 * hs <= 2^(RL-1), fixed powers of fixed roots, so they are immediates, and
 * e = 0 (a quarter to all of a stage's butterflies) skips the multiply.
 * Stage S of the round has half-size hs = 2^(RL-1-S), root order 2^(RL-S). */
template <int PI, int RL, bool DIT>
struct ConstTwF {
    static constexpr bool NEG = false;
    template <int S, int I>
    __host__ __device__ __forceinline__ TwPair get() const
    {
        constexpr int half = (1 << RL) >> (S + 1);
        constexpr int e = I & (half - 1);
        constexpr int lo = RL - S;
        using T = ConstTw<PI, lo, DIT ? (1 << lo) - e : e>;
        return TwPair{T::w, T::wp};
    }
};

}  // namespace nttc

#endif
