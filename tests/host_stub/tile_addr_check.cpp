/*
 * Host check of the address identities the tile passes rely on
 * (cryptography_amd/csrc/tile_geom.h, ntt_fast.hip):
 *  - tile_dma: word q0 + 2*lane of a 1 KB chunk (q0 a multiple of 128) lies
 *    at res_at(ft_addr(t, q0) + ft_addr(0, 2*lane)), the same index as
 *    res_at(ft_addr(t, q0 + 2*lane));
 *  - tile_span: the words a thread stores lie at base + k*step (for the row
 *    pads the engine pairs with each geometry);
 *  - last_group: the smallest DIF round's groups cover the tile once, rows
 *    m0 .. m0 + 2^RL - 1 of a column, stored at gb + i*gs, the next row
 *    2^logT residues on.
 * Every pass geometry the launcher makes (S = 2^3..2^10 points, T = 1..2^27
 * columns, C = min(T, 4096 / S), G = 4096 / (S C)), tiles at both ends and spread between, and the row pads
 * in use.  Prints "bad N" and exits non-zero on any mismatch.
 */
#include "../../cryptography_amd/csrc/tile_geom.h"

#include <stdio.h>

static long bad = 0, checks = 0;

static void check_tile(int LOGS, const FGeom &g, uint32_t t, uint32_t pad)
{
    const int CHUNKS = FT_SIZE / 128 / (FTPB / 64);
    uint32_t m, r;
    for (uint32_t w = 0; w < FTPB / 64; w++)
        for (int j = 0; j < CHUNKS; j++)
            for (uint32_t lane = 0; lane < 64; lane++) {
                const uint32_t q0 = (w * CHUNKS + j) * 128;
                const size_t whole = res_at(ft_addr(LOGS, g, t, q0 + 2 * lane, m, r), pad);
                const size_t split = res_at(ft_addr(LOGS, g, t, q0, m, r) + ft_addr(LOGS, g, 0, 2 * lane, m, r), pad);
                checks++;
                if (whole != split && bad++ < 8)
                    printf("dma: LOGS %d logC %u logT %u t %u q %u: %zu != %zu\n", LOGS, g.logC, g.logT, t,
                           q0 + 2 * lane, whole, split);
            }
    /* a thread's words tid + k*FTPB at base + k*step (tile_span), for the
     * pads the engine pairs with this geometry (pad_log_for: a pad after every
     * row of the first pass, T = 2^pad_log; the later passes work in blocks of
     * 2^(LOGS + logT) residues, at most one pad interval) */
    if (!(pad == NTT_PAD_NONE || pad == g.logT || pad >= LOGS + g.logT)) return;
    for (uint32_t tid = 0; tid < FTPB; tid += 37) {
        const size_t base = res_at(ft_addr(LOGS, g, t, tid, m, r), pad);
        const size_t step = res_at(ft_addr(LOGS, g, t, tid + FTPB, m, r), pad) - base;
        for (uint32_t k = 0; k < FT_PER; k++) {
            checks++;
            if (res_at(ft_addr(LOGS, g, t, tid + k * FTPB, m, r), pad) != base + k * step && bad++ < 8)
                printf("span: LOGS %d logC %u logT %u t %u tid %u k %u\n", LOGS, g.logC, g.logT, t, tid, k);
        }
    }
}

/* The smallest DIF round's groups (last_group, dif_last_store and the table
 * pass's register stores): the groups of all threads cover every tile word
 * once, and group rows m0 + i lie at gb + i*gs. */
template <int LOGS>
static void check_groups(const FGeom &g, uint32_t t, uint32_t pad)
{
    constexpr int RL = last_rl<LOGS>(), R = 1 << RL, NG = 8 >> RL;
    static unsigned char seen[FT_SIZE];
    for (uint32_t q = 0; q < FT_SIZE; q++) seen[q] = 0;
    const bool span = pad == NTT_PAD_NONE || pad == g.logT || pad >= LOGS + g.logT;
    uint32_t m, r;
    for (uint32_t tid = 0; tid < FTPB; tid++)
        for (int n = 0; n < NG; n++) {
            uint32_t c, m0;
            const uint32_t q0 = last_group_q0<LOGS>(g.logC, tid + (uint32_t)n * FTPB, c, m0);
            const size_t gb = res_at(ft_addr(LOGS, g, t, q0, m, r), pad);
            const size_t a0 = ft_addr(LOGS, g, t, q0, m, r);
            const size_t gs = res_at(a0 + ((size_t)1 << g.logT), pad) - gb;
            /* the next row of the group is 2^logT residues on (last_group's stride) */
            checks++;
            if (ft_addr(LOGS, g, t, q0 + (1u << g.logC), m, r) != a0 + ((size_t)1 << g.logT) && bad++ < 8)
                printf("group stride: LOGS %d logC %u logT %u tid %u n %d\n", LOGS, g.logC, g.logT, tid, n);
            for (int i = 0; i < R; i++) {
                const uint32_t q = q0 + ((uint32_t)i << g.logC);
                checks++;
                if ((q >= FT_SIZE || seen[q]++ || (ft_addr(LOGS, g, t, q, m, r), m != m0 + (uint32_t)i)) && bad++ < 8)
                    printf("groups: LOGS %d logC %u tid %u n %d i %d q %u\n", LOGS, g.logC, tid, n, i, q);
                if (span && q < FT_SIZE && res_at(ft_addr(LOGS, g, t, q, m, r), pad) != gb + i * gs && bad++ < 8)
                    printf("group span: LOGS %d logC %u logT %u t %u tid %u i %d\n", LOGS, g.logC, g.logT, t, tid, i);
            }
        }
}

static void check_groups_any(int LOGS, const FGeom &g, uint32_t t, uint32_t pad)
{
    switch (LOGS) {
    case 3: check_groups<3>(g, t, pad); break;
    case 4: check_groups<4>(g, t, pad); break;
    case 5: check_groups<5>(g, t, pad); break;
    case 6: check_groups<6>(g, t, pad); break;
    case 7: check_groups<7>(g, t, pad); break;
    case 8: check_groups<8>(g, t, pad); break;
    case 9: check_groups<9>(g, t, pad); break;
    case 10: check_groups<10>(g, t, pad); break;
    }
}

int main()
{
    const uint32_t pads[] = {12, 14, 17, 18, 20, 27, NTT_PAD_NONE};
    for (int LOGS = 3; LOGS <= 10; LOGS++)
        for (uint32_t logT = 0; logT <= 27; logT++) {
            /* the launcher's geometry (ntt_engine.hip): C = min(T, 4096 / S), G = 4096 / (S C) */
            const uint32_t room = FT_LOG - LOGS, logC = logT < room ? logT : room;
            const FGeom g{logT, logC, room - logC};
            for (uint32_t pad : pads)
                for (uint32_t s = 0; s < 24; s++) {
                    uint32_t t = s < 12 ? s : (uint32_t)(((uint64_t)s * 2654435761u) % (1u << 20));
                    if (s == 23) t = (1u << 20) - 1;
                    check_tile(LOGS, g, t, pad);
                    if (s % 6 == 0 || s == 23) check_groups_any(LOGS, g, t, pad);
                }
        }
    printf("checks %ld\nbad %ld\n", checks, bad);
    return bad != 0;
}
