/*
 * tile_geom.h — tile geometry of the LDS-tiled transform passes (ntt_fast.hip):
 * where residue q of a tile lives in a residue row.  Host and device, so the
 * address identities the kernels rely on are checked on the host
 * (tests/host_stub/tile_addr_check.cpp).
 */
#ifndef CRYPTOGRAPHY_AMD_TILE_GEOM_H
#define CRYPTOGRAPHY_AMD_TILE_GEOM_H

#include "ntt_kernels.h"

#define FT_LOG 12                 /* tile: 4096 residues */
#define FT_SIZE (1u << FT_LOG)
#define FTPB 512                  /* threads per block */
#define FT_PER (FT_SIZE / FTPB)   /* residues per thread */

/* Tile geometry: G blocks x S rows x C columns, LDS index ((g*S + m) << logC) + c. */
struct FGeom {
    uint32_t logT, logC, logG;
};

/* Residue index (before the row pad, res_at) of word q of tile t in a pass of
 * 2^LOGS-point sub-transforms: column c = q mod C, row m, block gg of the
 * tile; its column is col0 + c of T = 2^logT columns.  Fields c, m and gg are
 * bit ranges of q, so for q = q1 + q2 with q1 and q2 on disjoint bits,
 * ft_addr(t, q) = ft_addr(t, q1) + ft_addr(0, q2). */
__host__ __device__ __forceinline__ size_t ft_addr(int LOGS, const FGeom &g, uint32_t t, uint32_t q, uint32_t &m,
                                                   uint32_t &r)
{
    uint32_t c = q & ((1u << g.logC) - 1);
    m = (q >> g.logC) & ((1u << LOGS) - 1);
    uint32_t gg = q >> (g.logC + LOGS);
    uint32_t colTiles = g.logT - g.logC;
    size_t b0 = ((size_t)(t >> colTiles)) << g.logG;
    uint32_t col0 = (t & ((1u << colTiles) - 1)) << g.logC;
    r = col0 + c;
    return ((b0 + gg) << (LOGS + g.logT)) + r + ((size_t)m << g.logT);
}

/* The smallest DIF round of an S = 2^LOGS transform (LOGS >= 1) works on
 * groups of 2^last_rl(LOGS) consecutive rows, 8 >> last_rl of them per thread. */
__host__ __device__ constexpr int last_rl(int LOGS) { return ((LOGS - 1) % 3) + 1; }
template <int LOGS>
__host__ __device__ constexpr int last_rl() { return last_rl(LOGS); }

/* Group gid (thread tid's n-th: gid = tid + n*FTPB) of that round: rows
 * m0 .. m0 + 2^RL - 1 of tile column c, at tile words q0 + i*C. */
template <int LOGS>
__host__ __device__ __forceinline__ uint32_t last_group_q0(uint32_t logC, uint32_t gid, uint32_t &c, uint32_t &m0)
{
    constexpr int RL = last_rl(LOGS);
    c = gid & ((1u << logC) - 1);
    const uint32_t u = gid >> logC;
    m0 = (u & ((1u << (LOGS - RL)) - 1)) << RL;
    return ((((u >> (LOGS - RL)) << LOGS) + m0) << logC) + c;
}

#endif
