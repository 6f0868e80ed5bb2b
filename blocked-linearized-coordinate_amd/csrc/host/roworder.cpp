// roworder.cpp -- the row-ordered copies behind MTTKRP kernel 14.
//
// For an output mode n, the nonzeros of every BLCO block are stably sorted by
// their mode-n index; the BLCO words and values are stored again in that
// order (blocks keep their extent, so block coordinates and the decode are
// unchanged).  Equal rows keep the ALTO order, so the other modes' indices
// stay as local as the BLCO order made them.  The tiled form first groups
// the nonzeros by a tile of another ("split") mode's index, then by block:
// group (tile, block) is a contiguous run of the copy.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "radix.hpp"

namespace blco {

void row_order(int mode, const uint64_t* masks, const int* pos, uint64_t nnz, const uint64_t* coords,
               uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz, uint64_t* perm)
{
    const uint64_t mk = masks[mode];
    const int ps = pos[mode];
    const int row_bits = mk ? 64 - __builtin_clzll(mk) : 0;
    const int blk_bits = nblocks > 1 ? 64 - __builtin_clzll(nblocks - 1) : 0;
    if (row_bits + blk_bits <= 64) {
        // one radix sort on (block, row within the block)
        std::vector<uint64_t> key(nnz);
        #pragma omp parallel for schedule(static)
        for (uint64_t b = 0; b < nblocks; ++b)
            for (uint64_t i = block_start[b]; i < block_start[b] + block_nnz[b]; ++i)
                key[i] = (b << row_bits) | ((coords[i] >> ps) & mk);
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nnz; ++i) perm[i] = i;
        radix_sort(nnz, key.data(), perm, row_bits + blk_bits);
        return;
    }
    // too many key bits: sort each block on its own
    #pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t b = 0; b < nblocks; ++b) {
        uint64_t* p = perm + block_start[b];
        for (uint64_t i = 0; i < block_nnz[b]; ++i) p[i] = block_start[b] + i;
        std::stable_sort(p, p + block_nnz[b], [&](uint64_t x, uint64_t y) {
            return ((coords[x] >> ps) & mk) < ((coords[y] >> ps) & mk);
        });
    }
}

void row_order_tiled(int mode, int split, const uint8_t* tile_of, int ntiles, int N, const uint64_t* masks,
                     const int* pos, const uint64_t* bcoords, uint64_t nnz, const uint64_t* coords,
                     uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz, uint64_t* perm,
                     uint64_t* group_nnz)
{
    const uint64_t mk = masks[mode];
    const int ps = pos[mode];
    const int row_bits = mk ? 64 - __builtin_clzll(mk) : 0;
    const uint64_t ngroups = (uint64_t)ntiles * nblocks;
    const int grp_bits = ngroups > 1 ? 64 - __builtin_clzll(ngroups - 1) : 0;
    auto group = [&](uint64_t b, uint64_t i) -> uint64_t {
        if (ntiles <= 1) return b;
        const uint64_t si = ((coords[i] >> pos[split]) & masks[split]) | bcoords[b * N + split];
        return (uint64_t)tile_of[si] * nblocks + b;
    };
    std::vector<uint64_t> gid(nnz);
    #pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t b = 0; b < nblocks; ++b)
        for (uint64_t i = block_start[b]; i < block_start[b] + block_nnz[b]; ++i) gid[i] = group(b, i);
    for (uint64_t g = 0; g < ngroups; ++g) group_nnz[g] = 0;
    for (uint64_t i = 0; i < nnz; ++i) ++group_nnz[gid[i]];
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) perm[i] = i;
    if (row_bits + grp_bits <= 64) {
        // one radix sort on (group, row within the block)
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nnz; ++i) gid[i] = (gid[i] << row_bits) | ((coords[i] >> ps) & mk);
        radix_sort(nnz, gid.data(), perm, row_bits + grp_bits);
        return;
    }
    std::stable_sort(perm, perm + nnz, [&](uint64_t x, uint64_t y) {
        if (gid[x] != gid[y]) return gid[x] < gid[y];
        return ((coords[x] >> ps) & mk) < ((coords[y] >> ps) & mk);
    });
}

void gather_bytes(const void* src, size_t esz, const uint64_t* perm, uint64_t n, void* dst)
{
    const char* s = (const char*)src;
    char* d = (char*)dst;
    if (esz == 8) {
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i) ((uint64_t*)d)[i] = ((const uint64_t*)s)[perm[i]];
    } else if (esz == 4) {
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i) ((uint32_t*)d)[i] = ((const uint32_t*)s)[perm[i]];
    } else {
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < n; ++i) std::memcpy(d + i * esz, s + perm[i] * esz, esz);
    }
}

}  // namespace blco
