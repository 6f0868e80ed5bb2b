// roworder.cpp -- the row-ordered copies behind MTTKRP kernel 14.
//
// For an output mode n the nonzeros are grouped by (tile of a split mode,
// BLCO block) -- one tile when the copy is not tiled -- and stably sorted by
// their mode-n index within a group; the BLCO words and values are stored
// again in that order.  Every nonzero keeps its block, so block coordinates
// and the decode are unchanged.  Within a row the nonzeros are either left in
// ALTO order (which keeps the other modes' indices local) or ordered by the
// "outer" modes, which makes every fiber (equal row and outer indices)
// contiguous for the fiber kernel.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "internal.hpp"
#include "radix.hpp"

namespace blco {

void row_order(const RowOrderSpec& o, uint64_t nnz, const uint64_t* coords, uint64_t* perm, uint64_t* group_nnz)
{
    const uint64_t mk = o.masks[o.mode];
    const int ps = o.pos[o.mode];
    const int row_bits = mk ? 64 - __builtin_clzll(mk) : 0;
    const int ntiles = o.ntiles > 1 ? o.ntiles : 1;
    const uint64_t nb = o.nblocks, ngroups = (uint64_t)ntiles * nb;
    const int grp_bits = ngroups > 1 ? 64 - __builtin_clzll(ngroups - 1) : 0;
    // group of every nonzero
    std::vector<uint64_t> key(nnz);
    #pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t b = 0; b < nb; ++b) {
        const uint64_t lo = o.block_start[b], hi = lo + o.block_nnz[b];
        if (ntiles == 1) {
            for (uint64_t i = lo; i < hi; ++i) key[i] = b;
        } else {
            const uint64_t sm = o.masks[o.split], top = o.bcoords[b * o.nmodes + o.split];
            const int sp = o.pos[o.split];
            for (uint64_t i = lo; i < hi; ++i)
                key[i] = (uint64_t)o.tile_of[((coords[i] >> sp) & sm) | top] * nb + b;
        }
    }
    for (uint64_t g = 0; g < ngroups; ++g) group_nnz[g] = 0;
    for (uint64_t i = 0; i < nnz; ++i) ++group_nnz[key[i]];
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) perm[i] = i;
    // within a row: by the outer modes' local indices (LSD: this pass first)
    if (o.nouter > 0) {
        std::vector<uint64_t> k2(nnz);
        int bits = 0;
        for (int j = 0; j < o.nouter; ++j) {
            const uint64_t m = o.masks[o.outer[j]];
            bits += m ? 64 - __builtin_clzll(m) : 0;
        }
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nnz; ++i) {
            uint64_t k = 0;
            for (int j = 0; j < o.nouter; ++j) {
                const uint64_t m = o.masks[o.outer[j]];
                const int w = m ? 64 - __builtin_clzll(m) : 0;
                k = (w < 64 ? k << w : 0) | ((coords[i] >> o.pos[o.outer[j]]) & m);
            }
            k2[i] = k;
        }
        radix_sort(nnz, k2.data(), perm, bits);
    }
    if (row_bits + grp_bits <= 64) {
        // one stable radix sort on (group, row within the block)
        std::vector<uint64_t> k1(nnz);
        #pragma omp parallel for schedule(static)
        for (uint64_t j = 0; j < nnz; ++j) {
            const uint64_t i = perm[j];
            k1[j] = (grp_bits ? key[i] << row_bits : 0) | ((coords[i] >> ps) & mk);
        }
        radix_sort(nnz, k1.data(), perm, row_bits + grp_bits);
        return;
    }
    std::stable_sort(perm, perm + nnz, [&](uint64_t x, uint64_t y) {
        if (key[x] != key[y]) return key[x] < key[y];
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
