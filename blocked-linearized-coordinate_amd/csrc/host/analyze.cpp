// analyze.cpp -- host-side tables of a device tensor, computed once when it
// is created (in parallel over the nonzeros): the bounds check of every
// decoded index, the hot rows of each mode, and the per-mode row range of
// the non-hot rows of every fine chunk (the row-window kernel's windows).
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace blco {

constexpr uint64_t LOCAL_ROWS = 1u << 16;

// The block holding nonzero i (blocks are contiguous, in order).
uint64_t block_of(const uint64_t* block_start, uint64_t nblocks, uint64_t i)
{
    return (uint64_t)(std::upper_bound(block_start, block_start + nblocks, i) - block_start) - 1;
}

int analyze_tensor(int nmodes, const uint64_t* dims, const uint64_t* masks, const int* pos, uint64_t nnz,
                   const uint64_t* coords, uint64_t nblocks, const uint64_t* block_start,
                   const uint64_t* block_nnz, const uint64_t* block_coords, uint32_t fine_chunk, uint32_t hot_max,
                   TensorTables* t)
{
    for (uint64_t b = 0; b < nblocks; ++b)
        for (uint64_t s0 = 0; s0 < block_nnz[b]; s0 += fine_chunk) {
            t->fbeg.push_back(block_start[b] + s0);
            t->flen.push_back((uint32_t)std::min<uint64_t>(fine_chunk, block_nnz[b] - s0));
            t->fblk.push_back((uint32_t)b);
        }
    const uint64_t nfine = t->fbeg.size();
    t->hot.assign(nmodes, {});
    t->hot_cnt.assign(nmodes, {});
    // pass 1: bounds check and per-mode row counts (counts saturate at
    // 2^32 - 1; below 2^32 nonzeros none can, and atomic increments run in
    // parallel)
    const bool par = nnz < 0xFFFFFFFFull;
    for (int m = 0; m < nmodes; ++m) {
        std::vector<uint32_t> cnt;
        const bool count = dims[m] <= (1ull << 31);
        if (count) cnt.assign(dims[m], 0);
        uint32_t* c = cnt.data();
        const uint64_t mk = masks[m], dm = dims[m];
        const int ps = mk ? pos[m] : 0;
        // the lowest LOCAL_ROWS rows (where skewed modes keep their hot rows)
        // are counted per thread and merged, the rest with atomics
        const uint64_t local = std::min<uint64_t>(dm, LOCAL_ROWS);
        uint64_t worst = 0;
        bool out = false;
        if (par) {
            // each thread a contiguous range of nonzeros, walking the blocks
            // it crosses (one parallel region per mode, whatever the block
            // count)
            #pragma omp parallel reduction(max : worst) reduction(|| : out)
            {
                std::vector<uint32_t> mine(count ? local : 0, 0);
                const uint64_t nt = (uint64_t)omp_get_num_threads(), th = (uint64_t)omp_get_thread_num();
                const uint64_t i0 = nnz * th / nt, i1 = nnz * (th + 1) / nt;
                uint64_t b = i0 < i1 ? block_of(block_start, nblocks, i0) : 0;
                for (uint64_t i = i0; i < i1; ++i) {
                    while (i >= block_start[b] + block_nnz[b]) ++b;
                    const uint64_t v = ((coords[i] >> ps) & mk) | block_coords[b * nmodes + m];
                    if (v >= dm) {
                        out = true;
                        worst = v > worst ? v : worst;
                        continue;
                    }
                    if (!count) continue;
                    if (v < local) ++mine[v];
                    else __atomic_fetch_add(&c[v], 1u, __ATOMIC_RELAXED);
                }
                for (uint64_t v = 0; v < mine.size(); ++v)
                    if (mine[v]) __atomic_fetch_add(&c[v], mine[v], __ATOMIC_RELAXED);
            }
        } else {
            for (uint64_t b = 0; b < nblocks; ++b) {
                const uint64_t bc = block_coords[b * nmodes + m];
                for (uint64_t i = block_start[b]; i < block_start[b] + block_nnz[b]; ++i) {
                    const uint64_t v = ((coords[i] >> ps) & mk) | bc;
                    if (v >= dm) {
                        out = true;
                        worst = v > worst ? v : worst;
                    } else if (count && c[v] != 0xFFFFFFFFu) {
                        ++c[v];
                    }
                }
            }
        }
        if (out)
            return fail(BLCO_EINVAL, "a block decodes mode-%d index %llu >= %llu", m, (unsigned long long)worst,
                        (unsigned long long)dm);
        if (count && nnz) {
            const uint64_t thr = std::max<uint64_t>(2, nnz / 2048);
            std::vector<uint32_t> cand;
            for (uint64_t r = 0; r < dm; ++r)
                if (c[r] >= thr) cand.push_back((uint32_t)r);
            std::sort(cand.begin(), cand.end(),
                      [&](uint32_t x, uint32_t y) { return c[x] != c[y] ? c[x] > c[y] : x < y; });
            if (cand.size() > hot_max) cand.resize(hot_max);
            t->hot[m] = cand;
            for (uint32_t r : cand) t->hot_cnt[m].push_back(c[r]);
        }
    }
    // pass 2: per fine chunk, per mode, the row range of the non-hot rows
    t->flo.assign((uint64_t)nmodes * nfine, ~0ull);
    t->fhi.assign((uint64_t)nmodes * nfine, 0);
    for (int m = 0; m < nmodes; ++m) {
        std::vector<uint8_t> is_hot;
        if (!t->hot[m].empty()) {
            is_hot.assign(dims[m], 0);
            for (uint32_t r : t->hot[m]) is_hot[r] = 1;
        }
        const uint8_t* ih = is_hot.empty() ? nullptr : is_hot.data();
        const uint64_t mk = masks[m];
        const int ps = mk ? pos[m] : 0;
        #pragma omp parallel for schedule(dynamic, 256)
        for (uint64_t f = 0; f < nfine; ++f) {
            const uint64_t bc = block_coords[(uint64_t)t->fblk[f] * nmodes + m];
            uint64_t mn = ~0ull, mx = 0;
            for (uint64_t i = t->fbeg[f]; i < t->fbeg[f] + t->flen[f]; ++i) {
                const uint64_t v = ((coords[i] >> ps) & mk) | bc;
                if (ih && ih[v]) continue;
                mn = v < mn ? v : mn;
                mx = v > mx ? v : mx;
            }
            t->flo[m * nfine + f] = mn;
            t->fhi[m * nfine + f] = mx;
        }
    }
    return BLCO_OK;
}

}  // namespace blco
