// features.cpp -- rank- and type-independent features of each mode of a BLCO
// tensor, for the learned kernel selector (the feature list is in
// csrc/hip/selector.hip).  Computed once, from the host coordinates, when a
// device tensor is created.
//
// Slice (output-row) sizes are counted exactly when the mode has at most
// 2^28 rows; otherwise, like fiber sizes, they are estimated from the groups
// whose hashed key falls in a 1/S sample, S = max(1, nnz / 2^21): a group is
// kept or dropped whole, so the size distribution is not biased.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>

#include "internal.hpp"
#include "rng.hpp"

namespace blco {

namespace {

constexpr uint64_t SAMPLE_TARGET = 1ull << 21;

// size buckets by nonzeros, upper bounds exclusive
int bucket8(uint64_t c)
{
    if (c < 2) return 0;
    if (c < 4) return 1;
    if (c < 16) return 2;
    if (c < 64) return 3;
    if (c < 256) return 4;
    if (c < 1024) return 5;
    if (c < 4096) return 6;
    return 7;
}

int bucket4(uint64_t c) { return c < 2 ? 0 : c < 4 ? 1 : c < 16 ? 2 : 3; }

// Sizes of the groups of equal keys (sorted in place).
template <typename F>
void group_sizes(std::vector<uint64_t>& keys, F&& emit)
{
    std::sort(keys.begin(), keys.end());
    for (size_t i = 0; i < keys.size();) {
        size_t j = i + 1;
        while (j < keys.size() && keys[j] == keys[i]) ++j;
        emit((uint64_t)(j - i));
        i = j;
    }
}

// Keys key(b, i) of every nonzero with mix(key) % S == 0, gathered over
// threads (order does not matter: the keys are sorted afterwards).
template <typename K>
std::vector<uint64_t> sample_keys(uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz,
                                  uint64_t S, K&& key)
{
    std::vector<std::vector<uint64_t>> parts(omp_get_max_threads());
    const uint64_t nnz = nblocks ? block_start[nblocks - 1] + block_nnz[nblocks - 1] : 0;
    #pragma omp parallel
    {
        std::vector<uint64_t>& mine = parts[omp_get_thread_num()];
        const uint64_t nt = (uint64_t)omp_get_num_threads(), th = (uint64_t)omp_get_thread_num();
        const uint64_t i0 = (uint64_t)((unsigned __int128)nnz * th / nt);
        const uint64_t i1 = (uint64_t)((unsigned __int128)nnz * (th + 1) / nt);
        uint64_t b = i0 < i1 ? block_of(block_start, nblocks, i0) : 0;
        for (uint64_t i = i0; i < i1; ++i) {
            while (i >= block_start[b] + block_nnz[b]) ++b;
            const uint64_t k = key(b, i);
            if (S == 1 || mix64(k) % S == 0) mine.push_back(k);
        }
    }
    std::vector<uint64_t> keys;
    for (auto& p : parts) keys.insert(keys.end(), p.begin(), p.end());
    return keys;
}

}  // namespace

void compute_static_features(int N, const uint64_t* dims, const uint64_t* masks, const int* pos, uint64_t nnz,
                             const uint64_t* coords, uint64_t nblocks, const uint64_t* block_start,
                             const uint64_t* block_nnz, const uint64_t* block_coords, int stride, float* out)
{
    if (!nnz) return;
    const uint64_t S = std::max<uint64_t>(1, nnz / SAMPLE_TARGET);
    auto index = [&](uint64_t b, uint64_t i, int m) {
        return ((coords[i] >> pos[m]) & masks[m]) | block_coords[b * N + m];
    };
    for (int n = 0; n < N; ++n) {
        float* f = out + (size_t)n * stride;
        double slice[8] = {}, fiber[4] = {}, top = 0.0, kept = 0.0;
        uint64_t smax = 0;
        std::vector<uint64_t> sizes;
        if (dims[n] <= (1ull << 28)) {
            std::vector<uint32_t> cnt(dims[n], 0);
            uint32_t* cp = cnt.data();
            if (nnz < 0xFFFFFFFFull) {
                // no count can reach the saturation value: in parallel, the
                // lowest rows (where skewed modes keep their hot rows) counted
                // per thread and merged, the rest with atomic increments
                const uint64_t local = std::min<uint64_t>(dims[n], 1u << 16);
                #pragma omp parallel
                {
                    // a contiguous range of nonzeros per thread, walking its blocks
                    std::vector<uint32_t> mine(local, 0);
                    const uint64_t nt = (uint64_t)omp_get_num_threads(), th = (uint64_t)omp_get_thread_num();
                    const uint64_t i0 = nnz * th / nt, i1 = nnz * (th + 1) / nt;
                    uint64_t b = i0 < i1 ? block_of(block_start, nblocks, i0) : 0;
                    for (uint64_t i = i0; i < i1; ++i) {
                        while (i >= block_start[b] + block_nnz[b]) ++b;
                        const uint64_t v = index(b, i, n);
                        if (v < local) ++mine[v];
                        else __atomic_fetch_add(&cp[v], 1u, __ATOMIC_RELAXED);
                    }
                    for (uint64_t v = 0; v < local; ++v)
                        if (mine[v]) __atomic_fetch_add(&cp[v], mine[v], __ATOMIC_RELAXED);
                }
            } else {
                for (uint64_t b = 0; b < nblocks; ++b)
                    for (uint64_t i = block_start[b]; i < block_start[b] + block_nnz[b]; ++i) {
                        uint32_t& c = cp[index(b, i, n)];
                        c += c != 0xFFFFFFFFu;
                    }
            }
            for (uint32_t c : cnt)
                if (c) sizes.push_back(c);
        } else {
            std::vector<uint64_t> keys =
                sample_keys(nblocks, block_start, block_nnz, S, [&](uint64_t b, uint64_t i) { return index(b, i, n); });
            group_sizes(keys, [&](uint64_t c) { sizes.push_back(c); });
        }
        for (uint64_t c : sizes) {
            slice[bucket8(c)] += (double)c;
            kept += (double)c;
            smax = std::max(smax, c);
        }
        // the 64 largest slices of the tensor: 64 / S of them in a sample
        const size_t k = std::min<size_t>(std::max<uint64_t>(1, dims[n] <= (1ull << 28) ? 64 : 64 / S), sizes.size());
        std::partial_sort(sizes.begin(), sizes.begin() + k, sizes.end(), std::greater<uint64_t>());
        for (size_t j = 0; j < k; ++j) top += (double)sizes[j];
        if (kept > 0) {
            for (double& x : slice) x /= kept;
            top /= kept;
        }
        // fibers of mode n: all indices but n equal
        std::vector<uint64_t> fkeys = sample_keys(nblocks, block_start, block_nnz, S, [&](uint64_t b, uint64_t i) {
            uint64_t h = 0x243F6A8885A308D3ull;
            for (int m = 0; m < N; ++m)
                if (m != n) h = mix64(h ^ index(b, i, m));
            return h;
        });
        double fk = 0.0;
        group_sizes(fkeys, [&](uint64_t c) {
            fiber[bucket4(c)] += (double)c;
            fk += (double)c;
        });
        if (fk > 0)
            for (double& x : fiber) x /= fk;
        double other = 0.0;
        for (int m = 0; m < N; ++m)
            if (m != n) other += (double)dims[m];
        f[0] = (float)std::log2((double)dims[n]);
        f[1] = (float)std::log2((double)nnz);
        f[2] = (float)std::log2((double)std::max<uint64_t>(1, nblocks));
        f[3] = f[1] - f[2];
        f[5] = (float)N;
        f[7] = (float)std::log2(std::max(1.0, other));
        for (int j = 0; j < 8; ++j) f[8 + j] = (float)slice[j];
        f[16] = (float)std::log2((double)std::max<uint64_t>(1, smax));
        f[17] = (float)top;
        for (int j = 0; j < 4; ++j) f[18 + j] = (float)fiber[j];
        f[22] = f[1] - f[0];
    }
}

}  // namespace blco
