// compact.hpp -- stable stream compaction on the device: every index i in
// [0, n) whose predicate holds is handed, in increasing order of i, to an
// emitter together with its output position.  Two passes over tiles of
// COMPACT_TILE indices (a count per tile, an exclusive scan of the counts,
// then a pass that recounts each thread's contiguous run and writes), so
// the output order never depends on scheduling and n may exceed 2^32.
#pragma once

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "devcommon.hpp"

namespace blco {

constexpr int COMPACT_PER_THREAD = 16;
constexpr uint64_t COMPACT_TILE = (uint64_t)WG * COMPACT_PER_THREAD;

// Exclusive prefix of one value per thread over the workgroup; *total gets
// the workgroup's sum.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t s_wave[WAVES_PER_WG];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
    #pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o);
        if (lane >= o) x += y;
    }
    if (lane == WAVE - 1) s_wave[w] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    #pragma unroll
    for (int k = 0; k < WAVES_PER_WG; ++k) {
        if (k < w) before += s_wave[k];
        all += s_wave[k];
    }
    __syncthreads();  // s_wave may be reused by the next call
    *total = all;
    return before + x - v;
}

template <typename Pred>
__global__ __launch_bounds__(WG) void k_compact_count(uint64_t n, Pred pred, uint64_t* __restrict__ tile_cnt)
{
    const uint64_t t = blockIdx.x;
    const uint64_t i0 = t * COMPACT_TILE + (uint64_t)threadIdx.x * COMPACT_PER_THREAD;
    uint32_t c = 0;
    for (int k = 0; k < COMPACT_PER_THREAD; ++k) {
        const uint64_t i = i0 + k;
        if (i < n && pred(i)) ++c;
    }
    uint32_t total;
    (void)block_exclusive_scan(c, &total);
    if (threadIdx.x == 0) tile_cnt[t] = total;
}

template <typename Pred, typename Emit>
__global__ __launch_bounds__(WG) void k_compact_write(uint64_t n, Pred pred, Emit emit,
                                                      const uint64_t* __restrict__ tile_off)
{
    const uint64_t t = blockIdx.x;
    const uint64_t i0 = t * COMPACT_TILE + (uint64_t)threadIdx.x * COMPACT_PER_THREAD;
    uint32_t c = 0;
    for (int k = 0; k < COMPACT_PER_THREAD; ++k) {
        const uint64_t i = i0 + k;
        if (i < n && pred(i)) ++c;
    }
    uint32_t total;
    uint64_t pos = tile_off[t] + block_exclusive_scan(c, &total);
    for (int k = 0; k < COMPACT_PER_THREAD; ++k) {
        const uint64_t i = i0 + k;
        if (i < n && pred(i)) emit(i, pos++);
    }
}

// Scratch of compact(): tile counts and offsets (2 * tiles + 2 u64) and the
// scan's temporary storage; grown on demand, released by release().
struct CompactScratch {
    uint64_t* tiles = nullptr;
    uint64_t tiles_cap = 0;
    void* tmp = nullptr;
    size_t tmp_cap = 0;
    void release()
    {
        if (tiles) (void)hipFree(tiles);
        if (tmp) (void)hipFree(tmp);
        tiles = nullptr;
        tmp = nullptr;
        tiles_cap = 0;
        tmp_cap = 0;
    }
};

// Pass 1 of a compaction: counts and scans; *count (host) gets the number
// of i < n with pred(i).  Synchronous on `st`.
template <typename Pred>
int compact_count(uint64_t n, Pred pred, uint64_t* count, CompactScratch* s, hipStream_t st)
{
    *count = 0;
    if (n == 0) return BLCO_OK;
    const uint64_t tiles = (n + COMPACT_TILE - 1) / COMPACT_TILE;
    if (tiles > 0x7fffffffull) return fail(BLCO_EUNSUPPORTED, "compaction of %llu items", (unsigned long long)n);
    if (s->tiles_cap < tiles) {
        if (s->tiles) (void)hipFree(s->tiles);
        s->tiles = nullptr;
        s->tiles_cap = 0;
        if (hipMalloc(&s->tiles, sizeof(uint64_t) * (2 * tiles + 2)) != hipSuccess)
            return fail(BLCO_ENOMEM, "compaction scratch of %llu tiles", (unsigned long long)tiles);
        s->tiles_cap = tiles;
    }
    uint64_t* cnt = s->tiles;
    uint64_t* off = s->tiles + tiles + 1;
    size_t need = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, cnt, off, tiles + 1, st));
    if (s->tmp_cap < need) {
        if (s->tmp) (void)hipFree(s->tmp);
        s->tmp = nullptr;
        s->tmp_cap = 0;
        if (hipMalloc(&s->tmp, need) != hipSuccess) return fail(BLCO_ENOMEM, "compaction scan scratch");
        s->tmp_cap = need;
    }
    hipLaunchKernelGGL((k_compact_count<Pred>), dim3((uint32_t)tiles), dim3(WG), 0, st, n, pred, cnt);
    HIP_TRY(hipGetLastError());
    // one more zero count so that off[tiles] is the total
    HIP_TRY(hipMemsetAsync(cnt + tiles, 0, sizeof(uint64_t), st));
    size_t have = s->tmp_cap;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(s->tmp, have, cnt, off, tiles + 1, st));
    HIP_TRY(hipMemcpyAsync(count, off + tiles, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return BLCO_OK;
}

// Pass 2: emit(i, position) for every i < n with pred(i), in increasing
// order of i (the predicate must give what it gave in pass 1).
template <typename Pred, typename Emit>
int compact_write(uint64_t n, Pred pred, Emit emit, CompactScratch* s, hipStream_t st)
{
    if (n == 0) return BLCO_OK;
    const uint64_t tiles = (n + COMPACT_TILE - 1) / COMPACT_TILE;
    const uint64_t* off = s->tiles + tiles + 1;
    hipLaunchKernelGGL((k_compact_write<Pred, Emit>), dim3((uint32_t)tiles), dim3(WG), 0, st, n, pred, emit, off);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

// Both passes (the emitter's output must hold n entries, or as many as
// the predicate selects).
template <typename Pred, typename Emit>
int compact(uint64_t n, Pred pred, Emit emit, uint64_t* count, CompactScratch* s, hipStream_t st)
{
    int rc = compact_count(n, pred, count, s, st);
    if (!rc) rc = compact_write(n, pred, emit, s, st);
    return rc;
}

}  // namespace blco
