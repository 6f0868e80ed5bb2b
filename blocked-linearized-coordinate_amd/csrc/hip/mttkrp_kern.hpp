// mttkrp_kern.hpp -- the MTTKRP kernel templates (lvl1, lvl3, row window,
// atomic) and their per-shape selection; each precision instantiates them in
// its own translation unit (mttkrp_f64.hip, mttkrp_f32.hip, mttkrp_bf16.hip).
#pragma once

#include "devcommon.hpp"
#include "mttkrp.hpp"

namespace blco {

// ------------------------------------------------------------ row loads

template <typename FT, int C>
__device__ __forceinline__ void load_row(const FT* __restrict__ p,
                                         typename Acc<FT>::type (&x)[C])
{
    using AT = typename Acc<FT>::type;
    if constexpr (C * sizeof(FT) == 32) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
        uint4 u[2] = {q[0], q[1]};
        const FT* e = reinterpret_cast<const FT*>(u);
        #pragma unroll
        for (int c = 0; c < C; ++c) x[c] = to_acc(e[c]);
    } else if constexpr (C * sizeof(FT) == 16) {
        uint4 u = *reinterpret_cast<const uint4*>(p);
        const FT* e = reinterpret_cast<const FT*>(&u);
        #pragma unroll
        for (int c = 0; c < C; ++c) x[c] = to_acc(e[c]);
    } else if constexpr (C * sizeof(FT) == 8) {
        uint2 u = *reinterpret_cast<const uint2*>(p);
        const FT* e = reinterpret_cast<const FT*>(&u);
        #pragma unroll
        for (int c = 0; c < C; ++c) x[c] = to_acc(e[c]);
    } else {
        #pragma unroll
        for (int c = 0; c < C; ++c) x[c] = (AT)to_acc(p[c]);
    }
}

// A buffer resource over a factor matrix of fewer than 2^32 bytes: a row
// gather then takes a 32-bit byte offset in one VGPR beside the uniform
// resource, instead of a 64-bit address.  (Third descriptor word: the raw
// 32-bit data format of the gfx9 family; unbounded range.)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t factor_rsrc(const void* p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, -1, 0x00020000);
}

template <typename FT, int C>
__device__ __forceinline__ void load_row_buf(__amdgpu_buffer_rsrc_t rs, uint32_t byte_off,
                                             typename Acc<FT>::type (&x)[C])
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    FT e[C];
    if constexpr (C * sizeof(FT) == 32) {
        const u32x4 u0 = __builtin_amdgcn_raw_buffer_load_b128(rs, byte_off, 0, 0);
        const u32x4 u1 = __builtin_amdgcn_raw_buffer_load_b128(rs, byte_off + 16, 0, 0);
        __builtin_memcpy(e, &u0, 16);
        __builtin_memcpy(reinterpret_cast<char*>(e) + 16, &u1, 16);
    } else if constexpr (C * sizeof(FT) == 16) {
        const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(rs, byte_off, 0, 0);
        __builtin_memcpy(e, &u, 16);
    } else if constexpr (C * sizeof(FT) == 8) {
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(rs, byte_off, 0, 0);
        __builtin_memcpy(e, &u, 8);
    } else if constexpr (C * sizeof(FT) == 4) {
        const unsigned u = __builtin_amdgcn_raw_buffer_load_b32(rs, byte_off, 0, 0);
        __builtin_memcpy(e, &u, 4);
    } else {
        static_assert(C * sizeof(FT) == 2, "row piece of 2 to 32 bytes");
        const unsigned short u = __builtin_amdgcn_raw_buffer_load_b16(rs, byte_off, 0, 0);
        __builtin_memcpy(e, &u, 2);
    }
    #pragma unroll
    for (int c = 0; c < C; ++c) x[c] = to_acc(e[c]);
}

// Tensor coordinates and values are read once per MTTKRP: non-temporal
// loads keep them from evicting the factor rows that the gathers reuse.
template <typename T>
__device__ __forceinline__ T stream_load(const T* p)
{
    return __builtin_nontemporal_load(p);
}

template <typename AT>
__device__ __forceinline__ void atomic_add_global(AT* p, AT v)
{
    unsafeAtomicAdd(p, v);
}

template <typename AT>
__device__ __forceinline__ void atomic_add_lds(AT* p, AT v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ascending bitonic sort of one 32-bit key per lane across the wave.
__device__ __forceinline__ uint32_t wave_sort64(uint32_t key, int lane)
{
    #pragma unroll
    for (int k = 2; k <= WAVE; k <<= 1) {
        #pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint32_t other = (uint32_t)__shfl_xor((int)key, j);
            bool up = (lane & k) == 0;
            bool lower = (lane & j) == 0;
            uint32_t mn = key < other ? key : other;
            uint32_t mx = key < other ? other : key;
            key = (lower == up) ? mn : mx;
        }
    }
    return key;
}

// Loads, decodes and sorts one tile; leaves the sorted entries in the
// wave's LDS strip (s_idx[m][skewed position], s_val[...]).
template <int N, int L, typename VT, typename AT>
__device__ __forceinline__ void stage_tile(const MttkrpArgs& a, const uint64_t (&bc)[N],
                                           uint64_t e0, uint32_t nvalid, int lane,
                                           uint64_t (*s_idx)[WAVE + WAVE / L],
                                           AT* s_val)
{
    const bool valid = (uint32_t)lane < nvalid;
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    uint64_t coord = valid ? stream_load(a.coords + (e0 + lane - a.coord_base)) : 0;
    AT v = valid ? (AT)stream_load(vals + (e0 + lane - a.coord_base)) : (AT)0;
    uint64_t idx[N];
    uint64_t orow = 0;
    #pragma unroll
    for (int m = 0; m < N; ++m) {
        idx[m] = ((coord >> a.pos[m]) & a.mask[m]) | bc[m];
        if (m == a.mode) orow = idx[m];
    }
    uint64_t base = __shfl(orow, 0);
    uint64_t diff = orow - base + (1ull << 25);
    bool fits = diff < (1ull << 26);
    uint32_t key;
    if (__ballot(valid && !fits) == 0) {
        key = valid ? (((uint32_t)diff << 6) | (uint32_t)lane) : 0xFFFFFFFFu;
        key = wave_sort64(key, lane);
    } else {
        key = (uint32_t)lane;  // rows too far apart: keep tile order
    }
    const int src = (int)(key & 63u);
    const int q = lane + lane / L;
    #pragma unroll
    for (int m = 0; m < N; ++m) s_idx[m][q] = __shfl(idx[m], src);
    s_val[q] = __shfl(v, src);
    wave_lds_sync();
}

// Walks the sorted tile: slot `slot` covers positions [slot*L, slot*L + L).
// FLUSH(row, acc) is called once per segment.
template <int N, typename FT, int C, int L, bool EXACT, typename Flush>
__device__ __forceinline__ void segment_tile(const MttkrpArgs& a, uint32_t nvalid, int slot,
                                             int col, uint64_t (*s_idx)[WAVE + WAVE / L],
                                             const typename Acc<FT>::type* s_val, Flush flush)
{
    using AT = typename Acc<FT>::type;
    const int R = a.rank, mode = a.mode;
    int kmax = (int)nvalid - slot * L;
    kmax = kmax < 0 ? 0 : (kmax > L ? L : kmax);
    const int qb = slot * (L + 1);
    AT acc[C];
    #pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = 0;
    uint64_t cur = 0;
    bool have = false;
    for (int k = 0; k < kmax; ++k) {
        const int qq = qb + k;
        const uint64_t r = s_idx[mode][qq];
        if (have && r != cur) {
            flush(cur, acc);
            #pragma unroll
            for (int c = 0; c < C; ++c) acc[c] = 0;
        }
        cur = r;
        have = true;
        AT p[C];
        const AT vv = s_val[qq];
        #pragma unroll
        for (int c = 0; c < C; ++c) p[c] = vv;
        #pragma unroll
        for (int m = 0; m < N; ++m) {
            if (m == mode) continue;
            const FT* f = reinterpret_cast<const FT*>(a.f[m]) + s_idx[m][qq] * (uint64_t)R + col;
            AT x[C];
            if (EXACT || col + C <= R) {
                load_row<FT, C>(f, x);
            } else {
                #pragma unroll
                for (int c = 0; c < C; ++c) x[c] = (col + c < R) ? (AT)to_acc(f[c]) : (AT)0;
            }
            #pragma unroll
            for (int c = 0; c < C; ++c) p[c] *= x[c];
        }
        #pragma unroll
        for (int c = 0; c < C; ++c) acc[c] += p[c];
    }
    if (have) flush(cur, acc);
}

// --------------------------------------------------------------- lvl1

template <int N, typename FT, typename VT, int C, int L, bool EXACT>
__global__ __launch_bounds__(WG) void k_mttkrp_lvl1(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    __shared__ uint64_t s_idx[WAVES_PER_WG][N][WAVE + WAVE / L];
    __shared__ AT s_val[WAVES_PER_WG][WAVE + WAVE / L];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = a.chunk0 + xcd_remap(blockIdx.x, gridDim.x) * WAVES_PER_WG + w;
    if (c >= a.chunk0 + a.nchunks) return;  // whole wave leaves; no block barriers below
    const Chunk ch = a.chunks[c];
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const int slot = lane / L, sub = lane % L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const int R = a.rank;
    for (uint32_t t0 = 0; t0 < ch.len; t0 += WAVE) {
        const uint32_t nvalid = min((uint32_t)WAVE, ch.len - t0);
        stage_tile<N, L, VT, AT>(a, bc, ch.begin + t0, nvalid, lane, s_idx[w], s_val[w]);
        for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
            const int col = cb + sub * C;
            segment_tile<N, FT, C, L, EXACT>(a, nvalid, slot, col, s_idx[w], s_val[w],
                [&](uint64_t row, const AT (&acc)[C]) {
                    AT* o = out + row * (uint64_t)R + col;
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc)
                        if (EXACT || col + cc < R) atomic_add_global(o + cc, acc[cc]);
                });
        }
        wave_lds_sync();
    }
}

// --------------------------------------------------------------- lvl3

template <int N, typename FT, typename VT, int C, int L, bool EXACT>
__global__ __launch_bounds__(WG) void k_mttkrp_lvl3(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    __shared__ uint64_t s_idx[WAVES_PER_WG][N][WAVE + WAVE / L];
    __shared__ AT s_val[WAVES_PER_WG][WAVE + WAVE / L];
    extern __shared__ __align__(16) unsigned char s_dyn[];
    AT* s_out = reinterpret_cast<AT*>(s_dyn);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int R = a.rank;
    const uint64_t cells = a.out_rows * (uint64_t)R;
    for (uint64_t e = threadIdx.x; e < cells; e += WG) s_out[e] = 0;
    __syncthreads();
    const uint32_t g = blockIdx.x;
    const uint32_t cbeg = a.chunk0 + g * a.chunks_per_wg;
    const uint32_t cend = min(cbeg + a.chunks_per_wg, a.chunk0 + a.nchunks);
    const int slot = lane / L, sub = lane % L;
    for (uint32_t c = cbeg + w; c < cend; c += WAVES_PER_WG) {
        const Chunk ch = a.chunks[c];
        uint64_t bc[N];
        #pragma unroll
        for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
        for (uint32_t t0 = 0; t0 < ch.len; t0 += WAVE) {
            const uint32_t nvalid = min((uint32_t)WAVE, ch.len - t0);
            stage_tile<N, L, VT, AT>(a, bc, ch.begin + t0, nvalid, lane, s_idx[w], s_val[w]);
            for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
                const int col = cb + sub * C;
                segment_tile<N, FT, C, L, EXACT>(a, nvalid, slot, col, s_idx[w], s_val[w],
                    [&](uint64_t row, const AT (&acc)[C]) {
                        AT* o = s_out + row * (uint64_t)R + col;
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc)
                            if (EXACT || col + cc < R) atomic_add_lds(o + cc, acc[cc]);
                    });
            }
            wave_lds_sync();
        }
    }
    __syncthreads();
    AT* dst = reinterpret_cast<AT*>(a.nprtn > 1 ? a.partials : a.out);
    if (a.nprtn > 1) dst += (uint64_t)(g % (uint32_t)a.nprtn) * cells;
    for (uint64_t e = threadIdx.x; e < cells; e += WG) {
        AT v = s_out[e];
        if (v != (AT)0) atomic_add_global(dst + e, v);
    }
}

// --------------------------------------------------------- row window

// One workgroup of WIN_WAVES waves per super-chunk.  LDS holds only the two
// accumulators: the hot area (the output mode's most frequent rows, by
// frequency rank) and the row window [lo, lo + lds_rows) of the super-chunk's
// other rows.  Tiles are decoded and sorted in registers; the segment walk
// reads the sorted entries with cross-lane permutes and issues the factor-row
// gathers of WIN_BATCH positions before consuming them.  A segment goes to
// the hot area, else to the window, else straight to HBM (atomic).  At the end
// the window and the hot area are written, whole, with plain coalesced stores
// to the workgroup's own partial rows; k_window_reduce adds the partial rows
// of every output row into the output.  On a multi-XCD device an atomic is
// performed beyond the XCD's L2, so the flush avoids atomics altogether.
template <int N, typename FT, typename VT, int C, int L, bool EXACT, typename IT>
__global__ __launch_bounds__(WIN_WAVES * WAVE) void k_mttkrp_window(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    constexpr int PB = L < WIN_BATCH ? L : WIN_BATCH;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int R = a.rank, mode = a.mode;
    AT* s_hot = reinterpret_cast<AT*>(s_dyn);
    AT* s_win = s_hot + (uint64_t)a.hot_rows * R;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t e = a.chunk0 + g;  // grid == entry count
    const SuperChunk sc = a.sc[e];
    const uint64_t lo = a.sc_range[2 * (uint64_t)e];
    const uint64_t wrows = a.sc_poff[e + 1] - a.sc_poff[e];  // <= lds_rows; 0 = no window
    const uint64_t cells = ((uint64_t)a.hot_rows + wrows) * R;
    for (uint64_t i = threadIdx.x; i < cells; i += WIN_WAVES * WAVE) s_hot[i] = (AT)0;
    __syncthreads();
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)sc.block * N + m];
    const int slot = lane / L, sub = lane % L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    const uint32_t ntiles = (sc.len + WAVE - 1) / WAVE;

    // dest: LDS row (hot area rows first, then the window) or -1 for HBM
    auto flush = [&](uint64_t row, int32_t dest, int col, const AT (&acc)[C]) {
        const bool lds = dest >= 0;
        AT* o = lds ? s_hot + (uint64_t)dest * R + col : out + row * (uint64_t)R + col;
        #pragma unroll
        for (int cc = 0; cc < C; ++cc) {
            if (!(EXACT || col + cc < R)) continue;
            if (lds) atomic_add_lds(o + cc, acc[cc]);
            else atomic_add_global(o + cc, acc[cc]);
        }
    };

    for (uint32_t t = w; t < ntiles; t += WIN_WAVES) {
        const uint32_t t0 = t * WAVE;
        const uint32_t nvalid = min((uint32_t)WAVE, sc.len - t0);
        // decode and sort the tile by output row, all in registers
        const bool valid = (uint32_t)lane < nvalid;
        const uint64_t el = sc.begin + t0 + lane - a.coord_base;
        const uint64_t coord = valid ? stream_load(a.coords + el) : 0;
        const AT v = valid ? (AT)stream_load(vals + el) : (AT)0;
        IT idx[N];
        IT orow = 0;
        #pragma unroll
        for (int m = 0; m < N; ++m) {
            idx[m] = (IT)(((coord >> a.pos[m]) & a.mask[m]) | bc[m]);
            if (m == mode) orow = idx[m];
        }
        const IT base = __shfl(orow, 0);
        const uint64_t diff = (uint64_t)orow - (uint64_t)base + (1ull << 25);
        uint32_t key;
        if (__ballot(valid && diff >= (1ull << 26)) == 0) {
            key = valid ? (((uint32_t)diff << 6) | (uint32_t)lane) : 0xFFFFFFFFu;
            key = wave_sort64(key, lane);
        } else {
            key = (uint32_t)lane;
        }
        const int src = (int)(key & 63u);
        IT sidx[N];
        #pragma unroll
        for (int m = 0; m < N; ++m) sidx[m] = __shfl(idx[m], src);
        const AT sv = __shfl(v, src);
        IT srow = 0;
        #pragma unroll
        for (int m = 0; m < N; ++m)
            if (m == mode) srow = sidx[m];
        // each lane classifies its (sorted) nonzero once: one independent load
        // instead of one dependent load per flushed segment
        int32_t sdest = -1;
        if ((uint32_t)lane < nvalid) {
            const int32_t hr = a.hot_rank ? a.hot_rank[srow] : -1;
            if (hr >= 0 && (uint32_t)hr < a.hot_rows) sdest = hr;
            else if ((uint64_t)srow >= lo && (uint64_t)srow - lo < wrows) sdest = (int32_t)(a.hot_rows + ((uint64_t)srow - lo));
        }
        // segment walk over positions [slot*L, slot*L + L)
        int kmax = (int)nvalid - slot * L;
        kmax = kmax < 0 ? 0 : (kmax > L ? L : kmax);
        for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
            const int col = cb + sub * C;
            AT acc[C];
            #pragma unroll
            for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
            uint64_t cur = 0;
            int32_t cur_dest = -1;
            bool have = false;
            #pragma unroll
            for (int k0 = 0; k0 < L; k0 += PB) {
                IT rr[PB];
                int32_t dd[PB];
                AT pv[PB][C];
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    const int q = slot * L + k0 + b;
                    rr[b] = __shfl(srow, q);
                    dd[b] = __shfl(sdest, q);
                    const AT vq = __shfl(sv, q);
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) pv[b][cc] = vq;
                    #pragma unroll
                    for (int m = 0; m < N; ++m) {
                        if (m == mode) continue;
                        const IT im = __shfl(sidx[m], q);
                        AT x[C];
                        if (k0 + b < kmax) {
                            const FT* f = reinterpret_cast<const FT*>(a.f[m]) + (uint64_t)im * R + col;
                            if (EXACT || col + C <= R) {
                                load_row<FT, C>(f, x);
                            } else {
                                #pragma unroll
                                for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
                            }
                        } else {
                            #pragma unroll
                            for (int cc = 0; cc < C; ++cc) x[cc] = 0;
                        }
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) pv[b][cc] *= x[cc];
                    }
                }
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    if (k0 + b >= kmax) break;
                    if (have && (uint64_t)rr[b] != cur) {
                        flush(cur, cur_dest, col, acc);
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
                    }
                    cur = rr[b];
                    cur_dest = dd[b];
                    have = true;
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pv[b][cc];
                }
            }
            if (have) flush(cur, cur_dest, col, acc);
        }
    }
    __syncthreads();
    AT* part = reinterpret_cast<AT*>(a.part);
    const uint64_t hcells = (uint64_t)a.hot_rows * R;
    // partial rows are read once, by the reduce: stored non-temporally
    if (hcells) {
        AT* hp = part + (a.hot_poff + (uint64_t)e * a.hot_rows) * R;
        for (uint64_t i = threadIdx.x; i < hcells; i += WIN_WAVES * WAVE) __builtin_nontemporal_store(s_hot[i], hp + i);
    }
    if (wrows) {
        AT* wp = part + a.sc_poff[e] * R;
        for (uint64_t i = threadIdx.x; i < wrows * R; i += WIN_WAVES * WAVE) __builtin_nontemporal_store(s_win[i], wp + i);
    }
}

// ------------------------------------------------------------ row-ordered

// MTTKRP over a row-ordered copy (every block's nonzeros sorted by the output
// index): a wave takes a chunk, each of its S = 64 / L slots a contiguous run
// of the chunk.  Per round the L lanes of a slot decode L consecutive
// nonzeros, one each, into the output row and the element offsets of the
// input factor rows (IT: 32-bit when every index and offset fits); the slot
// then walks them in order through cross-lane permutes, gathering WIN_BATCH
// positions' factor rows at a time and keeping a register accumulator while
// the output row repeats.  The next round's BLCO words and values are loaded
// before this round's gathers, so the two latencies overlap.  Output rows are
// contiguous, so a slot adds into HBM (atomically: a row may continue in the
// next slot or chunk) about once per row it touches; the rows that end the
// slots' runs are first combined across the slots of the wave.
template <int N, typename FT, typename VT, int C, int L, bool EXACT, typename IT>
__global__ __launch_bounds__(WG) void k_mttkrp_sorted(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    constexpr int S = WAVE / L;
    constexpr int PB = L < WIN_BATCH ? L : WIN_BATCH;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = a.chunk0 + xcd_remap(blockIdx.x, gridDim.x) * WAVES_PER_WG + w;
    if (c >= a.chunk0 + a.nchunks) return;  // whole wave leaves; no block barriers below
    const Chunk ch = a.chunks[c];
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const int slot = lane / L, sub = lane % L;
    const int R = a.rank, mode = a.mode;
    const uint32_t run = (ch.len + S - 1) / S;
    const uint32_t p0 = min((uint32_t)slot * run, ch.len), p1 = min(p0 + run, ch.len);
    const uint32_t rounds = (run + L - 1) / L;  // the same for every slot
    AT* out = reinterpret_cast<AT*>(a.out);
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    // 32-bit offsets: every factor holds fewer than 2^32 bytes (host-checked)
    constexpr bool BUF = sizeof(IT) == 4;
    __amdgpu_buffer_rsrc_t rs[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) rs[m] = factor_rsrc(a.f[m]);
    for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
        const int col = cb + sub * C;
        AT acc[C];
        #pragma unroll
        for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
        uint64_t cur = 0;
        bool have = false;
        auto flush = [&]() {
            AT* o = out + cur * (uint64_t)R + col;
            #pragma unroll
            for (int cc = 0; cc < C; ++cc)
                if (EXACT || col + cc < R) atomic_add_global(o + cc, acc[cc]);
        };
        auto fetch = [&](uint32_t r, uint64_t& coord, AT& v) {
            const uint32_t q = p0 + r * L + sub;
            const bool ok = r < rounds && q < p1;
            const uint64_t el = ch.begin + q - a.coord_base;
            coord = ok ? stream_load(a.coords + el) : 0;
            v = ok ? (AT)stream_load(vals + el) : (AT)0;
        };
        uint64_t coord;
        AT v;
        fetch(0, coord, v);
        for (uint32_t r = 0; r < rounds; ++r) {
            const uint32_t base = p0 + r * L;
            // decode this round's nonzero: output row, factor-row offsets
            IT row = 0, off[N];
            #pragma unroll
            for (int m = 0; m < N; ++m) {
                const uint64_t i = ((coord >> a.pos[m]) & a.mask[m]) | bc[m];
                off[m] = (IT)(i * (uint64_t)R);
                if (m == mode) row = (IT)i;
            }
            const AT vr = v;
            fetch(r + 1, coord, v);  // in flight during this round's gathers
            const int nb = base < p1 ? (int)min((uint32_t)L, p1 - base) : 0;
            #pragma unroll
            for (int k0 = 0; k0 < L; k0 += PB) {
                IT rr[PB];
                AT pv[PB][C];
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    const int src = slot * L + k0 + b;
                    rr[b] = __shfl(row, src);
                    const AT vq = __shfl(vr, src);
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) pv[b][cc] = vq;
                    #pragma unroll
                    for (int m = 0; m < N; ++m) {
                        if (m == mode) continue;
                        const IT om = __shfl(off[m], src);
                        AT x[C];
                        if (k0 + b < nb) {
                            const FT* f = reinterpret_cast<const FT*>(a.f[m]) + (uint64_t)om + col;
                            if (BUF && (EXACT || col + C <= R)) {
                                load_row_buf<FT, C>(rs[m], ((uint32_t)om + col) * (uint32_t)sizeof(FT), x);
                            } else if (EXACT || col + C <= R) {
                                load_row<FT, C>(f, x);
                            } else {
                                #pragma unroll
                                for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
                            }
                        } else {
                            #pragma unroll
                            for (int cc = 0; cc < C; ++cc) x[cc] = 0;
                        }
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) pv[b][cc] *= x[cc];
                    }
                }
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    if (k0 + b >= nb) break;
                    if (have && (uint64_t)rr[b] != cur) {
                        flush();
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
                    }
                    cur = rr[b];
                    have = true;
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pv[b][cc];
                }
            }
        }
        // The slots' last segments: a row that runs over several slots (a
        // hot row spans whole chunks) is summed across them first, with an
        // inclusive segmented scan over the slots, and added by the slot
        // that ends it: one HBM atomic per column per wave for that row.
        const int below = lane - L, above = lane + L;
        const uint64_t prev_cur = __shfl(cur, below < 0 ? lane : below);
        const bool prev_have = __shfl(have, below < 0 ? lane : below);
        bool head = slot == 0 || !have || !prev_have || prev_cur != cur;
        #pragma unroll
        for (int d = 1; d < S; d <<= 1) {
            const int from = lane - d * L < 0 ? lane : lane - d * L;
            const bool pf = __shfl(head, from);
            AT pa[C];
            #pragma unroll
            for (int cc = 0; cc < C; ++cc) pa[cc] = __shfl(acc[cc], from);
            if (slot >= d) {
                if (!head) {
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pa[cc];
                }
                head = head || pf;
            }
        }
        const uint64_t next_cur = __shfl(cur, above >= WAVE ? lane : above);
        const bool next_have = __shfl(have, above >= WAVE ? lane : above);
        const bool tail = slot == S - 1 || !next_have || next_cur != cur;
        if (have && tail) flush();
    }
}

// ------------------------------------------------------------- fibers

// Loads C columns of a factor row (zero past the rank when not EXACT).
template <typename FT, int C, bool EXACT>
__device__ __forceinline__ void gather_row(const FT* f, int col, int R, typename Acc<FT>::type (&x)[C])
{
    using AT = typename Acc<FT>::type;
    if (EXACT || col + C <= R) {
        load_row<FT, C>(f, x);
    } else {
        #pragma unroll
        for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
    }
}

// MTTKRP over a fiber-ordered copy: the row-ordered copy whose nonzeros are
// further ordered, within each output row, by the outer modes (every input
// mode but the inner one, a.inner), so that a fiber -- the nonzeros sharing
// the output row and all outer indices -- is contiguous.  Per nonzero only
// the inner factor row is gathered; the outer factor rows are gathered once
// per fiber, with its first nonzero, and scale every inner term of it:
//   out[i] += sum over fibers (prod_outer U_m[i_m]) * (v * U_inner[i_inner]).
// Slots, rounds, the prefetch and the flush are those of k_mttkrp_sorted; a
// lane also marks whether its nonzero starts a fiber (its key differs from
// the previous position's, or it is the first of the slot's run).
template <int N, typename FT, typename VT, int C, int L, bool EXACT, typename IT>
__global__ __launch_bounds__(WG) void k_mttkrp_fiber(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    constexpr int S = WAVE / L;
    // two positions per batch: the outer rows of fiber starts need registers
    // beside the inner rows, and four positions lower the occupancy to 4 waves
    constexpr int PB = L < 2 ? L : 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = a.chunk0 + xcd_remap(blockIdx.x, gridDim.x) * WAVES_PER_WG + w;
    if (c >= a.chunk0 + a.nchunks) return;  // whole wave leaves; no block barriers below
    const Chunk ch = a.chunks[c];
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const int slot = lane / L, sub = lane % L;
    const int R = a.rank, mode = a.mode, inner = a.inner;
    const uint32_t run = (ch.len + S - 1) / S;
    const uint32_t p0 = min((uint32_t)slot * run, ch.len), p1 = min(p0 + run, ch.len);
    const uint32_t rounds = (run + L - 1) / L;  // the same for every slot
    AT* out = reinterpret_cast<AT*>(a.out);
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    const FT* fin = reinterpret_cast<const FT*>(a.f[inner]);
    // 32-bit offsets: every factor holds fewer than 2^32 bytes (host-checked)
    constexpr bool BUF = sizeof(IT) == 4;
    __amdgpu_buffer_rsrc_t rs[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) rs[m] = factor_rsrc(a.f[m]);
    const __amdgpu_buffer_rsrc_t rs_in = factor_rsrc(fin);
    for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
        const int col = cb + sub * C;
        AT acc[C], scale[C];
        #pragma unroll
        for (int cc = 0; cc < C; ++cc) acc[cc] = scale[cc] = 0;
        IT cur = 0;
        bool have = false;
        auto flush = [&]() {
            AT* o = out + (uint64_t)cur * R + col;
            #pragma unroll
            for (int cc = 0; cc < C; ++cc)
                if (EXACT || col + cc < R) atomic_add_global(o + cc, acc[cc]);
        };
        auto fetch = [&](uint32_t r, uint64_t& coord, AT& v) {
            const uint32_t q = p0 + r * L + sub;
            const bool ok = r < rounds && q < p1;
            const uint64_t el = ch.begin + q - a.coord_base;
            coord = ok ? stream_load(a.coords + el) : 0;
            v = ok ? (AT)stream_load(vals + el) : (AT)0;
        };
        uint64_t coord;
        AT v;
        fetch(0, coord, v);
        // fiber key of the last position decoded by the slot in the previous round
        IT carry[N];
        #pragma unroll
        for (int m = 0; m < N; ++m) carry[m] = 0;
        for (uint32_t r = 0; r < rounds; ++r) {
            const uint32_t base = p0 + r * L;
            // decode: output row, inner offset, outer offsets (off[m] for
            // m not the output or inner mode), and the fiber-start mark
            IT off[N], row = 0, ioff = 0;
            #pragma unroll
            for (int m = 0; m < N; ++m) {
                const uint64_t i = ((coord >> a.pos[m]) & a.mask[m]) | bc[m];
                off[m] = (IT)(i * (uint64_t)R);
                if (m == mode) row = (IT)i;
                if (m == inner) ioff = off[m];
            }
            bool start = base + sub == p0;
            #pragma unroll
            for (int m = 0; m < N; ++m) {
                if (m == inner) continue;
                const IT up = __shfl(off[m], sub > 0 ? lane - 1 : lane);
                start = start || (sub > 0 ? up : carry[m]) != off[m];
            }
            #pragma unroll
            for (int m = 0; m < N; ++m) carry[m] = __shfl(off[m], slot * L + L - 1);
            const AT vr = v;
            fetch(r + 1, coord, v);  // in flight during this round's gathers
            const int nb = base < p1 ? (int)min((uint32_t)L, p1 - base) : 0;
            #pragma unroll
            for (int k0 = 0; k0 < L; k0 += PB) {
                IT rr[PB];
                bool fs[PB];
                AT pv[PB][C], po[PB][C];
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    const int src = slot * L + k0 + b;
                    const bool ok = k0 + b < nb;
                    rr[b] = __shfl(row, src);
                    fs[b] = __shfl((int)start, src) != 0;
                    const AT vq = __shfl(vr, src);
                    const IT iq = __shfl(ioff, src);
                    AT x[C];
                    if (ok) {
                        if (BUF && (EXACT || col + C <= R))
                            load_row_buf<FT, C>(rs_in, ((uint32_t)iq + col) * (uint32_t)sizeof(FT), x);
                        else
                            gather_row<FT, C, EXACT>(fin + (uint64_t)iq + col, col, R, x);
                    } else {
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) x[cc] = 0;
                    }
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) {
                        pv[b][cc] = vq * x[cc];
                        po[b][cc] = 1;
                    }
                    #pragma unroll
                    for (int m = 0; m < N; ++m) {
                        if (m == mode || m == inner) continue;
                        const IT om = __shfl(off[m], src);
                        if (ok && fs[b]) {
                            if (BUF && (EXACT || col + C <= R))
                                load_row_buf<FT, C>(rs[m], ((uint32_t)om + col) * (uint32_t)sizeof(FT), x);
                            else
                                gather_row<FT, C, EXACT>(reinterpret_cast<const FT*>(a.f[m]) + (uint64_t)om + col,
                                                         col, R, x);
                            #pragma unroll
                            for (int cc = 0; cc < C; ++cc) po[b][cc] *= x[cc];
                        }
                    }
                }
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    if (k0 + b >= nb) break;
                    if (fs[b]) {
                        if (have && rr[b] != cur) {
                            flush();
                            #pragma unroll
                            for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
                        }
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) scale[cc] = po[b][cc];
                        cur = rr[b];
                        have = true;
                    }
                    // the fiber's outer product scales each inner term as it
                    // is added (no separate fiber sum: four fewer VGPRs)
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pv[b][cc] * scale[cc];
                }
            }
        }
        // the slots' last rows: segmented scan over the slots, as in k_mttkrp_sorted
        const int below = lane - L, above = lane + L;
        const IT prev_cur = __shfl(cur, below < 0 ? lane : below);
        const bool prev_have = __shfl(have, below < 0 ? lane : below);
        bool head = slot == 0 || !have || !prev_have || prev_cur != cur;
        #pragma unroll
        for (int d = 1; d < S; d <<= 1) {
            const int from = lane - d * L < 0 ? lane : lane - d * L;
            const bool pf = __shfl(head, from);
            AT pa[C];
            #pragma unroll
            for (int cc = 0; cc < C; ++cc) pa[cc] = __shfl(acc[cc], from);
            if (slot >= d) {
                if (!head) {
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pa[cc];
                }
                head = head || pf;
            }
        }
        const IT next_cur = __shfl(cur, above >= WAVE ? lane : above);
        const bool next_have = __shfl(have, above >= WAVE ? lane : above);
        const bool tail = slot == S - 1 || !next_have || next_cur != cur;
        if (have && tail) flush();
    }
}

// MTTKRP over a row-ordered copy with every round's decoded nonzeros staged
// in LDS (the form of kernel 14 for factors of fewer than FACTOR_BYTES_MAX
// bytes; k_mttkrp_sorted and k_mttkrp_fiber remain for larger ones).  Slots
// and rounds are those of k_mttkrp_sorted: each lane decodes the nonzero of
// its own position once and writes one record -- the output row (with the
// fiber-start bit when FIB), the 32-bit byte offsets of the input factor
// rows, the value -- to its wave's LDS area; the slot's lanes then read
// each position's record with LDS loads at immediate offsets, in place of a
// cross-lane permute per field.  Positions past the end of a slot's run read
// the run's last nonzero with value 0 (and no fiber start), so the walk has
// no masked branches.
//   FIB = false: every input factor row is gathered per nonzero (the plain
//   row-ordered kernel); the input modes in increasing order.
//   FIB = true: the copy is in fiber order (k_mttkrp_fiber): the inner row
//   is gathered per nonzero, the outer rows by the lanes whose position
//   starts a fiber, and their product scales the fiber's terms.
// The flush and the final segmented scan over the slots are those of
// k_mttkrp_sorted.
template <int N, typename FT, typename VT, int C, int L, bool EXACT, bool FIB, bool HOT>
__global__ __launch_bounds__(WG) void k_mttkrp_rows(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    constexpr int S = WAVE / L;
    constexpr int RW = N + (int)(sizeof(AT) / 4);  // row, N-1 offsets, value
    constexpr int RWP = RW <= 4 ? 4 : 8;
    // positions whose gathers are issued together (FP64 rows need twice the
    // registers: two positions keep 4 waves per SIMD)
    constexpr int PB = sizeof(AT) == 8 ? 2 : 4;
    static_assert(L % PB == 0, "positions per batch divide the slot width");
    // a slot's records, then 16 bytes of padding: the slots' records of one
    // position then start in different LDS banks
    constexpr int SLOT_W = L * RWP + 4;
    __shared__ __attribute__((aligned(16))) uint32_t s_rec[WAVES_PER_WG][S * SLOT_W];
    // the hot-row windows of the first two input modes (dynamic LDS), each a
    // contiguous run of factor rows copied as it lies in memory
    extern __shared__ __attribute__((aligned(16))) unsigned char s_tab[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int R = a.rank, mode = a.mode;
    const uint32_t row_bytes = (uint32_t)R * (uint32_t)sizeof(FT);
    // input modes: FIB, the inner one first, then the outer ones
    int in[N - 1] = {};
    {
        int j = 0;
        if (FIB) in[j++] = a.inner;
        #pragma unroll
        for (int m = 0; m < N; ++m)
            if (m != mode && !(FIB && m == a.inner) && j < N - 1) in[j++] = m;
    }
    constexpr int NT = !HOT ? 0 : N - 1 < 2 ? N - 1 : 2;  // input modes with a window
    uint32_t tab_off[NT > 0 ? NT : 1] = {};
    if (HOT) {
        uint32_t o = 0;
        #pragma unroll
        for (int t = 0; t < NT; ++t) {
            tab_off[t] = o;
            const uint32_t nb = a.hot_n[t] * row_bytes;  // a multiple of 16 (host-checked)
            const unsigned char* src = reinterpret_cast<const unsigned char*>(a.f[in[t]]) +
                                       (uint64_t)a.hot_lo[t] * row_bytes;
            for (uint32_t i = threadIdx.x * 16u; i < nb; i += WG * 16u)
                *reinterpret_cast<uint4*>(s_tab + o + i) = *reinterpret_cast<const uint4*>(src + i);
            o += nb;
        }
    }
    if (HOT) __syncthreads();  // the only block barrier: waves leave independently below
    const int slot = lane / L, sub = lane % L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    __amdgpu_buffer_rsrc_t rs[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) rs[m] = factor_rsrc(a.f[m]);
    uint32_t* my_rec = s_rec[w] + slot * SLOT_W + sub * RWP;
    const uint32_t* slot_rec = s_rec[w] + slot * SLOT_W;
    // the wave's chunks: a contiguous run of cpw chunks of the workgroup's
    // range (logical workgroups of one XCD take consecutive ranges)
    const uint32_t lw = xcd_remap(blockIdx.x, gridDim.x) * WAVES_PER_WG + w;
    for (uint32_t cj = 0; cj < a.cpw; ++cj) {
    // interleaved: the whole grid walks the chunks in step (the chunks in
    // flight stay neighbours); else a contiguous run per wave
    const uint32_t rel = a.cpw_interleave ? cj * gridDim.x * WAVES_PER_WG + lw : lw * a.cpw + cj;
    if (rel >= a.nchunks) break;
    const uint32_t c = a.chunk0 + rel;
    const Chunk ch = a.chunks[c];
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const uint32_t run = (ch.len + S - 1) / S;
    const uint32_t p0 = min((uint32_t)slot * run, ch.len), p1 = min(p0 + run, ch.len);
    const uint32_t rounds = (run + L - 1) / L;  // the same for every slot
    for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
        const int col = cb + sub * C;
        const uint32_t colb = (uint32_t)col * (uint32_t)sizeof(FT);
        AT acc[C], scale[C];
        #pragma unroll
        for (int cc = 0; cc < C; ++cc) acc[cc] = scale[cc] = 0;
        uint32_t cur = 0;
        bool have = false;
        auto flush = [&]() {
            AT* o = out + (uint64_t)cur * R + col;
            #pragma unroll
            for (int cc = 0; cc < C; ++cc)
                if (EXACT || col + cc < R) atomic_add_global(o + cc, acc[cc]);
        };
        // a row piece from the factor, or (offset with bit 31 set) from its
        // LDS window
        auto row_piece = [&](int m, uint32_t byte_off, AT (&x)[C]) {
            if (HOT && (byte_off >> 31)) {
                const unsigned char* t = s_tab + (byte_off & 0x7fffffffu);
                if (EXACT || col + C <= R) {
                    load_row<FT, C>(reinterpret_cast<const FT*>(t + colb), x);
                } else {
                    const FT* f = reinterpret_cast<const FT*>(t) + col;
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
                }
            } else if (EXACT || col + C <= R) {
                load_row_buf<FT, C>(rs[m], byte_off + colb, x);
            } else {
                const FT* f = reinterpret_cast<const FT*>(reinterpret_cast<const char*>(a.f[m]) + byte_off) + col;
                #pragma unroll
                for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
            }
        };
        // a position past the run reads the run's last nonzero (p1 >= 1)
        auto fetch = [&](uint32_t r, uint64_t& coord, AT& v, bool& valid) {
            const uint32_t q = p0 + r * L + sub;
            valid = r < rounds && q < p1;
            const uint64_t el = ch.begin + (valid ? q : p1 - 1) - a.coord_base;
            coord = stream_load(a.coords + el);
            const AT x = (AT)stream_load(vals + el);
            v = valid ? x : (AT)0;
        };
        uint64_t coord;
        AT v;
        bool valid;
        fetch(0, coord, v, valid);
        uint32_t carry[N];  // FIB: the fiber key of the slot's last position
        #pragma unroll
        for (int m = 0; m < N; ++m) carry[m] = 0;
        for (uint32_t r = 0; r < rounds; ++r) {
            const uint32_t base = p0 + r * L;
            uint32_t off[N], idx[N], row = 0;
            #pragma unroll
            for (int m = 0; m < N; ++m) {
                const uint64_t i = ((coord >> a.pos[m]) & a.mask[m]) | bc[m];
                idx[m] = (uint32_t)i;
                off[m] = (uint32_t)i * row_bytes;
                if (m == mode) row = (uint32_t)i;
            }
            uint32_t wv[RWP];
            #pragma unroll
            for (int k = 0; k < RWP; ++k) wv[k] = 0;
            wv[0] = row;
            if (FIB) {
                // a fiber starts where the output row or an outer index changes
                bool start = base + sub == p0;
                #pragma unroll
                for (int m = 0; m < N; ++m) {
                    const uint32_t up = __shfl(off[m], sub > 0 ? lane - 1 : lane);
                    const bool key = m != a.inner;
                    start = start || (key && (sub > 0 ? up : carry[m]) != off[m]);
                }
                #pragma unroll
                for (int m = 0; m < N; ++m) carry[m] = __shfl(off[m], slot * L + L - 1);
                wv[0] |= (start && valid) ? 0x80000000u : 0u;
            }
            #pragma unroll
            for (int j = 0; j < N - 1; ++j) {
                #pragma unroll
                for (int m = 0; m < N; ++m)
                    if (m == in[j]) wv[1 + j] = off[m];
            }
            // rows in a mode's LDS window: their offset there, bit 31 set
            #pragma unroll
            for (int t = 0; t < NT; ++t) {
                uint32_t ix = 0;
                #pragma unroll
                for (int m = 0; m < N; ++m)
                    if (m == in[t]) ix = idx[m];
                const uint32_t h = ix - a.hot_lo[t];
                if (h < a.hot_n[t]) wv[1 + t] = 0x80000000u | (tab_off[t] + h * row_bytes);
            }
            __builtin_memcpy(&wv[N], &v, sizeof(AT));
            wave_lds_sync();  // the previous round's reads are done
            #pragma unroll
            for (int k = 0; k < RWP; k += 4)
                *reinterpret_cast<uint4*>(my_rec + k) = uint4{wv[k], wv[k + 1], wv[k + 2], wv[k + 3]};
            wave_lds_sync();
            #pragma unroll
            for (int k0 = 0; k0 < L; k0 += PB) {
                uint32_t rv[PB][RWP];
                AT x[PB][C], po[PB][C];
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    #pragma unroll
                    for (int k = 0; k < RWP; k += 4) {
                        const uint4 u = *reinterpret_cast<const uint4*>(slot_rec + (k0 + b) * RWP + k);
                        rv[b][k] = u.x; rv[b][k + 1] = u.y; rv[b][k + 2] = u.z; rv[b][k + 3] = u.w;
                    }
                }
                if (FIB) {
                    #pragma unroll
                    for (int b = 0; b < PB; ++b) row_piece(in[0], rv[b][1], x[b]);
                    // outer rows, by the lanes whose position starts a fiber
                    #pragma unroll
                    for (int b = 0; b < PB; ++b) {
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) po[b][cc] = 1;
                        if (rv[b][0] >> 31) {
                            #pragma unroll
                            for (int j = 1; j < N - 1; ++j) {
                                AT y[C];
                                row_piece(in[j], rv[b][1 + j], y);
                                #pragma unroll
                                for (int cc = 0; cc < C; ++cc) po[b][cc] *= y[cc];
                            }
                        }
                    }
                } else {
                    // the product of the value and every input row
                    #pragma unroll
                    for (int b = 0; b < PB; ++b) {
                        AT vb;
                        __builtin_memcpy(&vb, &rv[b][N], sizeof(AT));
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) x[b][cc] = vb;
                        #pragma unroll
                        for (int j = 0; j < N - 1; ++j) {
                            AT y[C];
                            row_piece(in[j], rv[b][1 + j], y);
                            #pragma unroll
                            for (int cc = 0; cc < C; ++cc) x[b][cc] *= y[cc];
                        }
                    }
                }
                // the next round's nonzero, behind this batch's gathers
                if (k0 == 0) fetch(r + 1, coord, v, valid);
                #pragma unroll
                for (int b = 0; b < PB; ++b) {
                    const uint32_t rb = FIB ? rv[b][0] & 0x7fffffffu : rv[b][0];
                    const bool fs = FIB ? (rv[b][0] >> 31) != 0 : true;
                    if (fs && have && rb != cur) {
                        flush();
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) acc[cc] = 0;
                    }
                    cur = fs ? rb : cur;
                    have = have || fs;
                    if (FIB) {
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) scale[cc] = fs ? po[b][cc] : scale[cc];
                        AT vb;
                        __builtin_memcpy(&vb, &rv[b][N], sizeof(AT));
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) acc[cc] += (vb * x[b][cc]) * scale[cc];
                    } else {
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) acc[cc] += x[b][cc];
                    }
                }
            }
        }
        // the slots' last rows: segmented scan over the slots, as in k_mttkrp_sorted
        const int below = lane - L, above = lane + L;
        const uint32_t prev_cur = __shfl(cur, below < 0 ? lane : below);
        const bool prev_have = __shfl(have, below < 0 ? lane : below);
        bool head = slot == 0 || !have || !prev_have || prev_cur != cur;
        #pragma unroll
        for (int d = 1; d < S; d <<= 1) {
            const int from = lane - d * L < 0 ? lane : lane - d * L;
            const bool pf = __shfl(head, from);
            AT pa[C];
            #pragma unroll
            for (int cc = 0; cc < C; ++cc) pa[cc] = __shfl(acc[cc], from);
            if (slot >= d) {
                if (!head) {
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) acc[cc] += pa[cc];
                }
                head = head || pf;
            }
        }
        const uint32_t next_cur = __shfl(cur, above >= WAVE ? lane : above);
        const bool next_have = __shfl(have, above >= WAVE ? lane : above);
        const bool tail = slot == S - 1 || !next_have || next_cur != cur;
        if (have && tail) flush();
        wave_lds_sync();  // the next column block rewrites the records
    }
    }
}

// ------------------------------------------------------------- atomic

template <int N, typename FT, typename VT, int C, int L, bool EXACT>
__global__ __launch_bounds__(WG) void k_mttkrp_atomic(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t c = a.chunk0 + xcd_remap(blockIdx.x, gridDim.x) * WAVES_PER_WG + w;
    if (c >= a.chunk0 + a.nchunks) return;
    const Chunk ch = a.chunks[c];
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const int slot = lane / L, sub = lane % L;
    constexpr int S = WAVE / L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const VT* vals = reinterpret_cast<const VT*>(a.vals);
    const int R = a.rank, mode = a.mode;
    for (uint32_t t0 = 0; t0 < ch.len; t0 += WAVE) {
        const uint32_t nvalid = min((uint32_t)WAVE, ch.len - t0);
        const bool valid = (uint32_t)lane < nvalid;
        const uint64_t e = ch.begin + t0 + lane - a.coord_base;
        const uint64_t coord = valid ? stream_load(a.coords + e) : 0;
        const AT v = valid ? (AT)stream_load(vals + e) : (AT)0;
        uint64_t idx[N];
        #pragma unroll
        for (int m = 0; m < N; ++m) idx[m] = ((coord >> a.pos[m]) & a.mask[m]) | bc[m];
        for (int k = 0; k < L; ++k) {
            const int q = k * S + slot;  // nonzero handled by this slot in step k
            uint64_t rows[N];
            #pragma unroll
            for (int m = 0; m < N; ++m) rows[m] = __shfl(idx[m], q);
            const AT vv = __shfl(v, q);
            if ((uint32_t)q >= nvalid) continue;
            for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
                const int col = cb + sub * C;
                AT p[C];
                #pragma unroll
                for (int cc = 0; cc < C; ++cc) p[cc] = vv;
                uint64_t orow = 0;
                #pragma unroll
                for (int m = 0; m < N; ++m) {
                    if (m == mode) { orow = rows[m]; continue; }
                    const FT* f = reinterpret_cast<const FT*>(a.f[m]) + rows[m] * (uint64_t)R + col;
                    AT x[C];
                    if (EXACT || col + C <= R) {
                        load_row<FT, C>(f, x);
                    } else {
                        #pragma unroll
                        for (int cc = 0; cc < C; ++cc) x[cc] = (col + cc < R) ? (AT)to_acc(f[cc]) : (AT)0;
                    }
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) p[cc] *= x[cc];
                }
                AT* o = out + orow * (uint64_t)R + col;
                #pragma unroll
                for (int cc = 0; cc < C; ++cc)
                    if (EXACT || col + cc < R) atomic_add_global(o + cc, p[cc]);
            }
        }
    }
}

// Column layout (C columns per lane, L lanes per nonzero) for a rank and
// element size; EXACT when C * L == rank.
struct Shape {
    int C, L;
    bool exact;
};

template <int N, typename FT, typename VT, int C, int L, bool EX>
inline const void* kernel_ptr(int which)
{
    if (which == W_LVL1) return (const void*)&k_mttkrp_lvl1<N, FT, VT, C, L, EX>;
    if (which == W_LVL3) return (const void*)&k_mttkrp_lvl3<N, FT, VT, C, L, EX>;
    if (which == W_WINDOW) return (const void*)&k_mttkrp_window<N, FT, VT, C, L, EX, uint32_t>;
    if (which == W_WINDOW64) return (const void*)&k_mttkrp_window<N, FT, VT, C, L, EX, uint64_t>;
    if (which == W_SORTED) return (const void*)&k_mttkrp_rows<N, FT, VT, C, L, EX, false, false>;
    if (which == W_SORTED_LDS) return (const void*)&k_mttkrp_rows<N, FT, VT, C, L, EX, false, true>;
    if (which == W_SORTED64) return (const void*)&k_mttkrp_sorted<N, FT, VT, C, L, EX, uint64_t>;
    if (which == W_FIBER || which == W_FIBER64 || which == W_FIBER_LDS) {
        if constexpr (N >= 3)
            return which == W_FIBER       ? (const void*)&k_mttkrp_rows<N, FT, VT, C, L, EX, true, false>
                   : which == W_FIBER_LDS ? (const void*)&k_mttkrp_rows<N, FT, VT, C, L, EX, true, true>
                                          : (const void*)&k_mttkrp_fiber<N, FT, VT, C, L, EX, uint64_t>;
        return nullptr;  // an order-2 tensor has no fibers
    }
    return (const void*)&k_mttkrp_atomic<N, FT, VT, C, L, EX>;
}

template <int N, typename FT, typename VT>
inline const void* select_shape(const Shape& s, int which)
{
    if (s.exact) {
        if (s.C == 1 && s.L == 8) return kernel_ptr<N, FT, VT, 1, 8, true>(which);
        if (s.C == 2 && s.L == 8) return kernel_ptr<N, FT, VT, 2, 8, true>(which);
        if constexpr (sizeof(FT) == 8) {
            if (s.C == 4 && s.L == 8) return kernel_ptr<N, FT, VT, 4, 8, true>(which);
            if (s.C == 4 && s.L == 16) return kernel_ptr<N, FT, VT, 4, 16, true>(which);
        } else if constexpr (sizeof(FT) == 4) {
            if (s.C == 4 && s.L == 8) return kernel_ptr<N, FT, VT, 4, 8, true>(which);
            if (s.C == 4 && s.L == 16) return kernel_ptr<N, FT, VT, 4, 16, true>(which);
        } else {
            if (s.C == 4 && s.L == 8) return kernel_ptr<N, FT, VT, 4, 8, true>(which);
            if (s.C == 8 && s.L == 8) return kernel_ptr<N, FT, VT, 8, 8, true>(which);
        }
        return nullptr;
    }
    if (s.L == 8) return kernel_ptr<N, FT, VT, 1, 8, false>(which);
    if (s.L == 16) return kernel_ptr<N, FT, VT, 1, 16, false>(which);
    if (s.L == 32) return kernel_ptr<N, FT, VT, 1, 32, false>(which);
    return kernel_ptr<N, FT, VT, 1, 64, false>(which);
}

template <typename FT, typename VT>
inline const void* select_n(int N, const Shape& s, int which)
{
    switch (N) {
        case 2: return select_shape<2, FT, VT>(s, which);
        case 3: return select_shape<3, FT, VT>(s, which);
        case 4: return select_shape<4, FT, VT>(s, which);
        case 5: return select_shape<5, FT, VT>(s, which);
        default: return nullptr;
    }
}

inline Shape pick_shape(int rank, int esize)
{
    switch (rank) {
        case 8: return {1, 8, true};
        case 16: return {2, 8, true};
        case 32: return Shape{4, 8, true};
        case 64: return esize == 2 ? Shape{8, 8, true} : Shape{4, 16, true};
        default: break;
    }
    int L = 8;
    while (L < rank && L < WAVE) L <<= 1;
    return {1, L, false};
}

const void* mttkrp_kernel_f64(int N, const Shape& s, int which);
const void* mttkrp_kernel_f32(int N, const Shape& s, int which);
const void* mttkrp_kernel_bf16(int N, const Shape& s, int which);

}  // namespace blco
