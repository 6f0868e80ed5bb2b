// mttkrp.hip -- MTTKRP kernels on BLCO tensors for gfx950.
//
// Computes out[i_n, :] += sum over nonzeros x of val(x) * prod_{m != n}
// U_m[i_m(x), :] (src/alto_dev.cu:31-306 lvl1, :1247-1629 lvl3, :1680-1847
// atomic).  The work unit is a chunk: a run of 64 * tiles nonzeros of one
// BLCO block, handled by one wave64.
//
// lvl1 ("segmented"): per 64-nonzero tile the wave decodes every nonzero
// ((coord >> pos_m) & mask_m) | block_coord_m, sorts the 64 output rows with
// a bitonic network over the lanes (32-bit keys: row offset from lane 0's
// row, shifted left 6, | lane), gathers the sorted entries through
// cross-lane permutes into a wave-private LDS strip, and then lanes work as
// S slots x L lanes: slot s walks sorted positions [s*L, s*L + L), each lane
// owning C consecutive columns (one 16-byte row load per factor), keeps a
// register accumulator while the output row repeats and issues one HBM
// atomic per (segment, column) when it changes.  This replaces the
// reference's match_any / labeled_partition / shuffle-scan counting sort,
// which has no wave64 equivalent.
//
// lvl3 ("privatised"): the same per-tile segmentation, but segments are
// added into a workgroup-private copy of the whole output held in LDS (short
// modes only); at the end the workgroup adds its copy into partial matrix
// (workgroup % nprtn) -- or straight into the output when nprtn == 1 -- and
// a pull reduction sums the partial matrices (partial_matrix_reduction,
// src/alto_dev.cu:1589-1601).
//
// atomic: no segmentation; one atomic per (nonzero, column)
// (mttkrp_atomic_*, src/alto_dev.cu:1680-1829).
//
// Workgroups are remapped so that each XCD (own L2) walks a contiguous range
// of chunks, i.e. a contiguous range of the ALTO order, which bounds the
// factor rows each L2 sees.
#include "devcommon.hpp"
#include "mttkrp.hpp"

namespace blco {

// ------------------------------------------------------------ row loads

template <typename FT, int C>
__device__ __forceinline__ void load_row(const FT* __restrict__ p,
                                         typename Acc<FT>::type (&x)[C])
{
    using AT = typename Acc<FT>::type;
    if constexpr (C * sizeof(FT) == 16) {
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
    uint64_t coord = valid ? a.coords[e0 + lane - a.coord_base] : 0;
    AT v = valid ? (AT)vals[e0 + lane - a.coord_base] : (AT)0;
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

// One workgroup per super-chunk.  LDS holds two accumulators: the hot area
// (the output mode's most frequent rows, by frequency rank) and the row
// window [lo, lo + lds_rows) of the super-chunk's other rows.  A segment goes
// to the hot area, else to the window, else straight to HBM.  At the end the
// window is added into the output and the hot area into partial copy
// (workgroup % hot_parts), both with coalesced HBM atomics skipping zero
// cells; k_hot_reduce folds the copies into the output.
template <int N, typename FT, typename VT, int C, int L, bool EXACT>
__global__ __launch_bounds__(WG) void k_mttkrp_window(MttkrpArgs a)
{
    using AT = typename Acc<FT>::type;
    __shared__ uint64_t s_idx[WAVES_PER_WG][N][WAVE + WAVE / L];
    __shared__ AT s_val[WAVES_PER_WG][WAVE + WAVE / L];
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int R = a.rank;
    AT* s_hot = reinterpret_cast<AT*>(s_dyn);
    AT* s_win = s_hot + (uint64_t)a.hot_rows * R;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t e = a.chunk0 + g;  // grid == entry count
    const SuperChunk sc = a.sc[e];
    const uint64_t lo = a.sc_range[2 * (uint64_t)e], hi = a.sc_range[2 * (uint64_t)e + 1];
    const uint64_t wrows = hi >= lo ? min(hi - lo + 1, (uint64_t)a.lds_rows) : 0;
    const uint64_t cells = ((uint64_t)a.hot_rows + wrows) * R;
    for (uint64_t i = threadIdx.x; i < cells; i += WG) s_hot[i] = (AT)0;
    __syncthreads();
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)sc.block * N + m];
    const int slot = lane / L, sub = lane % L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const uint32_t ntiles = (sc.len + WAVE - 1) / WAVE;
    for (uint32_t t = w; t < ntiles; t += WAVES_PER_WG) {
        const uint32_t t0 = t * WAVE;
        const uint32_t nvalid = min((uint32_t)WAVE, sc.len - t0);
        stage_tile<N, L, VT, AT>(a, bc, sc.begin + t0, nvalid, lane, s_idx[w], s_val[w]);
        for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
            const int col = cb + sub * C;
            segment_tile<N, FT, C, L, EXACT>(a, nvalid, slot, col, s_idx[w], s_val[w],
                [&](uint64_t row, const AT (&acc)[C]) {
                    const int32_t hr = a.hot_rank ? a.hot_rank[row] : -1;
                    AT* o;
                    bool lds = true;
                    if (hr >= 0 && (uint32_t)hr < a.hot_rows) o = s_hot + (uint64_t)hr * R + col;
                    else if (row >= lo && row - lo < wrows) o = s_win + (row - lo) * (uint64_t)R + col;
                    else { o = out + row * (uint64_t)R + col; lds = false; }
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc) {
                        if (!(EXACT || col + cc < R)) continue;
                        if (lds) atomic_add_lds(o + cc, acc[cc]);
                        else atomic_add_global(o + cc, acc[cc]);
                    }
                });
        }
        wave_lds_sync();
    }
    __syncthreads();
    const uint64_t hcells = (uint64_t)a.hot_rows * R;
    if (hcells) {
        AT* hp = reinterpret_cast<AT*>(a.hotpart) + (uint64_t)(g % (uint32_t)a.hot_parts) * hcells;
        for (uint64_t i = threadIdx.x; i < hcells; i += WG) {
            const AT v = s_hot[i];
            if (v != (AT)0) atomic_add_global(hp + i, v);
        }
    }
    AT* dst = out + lo * (uint64_t)R;
    for (uint64_t i = threadIdx.x; i < wrows * R; i += WG) {
        const AT v = s_win[i];
        if (v != (AT)0) atomic_add_global(dst + i, v);
    }
}

// out[hot_list[h]][c] += sum_p part[p][h][c]; the copies are cleared for the
// next call.
template <typename AT>
__global__ __launch_bounds__(WG) void k_hot_reduce(AT* __restrict__ out, AT* __restrict__ part,
                                                   const uint32_t* __restrict__ hot_list,
                                                   uint32_t hot_rows, int R, int parts)
{
    const uint64_t cells = (uint64_t)hot_rows * R;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * WG) {
        AT s = 0;
        for (int p = 0; p < parts; ++p) {
            s += part[(uint64_t)p * cells + i];
            part[(uint64_t)p * cells + i] = 0;
        }
        const uint64_t h = i / R, c = i % R;
        out[(uint64_t)hot_list[h] * R + c] += s;
    }
}

int launch_hot_reduce(void* out, void* hotpart, const uint32_t* hot_list, uint32_t hot_rows,
                      int rank, int parts, int otype, hipStream_t st)
{
    const uint64_t cells = (uint64_t)hot_rows * rank;
    if (!cells) return BLCO_OK;
    uint32_t grid = (uint32_t)std::min<uint64_t>((cells + WG - 1) / WG, 1024);
    if (otype == BLCO_F64)
        hipLaunchKernelGGL(k_hot_reduce<double>, dim3(grid), dim3(WG), 0, st, (double*)out,
                           (double*)hotpart, hot_list, hot_rows, rank, parts);
    else
        hipLaunchKernelGGL(k_hot_reduce<float>, dim3(grid), dim3(WG), 0, st, (float*)out,
                           (float*)hotpart, hot_list, hot_rows, rank, parts);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
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
        const uint64_t coord = valid ? a.coords[e] : 0;
        const AT v = valid ? (AT)vals[e] : (AT)0;
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

// --------------------------------------------------- partial reduction

template <typename AT>
__global__ __launch_bounds__(WG) void k_reduce_partials(AT* __restrict__ out,
                                                        const AT* __restrict__ part,
                                                        uint64_t cells, int nprtn)
{
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < cells;
         e += (uint64_t)gridDim.x * WG) {
        AT s = out[e];
        for (int p = 0; p < nprtn; ++p) s += part[(uint64_t)p * cells + e];
        out[e] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(WG) void k_fill(T* __restrict__ x, uint64_t n, T v)
{
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < n; e += (uint64_t)gridDim.x * WG)
        x[e] = v;
}

// -------------------------------------------------------------- dispatch

// Column layout (C columns per lane, L lanes per nonzero) for a rank and
// element size; EXACT when C * L == rank.
struct Shape {
    int C, L;
    bool exact;
};

static Shape pick_shape(int rank, int esize)
{
    switch (rank) {
        case 8: return {1, 8, true};
        case 16: return {2, 8, true};
        case 32: return esize == 8 ? Shape{2, 16, true} : Shape{4, 8, true};
        case 64: return esize == 8 ? Shape{2, 32, true} : esize == 4 ? Shape{4, 16, true} : Shape{8, 8, true};
        default: break;
    }
    int L = 8;
    while (L < rank && L < WAVE) L <<= 1;
    return {1, L, false};
}



template <int N, typename FT, typename VT, int C, int L, bool EX>
static const void* kernel_ptr(int which)
{
    if (which == W_LVL1) return (const void*)&k_mttkrp_lvl1<N, FT, VT, C, L, EX>;
    if (which == W_LVL3) return (const void*)&k_mttkrp_lvl3<N, FT, VT, C, L, EX>;
    if (which == W_WINDOW) return (const void*)&k_mttkrp_window<N, FT, VT, C, L, EX>;
    return (const void*)&k_mttkrp_atomic<N, FT, VT, C, L, EX>;
}

template <int N, typename FT, typename VT>
static const void* select_shape(const Shape& s, int which)
{
    if (s.exact) {
        if (s.C == 1 && s.L == 8) return kernel_ptr<N, FT, VT, 1, 8, true>(which);
        if (s.C == 2 && s.L == 8) return kernel_ptr<N, FT, VT, 2, 8, true>(which);
        if constexpr (sizeof(FT) == 8) {
            if (s.C == 2 && s.L == 16) return kernel_ptr<N, FT, VT, 2, 16, true>(which);
            if (s.C == 2 && s.L == 32) return kernel_ptr<N, FT, VT, 2, 32, true>(which);
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
static const void* select_n(int N, const Shape& s, int which)
{
    switch (N) {
        case 2: return select_shape<2, FT, VT>(s, which);
        case 3: return select_shape<3, FT, VT>(s, which);
        case 4: return select_shape<4, FT, VT>(s, which);
        case 5: return select_shape<5, FT, VT>(s, which);
        default: return nullptr;
    }
}

const void* mttkrp_kernel(int which, int nmodes, int ftype, int rank, int* C, int* L)
{
    Shape s = pick_shape(rank, dtype_size(ftype));
    *C = s.C;
    *L = s.L;
    if (ftype == BLCO_F64) return select_n<double, double>(nmodes, s, which);
    if (ftype == BLCO_F32) return select_n<float, float>(nmodes, s, which);
    return select_n<bf16, float>(nmodes, s, which);
}

int launch_mttkrp(int which, const MttkrpArgs& a, int ftype, uint32_t grid, size_t dyn_lds,
                  hipStream_t st)
{
    int C, L;
    const void* k = mttkrp_kernel(which, a.nmodes, ftype, a.rank, &C, &L);
    if (!k) return fail(BLCO_EUNSUPPORTED, "no MTTKRP kernel for %d modes, rank %d", a.nmodes, a.rank);
    MttkrpArgs args = a;
    void* params[] = {&args};
    HIP_TRY(hipLaunchKernel(k, dim3(grid), dim3(WG), params, dyn_lds, st));
    return BLCO_OK;
}

int launch_reduce_partials(void* out, const void* part, uint64_t cells, int nprtn, int otype,
                           hipStream_t st)
{
    uint32_t grid = (uint32_t)std::min<uint64_t>((cells + WG - 1) / WG, 4096);
    if (grid == 0) return BLCO_OK;
    if (otype == BLCO_F64)
        hipLaunchKernelGGL(k_reduce_partials<double>, dim3(grid), dim3(WG), 0, st, (double*)out,
                           (const double*)part, cells, nprtn);
    else
        hipLaunchKernelGGL(k_reduce_partials<float>, dim3(grid), dim3(WG), 0, st, (float*)out,
                           (const float*)part, cells, nprtn);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_fill(void* x, uint64_t n, int dtype, double v, hipStream_t st)
{
    uint32_t grid = (uint32_t)std::min<uint64_t>((n + WG - 1) / WG, 8192);
    if (grid == 0) return BLCO_OK;
    if (dtype == BLCO_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(grid), dim3(WG), 0, st, (double*)x, n, v);
    else if (dtype == BLCO_F32)
        hipLaunchKernelGGL(k_fill<float>, dim3(grid), dim3(WG), 0, st, (float*)x, n, (float)v);
    else
        hipLaunchKernelGGL(k_fill<uint16_t>, dim3(grid), dim3(WG), 0, st, (uint16_t*)x, n,
                           (uint16_t)0);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

}  // namespace blco
