// mttkrp.hip -- MTTKRP on BLCO tensors for gfx950: kernel selection, launch
// and the reductions.  The kernel templates live in mttkrp_kern.hpp and are
// instantiated per precision in mttkrp_{f64,f32,bf16}.hip.
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
#include "mttkrp_kern.hpp"

namespace blco {

// out[row(u)][c] += sum over j in [ptr[u], ptr[u+1]) of part[idx[j]][c]:
// one lane per (segment, column); a segment holds at most a few partial rows,
// so a hot output row split over many segments is summed in parallel and
// added atomically.
template <typename AT>
__global__ __launch_bounds__(WG) void k_window_reduce(AT* __restrict__ out, const AT* __restrict__ part,
                                                      const uint64_t* __restrict__ rows,
                                                      const uint64_t* __restrict__ ptr,
                                                      const uint32_t* __restrict__ idx,
                                                      uint64_t nseg, int R)
{
    const uint64_t cells = nseg * (uint64_t)R;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * WG) {
        const uint64_t u = i / R, c = i % R;
        const uint64_t j1 = ptr[u + 1];
        AT s0 = 0, s1 = 0;
        uint64_t j = ptr[u];
        for (; j + 1 < j1; j += 2) {
            s0 += stream_load(part + (uint64_t)idx[j] * R + c);
            s1 += stream_load(part + (uint64_t)idx[j + 1] * R + c);
        }
        if (j < j1) s0 += stream_load(part + (uint64_t)idx[j] * R + c);
        const uint64_t rr = rows[u];
        AT* o = out + (rr & ~REDUCE_SHARED) * R + c;
        if (rr & REDUCE_SHARED) atomic_add_global(o, s0 + s1);
        else *o += s0 + s1;
    }
}

int launch_window_reduce(void* out, const void* part, const uint64_t* rows, const uint64_t* ptr,
                         const uint32_t* idx, uint64_t nrows, int rank, int otype, hipStream_t st)
{
    const uint64_t cells = nrows * (uint64_t)rank;
    if (!cells) return BLCO_OK;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((cells + WG - 1) / WG, 1u << 16);
    hipError_t e = hipGetLastError();  // a launch before this one
    if (e != hipSuccess)
        return fail(BLCO_EHIP, "error pending before the window reduce: %s", hipGetErrorString(e));
    if (otype == BLCO_F64)
        hipLaunchKernelGGL(k_window_reduce<double>, dim3(grid), dim3(WG), 0, st, (double*)out,
                           (const double*)part, rows, ptr, idx, nrows, rank);
    else
        hipLaunchKernelGGL(k_window_reduce<float>, dim3(grid), dim3(WG), 0, st, (float*)out,
                           (const float*)part, rows, ptr, idx, nrows, rank);
    e = hipGetLastError();
    if (e != hipSuccess)
        return fail(BLCO_EHIP, "window reduce launch (grid %u, %llu segments, rank %d): %s", grid,
                    (unsigned long long)nrows, rank, hipGetErrorString(e));
    return BLCO_OK;
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


const void* mttkrp_kernel(int which, int nmodes, int ftype, int rank, int* C, int* L)
{
    Shape s = pick_shape(rank, dtype_size(ftype));
    *C = s.C;
    *L = s.L;
    if (ftype == BLCO_F64) return mttkrp_kernel_f64(nmodes, s, which);
    if (ftype == BLCO_F32) return mttkrp_kernel_f32(nmodes, s, which);
    return mttkrp_kernel_bf16(nmodes, s, which);
}

int launch_mttkrp(int which, const MttkrpArgs& a, int ftype, uint32_t grid, size_t dyn_lds,
                  hipStream_t st)
{
    int C, L;
    const void* k = mttkrp_kernel(which, a.nmodes, ftype, a.rank, &C, &L);
    if (!k) return fail(BLCO_EUNSUPPORTED, "no MTTKRP kernel for %d modes, rank %d", a.nmodes, a.rank);
    MttkrpArgs args = a;
    void* params[] = {&args};
    const int block = (which == W_WINDOW || which == W_WINDOW64) ? WIN_WAVES * WAVE : WG;
    if (dyn_lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn_lds));
    hipError_t e = hipLaunchKernel(k, dim3(grid), dim3(block), params, dyn_lds, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess)
        return fail(BLCO_EHIP, "MTTKRP launch (kind %d, grid %u x %d, %zu bytes LDS, rank %d): %s", which,
                    grid, block, dyn_lds, a.rank, hipGetErrorString(e));
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
