// roworder_dev.hip -- the row-ordered copies of kernel 14 built on the GPU
// from the resident tensor, in the order the host builds them
// (roworder.cpp): nonzeros grouped by (tile of the split mode, block), by
// the output mode's index within a group and, for fiber order, by the outer
// modes' local indices within a row.  One stable radix sort of a composite
// key (group, row, outer key) gives the host's permutation: the host sorts
// stably by the outer key first and then by (group, row), and equal
// composite keys keep their input order in both.  The run starts of
// (group, row), the groups and the fiber count come from the sorted keys.
// Also here: the counts behind the tiling decision (nonzeros per split-mode
// index, and the tiles each output row meets).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "compact.hpp"
#include "devcommon.hpp"
#include "device.hpp"
#include "encode_dev.hpp"

namespace blco {
namespace {

struct Decode {
    const uint64_t* coords;
    const uint64_t* bstart;  // device, nb entries
    const uint64_t* bcoords; // nb x N
    uint64_t nb;
    int N;
    // block of nonzero i: the last b with bstart[b] <= i
    __device__ __forceinline__ uint64_t block(uint64_t i) const
    {
        uint64_t lo = 0, hi = nb;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (bstart[mid] <= i) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    __device__ __forceinline__ uint64_t index(uint64_t i, uint64_t b, uint64_t mask, int pos, int m) const
    {
        return ((coords[i] >> pos) & mask) | bcoords[b * N + m];
    }
};

uint32_t grid_for(uint64_t n)
{
    const uint64_t g = (n + WG - 1) / WG;
    return (uint32_t)(g < 65536 ? (g ? g : 1) : 65536);
}

// The lowest LDS_ROWS indices are counted in LDS first and added to the
// global counters once per workgroup: skewed (Zipf) modes put their hot rows
// at the low indices, whose global atomics would otherwise serialise.
constexpr uint32_t LDS_ROWS = 8192;
constexpr uint32_t HIST_GRID = 2048;

__global__ __launch_bounds__(WG) void k_split_hist(uint64_t n, Decode dc, int split, uint64_t mask, int pos,
                                                   unsigned long long* __restrict__ cnt)
{
    __shared__ unsigned int h[LDS_ROWS];
    for (uint32_t t = threadIdx.x; t < LDS_ROWS; t += WG) h[t] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = dc.block(i);
        const uint64_t x = dc.index(i, b, mask, pos, split);
        if (x < LDS_ROWS) atomicAdd(&h[x], 1u);
        else atomicAdd(&cnt[x], 1ull);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < LDS_ROWS; t += WG)
        if (h[t]) atomicAdd(&cnt[t], (unsigned long long)h[t]);
}

__global__ __launch_bounds__(WG) void k_row_tiles(uint64_t n, Decode dc, int mode, uint64_t rmask, int rpos, int split,
                                                  uint64_t smask, int spos, const uint8_t* __restrict__ tile_of,
                                                  unsigned int* __restrict__ met)
{
    __shared__ unsigned int h[LDS_ROWS];
    for (uint32_t t = threadIdx.x; t < LDS_ROWS; t += WG) h[t] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = dc.block(i);
        const uint64_t r = dc.index(i, b, rmask, rpos, mode);
        const unsigned int bit = 1u << tile_of[dc.index(i, b, smask, spos, split)];
        if (r < LDS_ROWS) atomicOr(&h[r], bit);
        else atomicOr(&met[r], bit);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < LDS_ROWS; t += WG)
        if (h[t]) atomicOr(&met[t], h[t]);
}

struct KeySpec {
    int ntiles;
    int split;
    uint64_t smask;
    int spos;
    const uint8_t* tile_of;
    uint64_t rmask;  // output mode's field (local row)
    int rpos;
    int row_bits;
    int nouter;
    uint64_t omask[BLCO_MAX_MODES];
    int opos[BLCO_MAX_MODES];
    int owidth[BLCO_MAX_MODES];
    int ob;  // outer key bits
};

// key = (group << (row_bits + ob)) | (row << ob) | outer key, group = tile * nb + block
template <typename PT>
__global__ __launch_bounds__(WG) void k_row_keys(uint64_t n, Decode dc, KeySpec ks, uint64_t* __restrict__ key,
                                                 PT* __restrict__ perm)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        const uint64_t b = dc.block(i);
        const uint64_t c = dc.coords[i];
        uint64_t g = b;
        if (ks.ntiles > 1) g = (uint64_t)ks.tile_of[dc.index(i, b, ks.smask, ks.spos, ks.split)] * dc.nb + b;
        uint64_t ok = 0;
        for (int j = 0; j < ks.nouter; ++j) {
            const int w = ks.owidth[j];
            ok = (w < 64 ? ok << w : 0) | ((c >> ks.opos[j]) & ks.omask[j]);
        }
        const uint64_t row = (c >> ks.rpos) & ks.rmask;
        const int low = ks.row_bits + ks.ob;
        key[i] = (low < 64 ? g << low : 0) | (ks.ob < 64 ? row << ks.ob : 0) | ok;
        perm[i] = (PT)i;
    }
}

template <typename PT>
__global__ __launch_bounds__(WG) void k_gather_copy(uint64_t n, const PT* __restrict__ perm,
                                                    const uint64_t* __restrict__ coords, const void* __restrict__ vals,
                                                    int vsz, uint64_t* __restrict__ sc, void* __restrict__ sv)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        const uint64_t p = perm[i];
        sc[i] = coords[p];
        if (vsz == 8) static_cast<double*>(sv)[i] = static_cast<const double*>(vals)[p];
        else static_cast<float*>(sv)[i] = static_cast<const float*>(vals)[p];
    }
}

// a run of equal (group, row) starts at i
struct RunStart {
    const uint64_t* key;
    int ob;
    __device__ bool operator()(uint64_t i) const
    {
        return i == 0 || (ob < 64 ? (key[i] >> ob) != (key[i - 1] >> ob) : false);
    }
};
struct RunEmit {
    const uint64_t* key;
    int gshift;  // row_bits + ob
    uint64_t* start;
    uint64_t* group;
    __device__ void operator()(uint64_t i, uint64_t p) const
    {
        start[p] = i;
        group[p] = gshift < 64 ? key[i] >> gshift : 0;
    }
};

// fibers: runs of equal full key
__global__ __launch_bounds__(WG) void k_count_changes(uint64_t n, const uint64_t* __restrict__ key,
                                                      unsigned long long* __restrict__ total)
{
    unsigned long long t = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        t += (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += (unsigned long long)__shfl_xor((long long)t, o);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(total, t);
}

int upload_bstart(const blco_dev* d, DevBufs* B, uint64_t** out)
{
    int rc = B->alloc(out, d->nblocks);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*out, d->bstart.data(), d->nblocks * 8, hipMemcpyHostToDevice));
    return BLCO_OK;
}

int width(uint64_t m) { return m ? 64 - __builtin_clzll(m) : 0; }

}  // namespace

int dev_split_counts(const blco_dev* d, int split, std::vector<uint64_t>* cnt)
{
    const uint64_t I = d->dims[split];
    DevBufs B;
    uint64_t* bs;
    unsigned long long* c;
    int rc;
    if ((rc = upload_bstart(d, &B, &bs)) || (rc = B.alloc(&c, I))) return rc;
    HIP_TRY(hipMemset(c, 0, I * 8));
    const Decode dc{d->d_coords, bs, d->d_bcoords, d->nblocks, d->nmodes};
    hipLaunchKernelGGL(k_split_hist, dim3(std::min(grid_for(d->nnz), HIST_GRID)), dim3(WG), 0, 0, d->nnz, dc, split,
                       d->masks[split], d->pos[split], c);
    HIP_TRY(hipGetLastError());
    cnt->resize(I);
    HIP_TRY(hipMemcpy(cnt->data(), c, I * 8, hipMemcpyDeviceToHost));
    return BLCO_OK;
}

int dev_row_tiles(const blco_dev* d, int mode, int split, const std::vector<uint8_t>& tile_of,
                  std::vector<uint8_t>* met)
{
    const uint64_t rows = d->dims[mode];
    DevBufs B;
    uint64_t* bs;
    uint8_t* to;
    unsigned int* m;
    int rc;
    if ((rc = upload_bstart(d, &B, &bs)) || (rc = B.alloc(&to, tile_of.size())) || (rc = B.alloc(&m, rows)))
        return rc;
    HIP_TRY(hipMemcpy(to, tile_of.data(), tile_of.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(m, 0, rows * 4));
    const Decode dc{d->d_coords, bs, d->d_bcoords, d->nblocks, d->nmodes};
    hipLaunchKernelGGL(k_row_tiles, dim3(std::min(grid_for(d->nnz), HIST_GRID)), dim3(WG), 0, 0, d->nnz, dc, mode,
                       d->masks[mode], d->pos[mode], split, d->masks[split], d->pos[split], to, m);
    HIP_TRY(hipGetLastError());
    std::vector<unsigned int> h(rows);
    HIP_TRY(hipMemcpy(h.data(), m, rows * 4, hipMemcpyDeviceToHost));
    met->resize(rows);
    for (uint64_t r = 0; r < rows; ++r) (*met)[r] = (uint8_t)h[r];
    return BLCO_OK;
}

int dev_row_copy(const blco_dev* d, const RowOrderSpec& o, std::vector<SGroup>* groups, std::vector<uint64_t>* runs,
                 uint64_t* fibers, uint64_t** sc_out, void** sv_out)
{
    *sc_out = nullptr;
    *sv_out = nullptr;
    const uint64_t n = d->nnz, nb = d->nblocks;
    const int ntiles = o.ntiles > 1 ? o.ntiles : 1;
    const uint64_t ngroups = (uint64_t)ntiles * nb;
    const int grp_bits = ngroups > 1 ? 64 - __builtin_clzll(ngroups - 1) : 0;
    KeySpec ks;
    std::memset(&ks, 0, sizeof(ks));
    ks.ntiles = ntiles;
    ks.split = o.split;
    if (ntiles > 1) {
        ks.smask = d->masks[o.split];
        ks.spos = d->pos[o.split];
    }
    ks.rmask = d->masks[o.mode];
    ks.rpos = d->pos[o.mode];
    ks.row_bits = width(ks.rmask);
    ks.nouter = o.nouter;
    for (int j = 0; j < o.nouter; ++j) {
        ks.omask[j] = d->masks[o.outer[j]];
        ks.opos[j] = d->pos[o.outer[j]];
        ks.owidth[j] = width(ks.omask[j]);
        ks.ob += ks.owidth[j];
    }
    const int bits = grp_bits + ks.row_bits + ks.ob;
    if (bits > 64) return fail(BLCO_EUNSUPPORTED, "row-ordered key of %d bits", bits);
    DevBufs B;
    uint64_t *bs, *k1, *k2;
    uint8_t* to = nullptr;
    int rc;
    if ((rc = upload_bstart(d, &B, &bs)) || (rc = B.alloc(&k1, n)) || (rc = B.alloc(&k2, n))) return rc;
    if (ntiles > 1) {
        if ((rc = B.alloc(&to, d->dims[o.split]))) return rc;
        HIP_TRY(hipMemcpy(to, o.tile_of, d->dims[o.split], hipMemcpyHostToDevice));
        ks.tile_of = to;
    }
    const Decode dc{d->d_coords, bs, d->d_bcoords, nb, d->nmodes};
    const int vsz = d->vdtype == BLCO_F64 ? 8 : 4;
    uint64_t* sc = nullptr;
    void* sv = nullptr;
    auto sort_and_gather = [&](auto* p1, auto* p2) -> int {
        using PT = std::remove_pointer_t<decltype(p1)>;
        hipLaunchKernelGGL((k_row_keys<PT>), dim3(grid_for(n)), dim3(WG), 0, 0, n, dc, ks, k1, p1);
        HIP_TRY(hipGetLastError());
        size_t tmp_bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k1, k2, p1, p2, n, 0, bits > 0 ? bits : 1));
        unsigned char* tmp;
        int r2 = B.alloc(&tmp, tmp_bytes ? tmp_bytes : 1);
        if (r2) return r2;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k1, k2, p1, p2, n, 0, bits > 0 ? bits : 1));
        B.release(tmp);
        B.release(k1);
        k1 = nullptr;
        if (hipMalloc(&sc, n * 8) != hipSuccess || hipMalloc(&sv, n * vsz) != hipSuccess)
            return fail(BLCO_ENOMEM, "no device memory for the row-ordered copy of mode %d (%llu bytes)", o.mode,
                        (unsigned long long)(n * (8 + vsz)));
        hipLaunchKernelGGL((k_gather_copy<PT>), dim3(grid_for(n)), dim3(WG), 0, 0, n, p2, d->d_coords, d->d_vals,
                           vsz, sc, sv);
        HIP_TRY(hipGetLastError());
        return BLCO_OK;
    };
    if (n < (1ull << 32)) {
        uint32_t *p1, *p2;
        if ((rc = B.alloc(&p1, n)) || (rc = B.alloc(&p2, n))) return rc;
        rc = sort_and_gather(p1, p2);
        B.release(p1);
        B.release(p2);
    } else {
        uint64_t *p1, *p2;
        if ((rc = B.alloc(&p1, n)) || (rc = B.alloc(&p2, n))) return rc;
        rc = sort_and_gather(p1, p2);
        B.release(p1);
        B.release(p2);
    }
    auto drop = [&](int r) {
        if (sc) (void)hipFree(sc);
        if (sv) (void)hipFree(sv);
        return r;
    };
    if (rc) return drop(rc);
    // run starts of (group, row) and the group of each run
    CompactScratch cs;
    uint64_t nruns = 0;
    uint64_t *rs = nullptr, *rg = nullptr;
    rc = compact_count(n, RunStart{k2, ks.ob}, &nruns, &cs, 0);
    if (!rc) rc = B.alloc(&rs, nruns);
    if (!rc) rc = B.alloc(&rg, nruns);
    if (!rc) rc = compact_write(n, RunStart{k2, ks.ob}, RunEmit{k2, ks.row_bits + ks.ob, rs, rg}, &cs, 0);
    cs.release();
    if (rc) return drop(rc);
    std::vector<uint64_t> h_rs(nruns), h_rg(nruns);
    if (hipMemcpy(h_rs.data(), rs, nruns * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h_rg.data(), rg, nruns * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return drop(fail(BLCO_EHIP, "download of the row runs failed"));
    if (fibers) {
        unsigned long long* t;
        if ((rc = B.alloc(&t, 1))) return drop(rc);
        HIP_TRY(hipMemset(t, 0, 8));
        hipLaunchKernelGGL(k_count_changes, dim3(grid_for(n)), dim3(WG), 0, 0, n, k2, t);
        unsigned long long h = 0;
        if (hipGetLastError() != hipSuccess || hipMemcpy(&h, t, 8, hipMemcpyDeviceToHost) != hipSuccess)
            return drop(fail(BLCO_EHIP, "fiber count failed"));
        *fibers = h;
    }
    // groups in storage order: a group starts where the run's group changes
    groups->clear();
    for (uint64_t r = 0; r < nruns; ++r) {
        if (r > 0 && h_rg[r] == h_rg[r - 1]) continue;
        const uint64_t g = h_rg[r];
        uint64_t end = n;
        for (uint64_t q = r + 1; q < nruns; ++q)
            if (h_rg[q] != g) {
                end = h_rs[q];
                break;
            }
        groups->push_back(SGroup{h_rs[r], end - h_rs[r], (uint32_t)(g % nb), (uint32_t)(g / nb)});
    }
    runs->swap(h_rs);
    runs->push_back(n);
    *sc_out = sc;
    *sv_out = sv;
    return BLCO_OK;
}

}  // namespace blco
