// encode_dev.hip -- BLCO construction on the GPU.
//
// The same tensor as the host encoder (gen_blcotensor_host,
// include/blco.hpp:196-331, over the ALTO order of alto.hpp:107-186,697-754),
// bit for bit: ALTO keys are formed by depositing every index bit at its
// interleaved position (pdep over the 128-bit mode masks), sorted stably
// with a device LSD radix sort (low 64 bits first, then the upper half),
// re-encoded to BLCO words by extracting each mode's low-half bits (pext)
// to its packed field, and the values follow the permutation.  Blocks are
// runs of equal upper key half, built on the host from the sorted upper
// halves exactly as the host encoder builds them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <new>
#include <vector>

#include "compact.hpp"
#include "devcommon.hpp"
#include "encode_dev.hpp"
#include "../host/internal.hpp"

namespace blco {

KeyLayout key_layout(const Layout& L)
{
    KeyLayout k;
    std::memset(&k, 0, sizeof(k));
    k.nmodes = L.nmodes;
    for (int n = 0; n < L.nmodes; ++n) {
        k.nbits[n] = L.bits[n];
        k.trunc[n] = L.trunc[n];
        k.pos[n] = L.pos[n];
        k.dims[n] = L.dims[n];
        int j = 0;
        for (int p = 0; p < 128 && j < 64; ++p) {
            const uint64_t word = p < 64 ? L.alto_lo[n] : L.alto_hi[n];
            if ((word >> (p & 63)) & 1u) k.at[n][j++] = (uint8_t)p;
        }
    }
    return k;
}

namespace {

__global__ __launch_bounds__(WG) void k_alto_keys(KeyLayout k, uint64_t nnz, const uint64_t* __restrict__ cidx,
                                                  uint64_t* __restrict__ lo, uint64_t* __restrict__ hi,
                                                  uint64_t* __restrict__ idx, int* __restrict__ bad)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * WG) {
        uint64_t x[BLCO_MAX_MODES];
        bool out = false;
        for (int n = 0; n < k.nmodes; ++n) {
            x[n] = cidx[(uint64_t)n * nnz + i];
            out |= x[n] >= k.dims[n];
        }
        uint64_t l, h;
        alto_key(k, x, &l, &h);
        lo[i] = l;
        if (hi) hi[i] = h;
        idx[i] = i;
        if (out) atomicOr(bad, 1);
    }
}

// hi_g[j] = hi[idx[j]]: the upper halves in the order of the low-half sort.
__global__ __launch_bounds__(WG) void k_gather_u64(uint64_t n, const uint64_t* __restrict__ src,
                                                   const uint64_t* __restrict__ idx, uint64_t* __restrict__ dst,
                                                   uint64_t* __restrict__ iota)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) {
        dst[i] = src[idx[i]];
        if (iota) iota[i] = i;
    }
}

// BLCO words and values of the sorted nonzeros; the sorted nonzero i is
// lo[q] / input nonzero perm[q] with q = order ? order[i] : i.
__global__ __launch_bounds__(WG) void k_reencode(KeyLayout k, uint64_t nnz, const uint64_t* __restrict__ lo,
                                                 const uint64_t* __restrict__ perm,
                                                 const uint64_t* __restrict__ order,
                                                 const double* __restrict__ vals, uint64_t* __restrict__ coords,
                                                 double* __restrict__ vout)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * WG) {
        const uint64_t q = order ? order[i] : i;
        const uint64_t key = lo[q];
        uint64_t c = 0;
        for (int n = 0; n < k.nmodes; ++n)
            for (int j = 0; j < k.trunc[n]; ++j) c |= ((key >> k.at[n][j]) & 1u) << (k.pos[n] + j);
        coords[i] = c;
        vout[i] = vals[perm[q]];
    }
}

// Run starts of the sorted upper key halves (a block boundary candidate).
struct HiRunStart {
    const uint64_t* hi;
    __device__ bool operator()(uint64_t i) const { return i == 0 || hi[i] != hi[i - 1]; }
};
struct HiRunEmit {
    const uint64_t* hi;
    uint64_t* start;
    uint64_t* val;
    __device__ void operator()(uint64_t i, uint64_t p) const
    {
        start[p] = i;
        val[p] = hi[i];
    }
};

uint32_t grid_for(uint64_t n)
{
    const uint64_t g = (n + WG - 1) / WG;
    return (uint32_t)(g < 8192 ? (g ? g : 1) : 8192);
}

}  // namespace

int encode_sorted_device(const Layout& L, uint64_t nnz, uint64_t* lo, uint64_t* hi, uint64_t* idx, double* vals,
                         uint64_t mbs, blco_host* h, hipStream_t st)
{
    if (nnz == 0) return BLCO_OK;
    const KeyLayout k = key_layout(L);
    const bool wide = L.total_bits > 64;
    DevBufs B;
    int rc;
    uint64_t *lo2, *idx2, *hi2 = nullptr, *ord = nullptr, *ord2 = nullptr;
    if ((rc = B.alloc(&lo2, nnz)) || (rc = B.alloc(&idx2, nnz))) return rc;
    if (wide && ((rc = B.alloc(&hi2, nnz)) || (rc = B.alloc(&ord, nnz)) || (rc = B.alloc(&ord2, nnz)))) return rc;
    // stable LSD radix sort: the low key half, then (wide keys) the upper half
    size_t tmp_bytes = 0, t2 = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, lo, lo2, idx, idx2, nnz, 0,
                                               wide ? 64 : L.total_bits, st));
    if (wide)
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, hi2, hi, ord, ord2, nnz, 0, L.total_bits - 64, st));
    if (t2 > tmp_bytes) tmp_bytes = t2;
    unsigned char* tmp;
    if ((rc = B.alloc(&tmp, tmp_bytes ? tmp_bytes : 1))) return rc;
    if (L.total_bits == 0) {  // every mode has length 1: nothing to sort
        HIP_TRY(hipMemcpyAsync(lo2, lo, nnz * 8, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(idx2, idx, nnz * 8, hipMemcpyDeviceToDevice, st));
    } else {
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, lo, lo2, idx, idx2, nnz, 0,
                                                   wide ? 64 : L.total_bits, st));
    }
    const uint64_t* order = nullptr;
    if (wide) {
        hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(nnz)), dim3(WG), 0, st, nnz, hi, idx2, hi2, ord);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, hi2, hi, ord, ord2, nnz, 0, L.total_bits - 64,
                                                   st));
        order = ord2;  // sorted nonzero i is lo2[ord2[i]], input nonzero idx2[ord2[i]]
    }
    progress("device encode: keys sorted");
    // the outputs take the place of the inputs: BLCO words into lo, values into idx
    double* vout = reinterpret_cast<double*>(idx);
    hipLaunchKernelGGL(k_reencode, dim3(grid_for(nnz)), dim3(WG), 0, st, k, nnz, lo2, idx2, order, vals, lo, vout);
    HIP_TRY(hipGetLastError());
    // blocks: runs of equal upper half (sorted in hi), cut at mbs nonzeros
    std::vector<uint64_t> starts(1, 0), his(1, 0);
    if (wide) {
        CompactScratch cs;
        uint64_t nruns = 0;
        rc = compact_count(nnz, HiRunStart{hi}, &nruns, &cs, st);
        uint64_t *d_start = nullptr, *d_val = nullptr;
        if (!rc) rc = B.alloc(&d_start, nruns);
        if (!rc) rc = B.alloc(&d_val, nruns);
        if (!rc) rc = compact_write(nnz, HiRunStart{hi}, HiRunEmit{hi, d_start, d_val}, &cs, st);
        cs.release();
        if (rc) return rc;
        starts.resize(nruns);
        his.resize(nruns);
        HIP_TRY(hipMemcpyAsync(starts.data(), d_start, nruns * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(his.data(), d_val, nruns * 8, hipMemcpyDeviceToHost, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    build_blocks_runs(L, nnz, starts.size(), starts.data(), his.data(), mbs, h);
    return BLCO_OK;
}

}  // namespace blco

using namespace blco;

extern "C" int blco_device_encode(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* cidx,
                                  const double* vals, uint64_t max_block_size, blco_host** out)
{
    if (!out || !dims || (nnz && (!cidx || !vals))) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BLCO_ENODEV, "no GPU visible");
    blco_host* h = new (std::nothrow) blco_host();
    if (!h) return fail(BLCO_ENOMEM, "out of host memory");
    auto done = [&](int rc) {
        if (rc) delete h;
        else *out = h;
        return rc;
    };
    int rc = make_layout(nmodes, dims, &h->L);
    if (rc) return done(rc);
    const Layout& L = h->L;
    h->nnz = nnz;
    h->max_block_size = max_block_size;
    const uint64_t mbs = max_block_size ? max_block_size : (nnz ? nnz : 1);  // blco.hpp:264
    if (nnz == 0) return done(BLCO_OK);
    try {
        h->coords.resize(nnz);
        h->vals.resize(nnz);
    } catch (const std::bad_alloc&) {
        return done(fail(BLCO_ENOMEM, "host allocation for %llu nonzeros failed", (unsigned long long)nnz));
    }
    const KeyLayout k = key_layout(L);
    const bool wide = L.total_bits > 64;
    DevBufs B;
    uint64_t *d_cidx, *lo, *idx, *hi = nullptr;
    double* d_vals;
    int* d_bad;
    if ((rc = B.alloc(&d_cidx, (uint64_t)nmodes * nnz)) || (rc = B.alloc(&lo, nnz)) || (rc = B.alloc(&idx, nnz)) ||
        (rc = B.alloc(&d_vals, nnz)) || (rc = B.alloc(&d_bad, 1)))
        return done(rc);
    if (wide && (rc = B.alloc(&hi, nnz))) return done(rc);
    if (hipMemcpy(d_cidx, cidx, (uint64_t)nmodes * nnz * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_vals, vals, nnz * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d_bad, 0, sizeof(int)) != hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: upload failed"));
    progress("device encode: %llu nonzeros uploaded", (unsigned long long)nnz);
    k_alto_keys<<<grid_for(nnz), WG>>>(k, nnz, d_cidx, lo, hi, idx, d_bad);
    int bad = 0;
    if (hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: key formation failed: %s", hipGetErrorString(hipGetLastError())));
    if (bad) return done(fail(BLCO_EINVAL, "a coordinate is outside its mode length"));
    B.release(d_cidx);  // the sort needs the room
    rc = encode_sorted_device(L, nnz, lo, hi, idx, d_vals, mbs, h, nullptr);
    if (rc) return done(rc);
    if (hipMemcpy(h->coords.data(), lo, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h->vals.data(), idx, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: re-encoding failed: %s", hipGetErrorString(hipGetLastError())));
    progress("device encode: %zu block(s)", h->block_start.size());
    return done(BLCO_OK);
}
