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

#include "devcommon.hpp"
#include "../host/internal.hpp"

namespace blco {
namespace {

// Bit positions of each mode's index bits in the 128-bit ALTO key (bit j of
// the index goes to position at[n][j]), and of its BLCO field.
struct KeyLayout {
    int nmodes;
    int nbits[BLCO_MAX_MODES];  // index bits of the mode
    int trunc[BLCO_MAX_MODES];  // of which in the low key half
    int pos[BLCO_MAX_MODES];    // BLCO field offset
    uint64_t dims[BLCO_MAX_MODES];
    uint8_t at[BLCO_MAX_MODES][64];
};

__global__ __launch_bounds__(WG) void k_alto_keys(KeyLayout k, uint64_t nnz, const uint64_t* __restrict__ cidx,
                                                  uint64_t* __restrict__ lo, uint64_t* __restrict__ hi,
                                                  uint64_t* __restrict__ idx, int* __restrict__ bad)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * WG) {
        uint64_t l = 0, h = 0;
        bool out = false;
        for (int n = 0; n < k.nmodes; ++n) {
            const uint64_t x = cidx[(uint64_t)n * nnz + i];
            out |= x >= k.dims[n];
            for (int j = 0; j < k.nbits[n]; ++j) {
                const uint64_t b = (x >> j) & 1u;
                const int p = k.at[n][j];
                if (p < 64) l |= b << p;
                else h |= b << (p - 64);
            }
        }
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

uint32_t grid_for(uint64_t n)
{
    const uint64_t g = (n + WG - 1) / WG;
    return (uint32_t)(g < 8192 ? (g ? g : 1) : 8192);
}

// Device buffers released on every exit path.
struct DevBufs {
    std::vector<void*> p;
    ~DevBufs()
    {
        for (void* x : p)
            if (x) (void)hipFree(x);
    }
    template <typename T>
    int alloc(T** out, uint64_t count)
    {
        *out = nullptr;
        if (hipMalloc((void**)out, count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(BLCO_ENOMEM, "device encoder: allocation of %llu bytes failed",
                        (unsigned long long)(count * sizeof(T)));
        }
        p.push_back(*out);
        return BLCO_OK;
    }
};

}  // namespace
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
    // bit positions: the j-th set bit of mode n's 128-bit mask (low half first)
    KeyLayout k;
    std::memset(&k, 0, sizeof(k));
    k.nmodes = nmodes;
    for (int n = 0; n < nmodes; ++n) {
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
    const bool wide = L.total_bits > 64;
    DevBufs B;
    uint64_t *d_cidx, *lo, *lo2, *idx, *idx2, *hi = nullptr, *hi2 = nullptr, *ord = nullptr, *ord2 = nullptr;
    uint64_t* d_coords;
    double *d_vals, *d_vout;
    int* d_bad;
    if ((rc = B.alloc(&d_cidx, (uint64_t)nmodes * nnz)) || (rc = B.alloc(&lo, nnz)) || (rc = B.alloc(&lo2, nnz)) ||
        (rc = B.alloc(&idx, nnz)) || (rc = B.alloc(&idx2, nnz)) || (rc = B.alloc(&d_vals, nnz)) ||
        (rc = B.alloc(&d_bad, 1)))
        return done(rc);
    if (wide && ((rc = B.alloc(&hi, nnz)) || (rc = B.alloc(&hi2, nnz)) || (rc = B.alloc(&ord, nnz)) ||
                 (rc = B.alloc(&ord2, nnz))))
        return done(rc);
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
    // stable LSD radix sort: the low key half, then (wide keys) the upper half
    size_t tmp_bytes = 0, t2 = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, lo, lo2, idx, idx2, nnz, 0,
                                           wide ? 64 : L.total_bits) != hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: sort setup failed"));
    if (wide && hipcub::DeviceRadixSort::SortPairs(nullptr, t2, hi2, hi, ord, ord2, nnz, 0, L.total_bits - 64) !=
                    hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: sort setup failed"));
    if (t2 > tmp_bytes) tmp_bytes = t2;
    unsigned char* tmp;
    if ((rc = B.alloc(&tmp, tmp_bytes ? tmp_bytes : 1))) return done(rc);
    if (L.total_bits == 0) {  // every mode has length 1: nothing to sort
        if (hipMemcpy(lo2, lo, nnz * 8, hipMemcpyDeviceToDevice) != hipSuccess ||
            hipMemcpy(idx2, idx, nnz * 8, hipMemcpyDeviceToDevice) != hipSuccess)
            return done(fail(BLCO_EHIP, "device encoder: copy failed"));
    } else if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, lo, lo2, idx, idx2, nnz, 0,
                                                  wide ? 64 : L.total_bits) != hipSuccess) {
        return done(fail(BLCO_EHIP, "device encoder: sort failed"));
    }
    const uint64_t* order = nullptr;
    if (wide) {
        k_gather_u64<<<grid_for(nnz), WG>>>(nnz, hi, idx2, hi2, ord);
        if (hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, hi2, hi, ord, ord2, nnz, 0, L.total_bits - 64) !=
            hipSuccess)
            return done(fail(BLCO_EHIP, "device encoder: sort failed"));
        order = ord2;  // sorted nonzero i is lo2[ord2[i]], input nonzero idx2[ord2[i]]
    }
    progress("device encode: keys sorted");
    // outputs reuse the key buffers that are free now
    d_coords = lo;
    d_vout = reinterpret_cast<double*>(idx);
    k_reencode<<<grid_for(nnz), WG>>>(k, nnz, lo2, idx2, order, d_vals, d_coords, d_vout);
    if (hipMemcpy(h->coords.data(), d_coords, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(h->vals.data(), d_vout, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
        return done(fail(BLCO_EHIP, "device encoder: re-encoding failed: %s", hipGetErrorString(hipGetLastError())));
    std::vector<uint64_t> hi_h;
    if (wide) {
        try {
            hi_h.resize(nnz);
        } catch (const std::bad_alloc&) {
            return done(fail(BLCO_ENOMEM, "host allocation for %llu nonzeros failed", (unsigned long long)nnz));
        }
        if (hipMemcpy(hi_h.data(), hi, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess)
            return done(fail(BLCO_EHIP, "device encoder: download failed"));
    }
    build_blocks(L, nnz, wide ? hi_h.data() : nullptr, mbs, h);
    progress("device encode: %zu block(s)", h->block_start.size());
    return done(BLCO_OK);
}
