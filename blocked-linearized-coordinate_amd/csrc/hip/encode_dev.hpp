// encode_dev.hpp -- pieces of the GPU BLCO encoder shared with the GPU
// generator: the ALTO key layout, key formation, and the sort + re-encode
// core that works on device-resident keys.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "devcommon.hpp"
#include "../host/internal.hpp"

namespace blco {

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

KeyLayout key_layout(const Layout& L);

// ALTO key (pdep of every index over its interleave mask) of one nonzero.
__device__ __forceinline__ void alto_key(const KeyLayout& k, const uint64_t* x, uint64_t* lo, uint64_t* hi)
{
    uint64_t l = 0, h = 0;
    for (int n = 0; n < k.nmodes; ++n)
        for (int j = 0; j < k.nbits[n]; ++j) {
            const uint64_t b = (x[n] >> j) & 1u;
            const int p = k.at[n][j];
            if (p < 64) l |= b << p;
            else h |= b << (p - 64);
        }
    *lo = l;
    *hi = h;
}

// Device buffers released on every exit path.
struct DevBufs {
    std::vector<void*> p;
    DevBufs() = default;
    DevBufs(const DevBufs&) = delete;
    DevBufs& operator=(const DevBufs&) = delete;
    ~DevBufs()
    {
        for (void* x : p)
            if (x) (void)hipFree(x);
    }
    template <typename T>
    int alloc(T** out, uint64_t count)
    {
        *out = nullptr;
        if (hipMalloc((void**)out, (count ? count : 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return fail(BLCO_ENOMEM, "device allocation of %llu bytes failed",
                        (unsigned long long)(count * sizeof(T)));
        }
        p.push_back(*out);
        return BLCO_OK;
    }
    // Frees one buffer early (it is no longer released at exit).
    void release(void* x)
    {
        for (void*& q : p)
            if (q == x) {
                (void)hipFree(q);
                q = nullptr;
            }
    }
};

// Sorts nnz nonzeros by their ALTO keys (device lo, and hi for keys wider
// than 64 bits, else nullptr; idx = 0, 1, ..., nnz - 1) stably and
// re-encodes them: on return lo holds the BLCO words and idx (as doubles)
// the values of `vals` (f64, input order), both in BLCO order, and the
// blocks (runs of equal upper half cut at mbs nonzeros) are appended to h.
// hi and vals are left as scratch.  Synchronises `st`.
int encode_sorted_device(const Layout& L, uint64_t nnz, uint64_t* lo, uint64_t* hi, uint64_t* idx, double* vals,
                         uint64_t mbs, blco_host* h, hipStream_t st);

}  // namespace blco
