// devcommon.hpp -- shared device helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../host/internal.hpp"
#include "blco_hip.h"

namespace blco {

constexpr int WAVE = 64;
constexpr int WG = 256;            // threads per workgroup (4 waves)
constexpr int WAVES_PER_WG = WG / WAVE;
constexpr int NUM_XCD = 8;

// Launch check: a failing hip call returns BLCO_EHIP with its message.
#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess)                                                  \
            return ::blco::fail(BLCO_EHIP, "%s: %s (%s:%d)", #expr,            \
                                hipGetErrorString(e_), __FILE__, __LINE__);    \
    } while (0)

// Storage for bf16 factors: the upper 16 bits of an IEEE float.
struct bf16 {
    uint16_t bits;
};

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

template <typename T> struct Acc { using type = float; };
template <> struct Acc<double> { using type = double; };

template <typename T> __device__ __forceinline__ typename Acc<T>::type to_acc(T x) { return x; }
template <> __device__ __forceinline__ float to_acc<bf16>(bf16 x) { return bf16_to_f32(x.bits); }

template <typename T> __device__ __forceinline__ T from_f64(double x) { return (T)x; }
template <> __device__ __forceinline__ bf16 from_f64<bf16>(double x) { return bf16{f32_to_bf16((float)x)}; }

// Logical workgroup id so that consecutive logical ids sit on one XCD:
// hardware dispatches workgroup b to XCD b % 8, so XCD x receives the
// contiguous logical range [start(x), start(x) + count(x)).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G)
{
    uint32_t x = b % NUM_XCD, local = b / NUM_XCD;
    uint32_t q = G / NUM_XCD, r = G % NUM_XCD;
    return x * q + (x < r ? x : r) + local;
}

inline int dtype_size(int dt) { return dt == BLCO_F64 ? 8 : dt == BLCO_F32 ? 4 : 2; }

}  // namespace blco
