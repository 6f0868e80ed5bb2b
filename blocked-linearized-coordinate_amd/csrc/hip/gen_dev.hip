// gen_dev.hip -- the synthetic Poisson tensor generator on the GPU, feeding
// the GPU encoder: the tensor goes from the generator's spec to a resident
// BLCO device tensor without passing through host memory.
//
// It produces the tensor of the host generator (generate.cpp) for the same
// spec: the same counter-based draws (SplitMix64 keyed by seed, stream and
// counter), the same Zipf inversion over the host's own cdf and guide tables,
// duplicates merged keeping the first draw, the same top-up rounds (the first
// target_nnz unique draws in draw order), the same Exp(1) generating factors
// and the same zero-truncated Poisson sampler with the same double operations
// (this file, like generate.cpp, is compiled without floating-point
// contraction; the sampler's fused multiply-adds are explicit fma).  Counts
// below 65536 take log1p from a host-computed table.  The library functions
// exp, log and lgamma of the sampler are the device's own: where one of
// them differs from the host's by an ulp, a draw whose Poisson decision
// falls within that ulp could differ (none has in the tests, which compare
// the two generators bit for bit).
//
// Duplicates: each draw's linear key (per-mode bit fields, mode 0 lowest) is
// stored as its low 8 bits and the rest; the draws are split into 2^b
// partitions by the key's low b bits (b chosen from the measured histogram
// so that every key fits 64 bits and every partition fits the device), and
// each partition is compacted in draw order, radix-sorted stably by key,
// and the first draw of every run of equal keys is marked in a bitmap over
// all draws.  The first target_nnz marked draws are the nonzeros.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

#include "compact.hpp"
#include "devcommon.hpp"
#include "encode_dev.hpp"
#include "../host/internal.hpp"
#include "../host/rng.hpp"

namespace blco {

int create_device_tensor(int nmodes, const uint64_t* dims, const uint64_t* masks, const int* pos,
                         uint64_t nnz, const uint64_t* coords, const double* vals,
                         uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz,
                         const uint64_t* block_coords, int value_dtype, int stream_data,
                         uint64_t slot_nnz, uint64_t device_budget, uint64_t* adopt_coords, void* adopt_vals,
                         blco_dev** out);

namespace {

constexpr uint32_t LOG1P_TABLE = 1u << 16;

struct ModeSampler {
    const double* cdf;     // Zipf cdf (nullptr: uniform)
    const uint64_t* guide; // SAMPLER_GUIDE + 1 entries
    uint64_t I;
};

struct GenCtx {
    int N;
    int sh[BLCO_MAX_MODES];          // bit offset of each mode in the linear draw key
    uint64_t key_index[BLCO_MAX_MODES]; // stream keys of the index draws
    uint64_t key_value;                 // stream key of the value draws
    ModeSampler m[BLCO_MAX_MODES];
    int Rg;
    const double* A[BLCO_MAX_MODES];  // generating factors, dims[n] x Rg
    const double* log1p_tab;          // log1p(c), c < LOG1P_TABLE
    int value_kind;
};

__device__ __forceinline__ uint64_t d_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// rng_u01 (rng.hpp) with the stream's key precomputed on the host
__device__ __forceinline__ double d_u01(uint64_t key, uint64_t ctr)
{
    return (double)(d_mix64(key + ctr * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}

// IndexSampler::draw (generate.cpp): the first k with cdf[k] > u, searched in
// the guide interval [guide[g], min(guide[g + 1] + 1, I)).
__device__ __forceinline__ uint64_t d_draw(const ModeSampler& s, double u)
{
    if (s.cdf) {
        const uint32_t g = (uint32_t)(u * (double)SAMPLER_GUIDE);
        uint64_t lo = s.guide[g], hi = s.guide[g + 1] + 1;
        if (hi > s.I) hi = s.I;
        while (lo < hi) {
            const uint64_t mid = lo + ((hi - lo) >> 1);
            if (u < s.cdf[mid]) hi = mid;
            else lo = mid + 1;
        }
        return lo < s.I ? lo : s.I - 1;
    }
    const uint64_t k = (uint64_t)(u * (double)s.I);
    return k < s.I ? k : s.I - 1;
}

__device__ __forceinline__ void d_indices(const GenCtx& g, uint64_t d, uint64_t* ix)
{
    for (int n = 0; n < g.N; ++n) ix[n] = d_draw(g.m[n], d_u01(g.key_index[n], d));
}

// ztpoisson (rng.hpp), step for step.  The rejection loop of the PTRS branch
// is bounded on the device (the host's is not; its acceptance rate is above
// 0.9, so the bound is never met in practice): *bad gets bit 2 if it is.
__device__ uint64_t d_ztpoisson(double lam, uint64_t vkey, uint64_t base, int* bad)
{
    uint64_t j = 0;
    if (!(lam > 0.0)) return 1;
    if (lam <= 30.0) {
        const double p0 = exp(-lam);
        const double u = d_u01(vkey, base + j++);
        const double target = fma(u, 1.0 - p0, p0);
        double p = p0, c = p0;
        uint64_t k = 0;
        while (c < target && k < 10000) {
            ++k;
            p *= lam / (double)k;
            c += p;
        }
        return k < 1 ? 1 : k;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = fma(slam, 2.53, 0.931);
    const double a = fma(b, 0.02483, -0.059);
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < (1 << 20); ++it) {
        const double U = d_u01(vkey, base + j++) - 0.5;
        const double V = d_u01(vkey, base + j++);
        const double us = 0.5 - fabs(U);
        const double kf = floor(fma(2.0 * a / us + b, U, lam) + 0.43);
        if (us >= 0.07 && V <= vr) {
            if (kf >= 1.0) return (uint64_t)kf;
            continue;
        }
        if (kf < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(invalpha) - log(a / (us * us) + b) <= fma(kf, loglam, -lam) - lgamma(kf + 1.0)) {
            if (kf >= 1.0) return (uint64_t)kf;
        }
    }
    atomicOr(bad, 2);
    return 1;
}

// Linear keys of draws [d0, d0 + n): low 8 bits and the rest.
__global__ __launch_bounds__(WG) void k_draw_keys(GenCtx g, uint64_t n, uint8_t* __restrict__ lo8,
                                                  uint64_t* __restrict__ rest)
{
    for (uint64_t d = (uint64_t)blockIdx.x * WG + threadIdx.x; d < n; d += (uint64_t)gridDim.x * WG) {
        uint64_t ix[BLCO_MAX_MODES];
        d_indices(g, d, ix);
        unsigned __int128 k = 0;
        for (int m = 0; m < g.N; ++m) k |= (unsigned __int128)ix[m] << g.sh[m];
        lo8[d] = (uint8_t)k;
        rest[d] = (uint64_t)(k >> 8);
    }
}

__global__ __launch_bounds__(WG) void k_hist256(uint64_t n, const uint8_t* __restrict__ lo8,
                                                unsigned long long* __restrict__ hist)
{
    __shared__ unsigned int h[256];
    h[threadIdx.x] = 0;  // WG == 256
    __syncthreads();
    for (uint64_t d = (uint64_t)blockIdx.x * WG + threadIdx.x; d < n; d += (uint64_t)gridDim.x * WG)
        atomicAdd(&h[lo8[d]], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

struct PartPred {
    const uint8_t* lo8;
    uint32_t mask, p;
    __device__ bool operator()(uint64_t d) const { return ((uint32_t)lo8[d] & mask) == p; }
};
// key without the partition's b low bits: (rest << (8 - b)) | (lo8 >> b)
struct PartEmit {
    const uint8_t* lo8;
    const uint64_t* rest;
    int b;
    uint64_t* key;
    uint64_t* id;
    __device__ void operator()(uint64_t d, uint64_t pos) const
    {
        key[pos] = (rest[d] << (8 - b)) | ((uint64_t)lo8[d] >> b);
        id[pos] = d;
    }
};

// The first draw of each run of equal sorted keys is a first occurrence.
__global__ __launch_bounds__(WG) void k_mark_first(uint64_t n, const uint64_t* __restrict__ key,
                                                   const uint64_t* __restrict__ id, uint32_t* __restrict__ bitmap)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        if (i == 0 || key[i] != key[i - 1]) {
            const uint64_t d = id[i];
            atomicOr(&bitmap[d >> 5], 1u << (d & 31));
        }
}

__global__ __launch_bounds__(WG) void k_popcount(uint64_t words, const uint32_t* __restrict__ bitmap,
                                                 uint64_t* __restrict__ cnt, unsigned long long* __restrict__ total)
{
    unsigned long long t = 0;
    for (uint64_t w = (uint64_t)blockIdx.x * WG + threadIdx.x; w < words; w += (uint64_t)gridDim.x * WG) {
        const uint64_t c = (uint64_t)__popc(bitmap[w]);
        if (cnt) cnt[w] = c;
        t += c;
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += (unsigned long long)__shfl_xor((long long)t, o);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(total, t);
}

// kept[pos] = the draw of the pos-th first occurrence, pos < want.
__global__ __launch_bounds__(WG) void k_select_first(uint64_t words, const uint32_t* __restrict__ bitmap,
                                                     const uint64_t* __restrict__ woff, uint64_t want,
                                                     uint64_t* __restrict__ kept)
{
    for (uint64_t w = (uint64_t)blockIdx.x * WG + threadIdx.x; w < words; w += (uint64_t)gridDim.x * WG) {
        uint64_t pos = woff[w];
        uint32_t bits = bitmap[w];
        while (bits && pos < want) {
            const int j = __ffs(bits) - 1;
            kept[pos++] = (w << 5) | (uint64_t)j;
            bits &= bits - 1;
        }
    }
}

// Per kept draw: indices, rate, count, value, ALTO key (lo, hi), and the
// input position for the encoder's sort.
__global__ __launch_bounds__(WG) void k_values(GenCtx g, KeyLayout kl, uint64_t nnz,
                                               const uint64_t* __restrict__ kept, uint64_t* __restrict__ lo,
                                               uint64_t* __restrict__ hi, uint64_t* __restrict__ idx,
                                               double* __restrict__ vals, int* __restrict__ bad)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * WG) {
        const uint64_t d = kept[i];
        uint64_t ix[BLCO_MAX_MODES];
        d_indices(g, d, ix);
        double rho = 0.0;
        for (int r = 0; r < g.Rg; ++r) {
            double p = 1.0;
            for (int n = 0; n < g.N; ++n) p *= g.A[n][ix[n] * (uint64_t)g.Rg + r];
            rho += p;
        }
        const uint64_t c = d_ztpoisson(rho, g.key_value, d * 64, bad);
        double v;
        if (g.value_kind == 1) v = c < LOG1P_TABLE ? g.log1p_tab[c] : log1p((double)c);
        else v = (double)c;
        uint64_t l, h;
        alto_key(kl, ix, &l, &h);
        lo[i] = l;
        if (hi) hi[i] = h;
        idx[i] = i;
        vals[i] = v;
    }
}

__global__ __launch_bounds__(WG) void k_to_f32(uint64_t n, const double* __restrict__ src, float* __restrict__ dst)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        dst[i] = (float)src[i];
}

uint32_t grid_for(uint64_t n)
{
    const uint64_t g = (n + WG - 1) / WG;
    return (uint32_t)(g < 65536 ? (g ? g : 1) : 65536);
}

uint64_t device_free_bytes()
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return fr;
}

}  // namespace

// The first draws of the unique coordinates of the spec (in draw order):
// kept (device, *nnz entries, owned by the caller).
static int unique_draws_device(const blco_gen_spec* sp, const GenCtx& g, int total_bits, uint64_t** kept_out,
                               uint64_t* nnz_out, uint64_t* ndraw_out, hipStream_t st)
{
    *kept_out = nullptr;
    const uint64_t want = sp->target_nnz;
    // the host's round sizes and the limit of its draw-id type
    const double id_limit = (double)want * 8.0 < (double)std::numeric_limits<uint32_t>::max()
                                ? (double)std::numeric_limits<uint32_t>::max()
                                : (double)std::numeric_limits<uint64_t>::max();
    const int bmin = total_bits > 64 ? total_bits - 64 : 0;
    if (bmin > 8) return fail(BLCO_EUNSUPPORTED, "the GPU generator takes keys of at most 72 bits (%d)", total_bits);
    uint64_t ndraw = 0, unique = 0;
    DevBufs keep;  // released at exit
    uint32_t* bitmap = nullptr;
    uint64_t words = 0;
    for (int round = 0; round < 64; ++round) {
        uint64_t add = want;
        if (round > 0) {
            const double yield = std::max(0.01, (double)unique / (double)ndraw);
            add = (uint64_t)((double)(want - unique) / yield * 1.15) + 1024;
        }
        if ((double)ndraw + (double)add > id_limit) break;
        ndraw += add;
        DevBufs B;
        uint8_t* lo8;
        uint64_t* rest;
        int rc;
        if ((rc = B.alloc(&lo8, ndraw)) || (rc = B.alloc(&rest, ndraw))) return rc;
        hipLaunchKernelGGL(k_draw_keys, dim3(grid_for(ndraw)), dim3(WG), 0, st, g, ndraw, lo8, rest);
        HIP_TRY(hipGetLastError());
        unsigned long long* d_hist;
        if ((rc = B.alloc(&d_hist, 256))) return rc;
        HIP_TRY(hipMemsetAsync(d_hist, 0, 256 * sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_hist256, dim3(grid_for(ndraw)), dim3(WG), 0, st, ndraw, lo8, d_hist);
        HIP_TRY(hipGetLastError());
        unsigned long long hist[256];
        HIP_TRY(hipMemcpyAsync(hist, d_hist, sizeof(hist), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        // partition bits: keys fit 64 bits, partitions fit the sort's item
        // count and the device (keys and ids, in and out: 32 bytes each)
        words = (ndraw + 31) / 32;
        if (bitmap) keep.release(bitmap);
        if ((rc = keep.alloc(&bitmap, words))) return rc;
        HIP_TRY(hipMemsetAsync(bitmap, 0, words * 4, st));
        const uint64_t avail = device_free_bytes();
        int b = -1;
        uint64_t most = 0;
        for (int bb = bmin; bb <= 8 && b < 0; ++bb) {
            const uint32_t P = 1u << bb;
            uint64_t mx = 0;
            for (uint32_t p = 0; p < P; ++p) {
                uint64_t c = 0;
                for (uint32_t v = p; v < 256; v += P) c += hist[v];
                mx = std::max(mx, c);
            }
            if (mx <= 0x7fffffffull && (double)mx * 32.0 + (double)(64 << 20) < 0.9 * (double)avail) {
                b = bb;
                most = mx;
            }
        }
        if (b < 0)
            return fail(BLCO_ENOMEM, "GPU generator: %llu draws do not fit the device in 256 partitions",
                        (unsigned long long)ndraw);
        uint64_t *k1, *k2, *i1, *i2;
        if ((rc = B.alloc(&k1, most)) || (rc = B.alloc(&k2, most)) || (rc = B.alloc(&i1, most)) ||
            (rc = B.alloc(&i2, most)))
            return rc;
        const int sort_bits = std::max(1, total_bits - b);
        size_t tmp_bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k1, k2, i1, i2, std::max<uint64_t>(most, 1), 0,
                                                   sort_bits, st));
        unsigned char* tmp;
        if ((rc = B.alloc(&tmp, tmp_bytes ? tmp_bytes : 1))) return rc;
        CompactScratch cs;
        const uint32_t P = 1u << b;
        for (uint32_t p = 0; p < P; ++p) {
            uint64_t cnt = 0;
            rc = compact(ndraw, PartPred{lo8, P - 1, p}, PartEmit{lo8, rest, b, k1, i1}, &cnt, &cs, st);
            if (rc) {
                cs.release();
                return rc;
            }
            if (cnt == 0) continue;
            size_t have = tmp_bytes;
            HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, have, k1, k2, i1, i2, cnt, 0, sort_bits, st));
            hipLaunchKernelGGL(k_mark_first, dim3(grid_for(cnt)), dim3(WG), 0, st, cnt, k2, i2, bitmap);
            HIP_TRY(hipGetLastError());
        }
        cs.release();
        unsigned long long* d_total;
        if ((rc = B.alloc(&d_total, 1))) return rc;
        HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(unsigned long long), st));
        hipLaunchKernelGGL(k_popcount, dim3(grid_for(words)), dim3(WG), 0, st, words, bitmap, nullptr, d_total);
        HIP_TRY(hipGetLastError());
        unsigned long long u = 0;
        HIP_TRY(hipMemcpyAsync(&u, d_total, sizeof(u), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        unique = u;
        progress("GPU generator: round %d, %llu draws, %llu unique (%u partitions)", round,
                 (unsigned long long)ndraw, (unsigned long long)unique, P);
        if (!sp->topup || unique >= want) break;
    }
    const uint64_t nnz = sp->topup ? std::min(unique, want) : unique;
    *nnz_out = nnz;
    *ndraw_out = ndraw;
    if (nnz == 0) return BLCO_OK;
    // the first nnz first occurrences in draw order
    DevBufs B;
    uint64_t *wc, *woff, *kept;
    int rc;
    if ((rc = B.alloc(&wc, words + 1)) || (rc = B.alloc(&woff, words + 1))) return rc;
    unsigned long long* d_total;
    if ((rc = B.alloc(&d_total, 1))) return rc;
    HIP_TRY(hipMemsetAsync(d_total, 0, sizeof(unsigned long long), st));
    HIP_TRY(hipMemsetAsync(wc + words, 0, sizeof(uint64_t), st));
    hipLaunchKernelGGL(k_popcount, dim3(grid_for(words)), dim3(WG), 0, st, words, bitmap, wc, d_total);
    HIP_TRY(hipGetLastError());
    size_t tmp_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, wc, woff, words + 1, st));
    unsigned char* tmp;
    if ((rc = B.alloc(&tmp, tmp_bytes ? tmp_bytes : 1))) return rc;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, wc, woff, words + 1, st));
    if (hipMalloc(&kept, nnz * sizeof(uint64_t)) != hipSuccess)
        return fail(BLCO_ENOMEM, "GPU generator: %llu nonzeros", (unsigned long long)nnz);
    hipLaunchKernelGGL(k_select_first, dim3(grid_for(words)), dim3(WG), 0, st, words, bitmap, woff, nnz, kept);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipFree(kept);
        return fail(BLCO_EHIP, "GPU generator: selection failed");
    }
    *kept_out = kept;
    return BLCO_OK;
}

}  // namespace blco

using namespace blco;

extern "C" int blco_generate_device(const blco_gen_spec* sp, uint64_t max_block_size, int value_dtype,
                                    blco_dev** out)
{
    if (!sp || !out) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    if (value_dtype != BLCO_F64 && value_dtype != BLCO_F32) return fail(BLCO_EINVAL, "value dtype %d", value_dtype);
    const int N = sp->nmodes;
    Layout L;
    int rc = make_layout(N, sp->dims, &L);
    if (rc) return rc;
    if (sp->gen_rank < 1) return fail(BLCO_EINVAL, "gen_rank must be >= 1");
    if (sp->target_nnz == 0) return fail(BLCO_EINVAL, "target_nnz must be >= 1");
    long double cells = 1.0L;
    for (int n = 0; n < N; ++n) cells *= (long double)sp->dims[n];
    if ((long double)sp->target_nnz > cells) return fail(BLCO_EINVAL, "target_nnz exceeds the number of cells");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BLCO_ENODEV, "no GPU visible");
    hipStream_t st = nullptr;
    DevBufs T;  // tables
    GenCtx g;
    std::memset(&g, 0, sizeof(g));
    g.N = N;
    g.Rg = sp->gen_rank;
    g.value_kind = sp->value_kind;
    int sh = 0;
    for (int n = 0; n < N; ++n) {
        g.sh[n] = sh;
        sh += L.bits[n];
        g.key_index[n] = mix64(sp->seed ^ mix64((STREAM_INDEX + (uint64_t)n) * 0xD6E8FEB86659FD93ull));
        std::vector<double> cdf;
        std::vector<uint64_t> guide;
        sampler_tables(sp->dims[n], sp->zipf_s[n], &cdf, &guide);
        g.m[n].I = sp->dims[n];
        if (!cdf.empty()) {
            double* dc;
            uint64_t* dg;
            if ((rc = T.alloc(&dc, cdf.size())) || (rc = T.alloc(&dg, guide.size()))) return rc;
            HIP_TRY(hipMemcpy(dc, cdf.data(), cdf.size() * 8, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(dg, guide.data(), guide.size() * 8, hipMemcpyHostToDevice));
            g.m[n].cdf = dc;
            g.m[n].guide = dg;
        }
    }
    g.key_value = mix64(sp->seed ^ mix64(STREAM_VALUE * 0xD6E8FEB86659FD93ull));
    progress("GPU generator: tables ready");
    uint64_t* kept = nullptr;
    uint64_t nnz = 0, ndraw = 0;
    rc = unique_draws_device(sp, g, L.total_bits, &kept, &nnz, &ndraw, st);
    if (rc) return rc;
    progress("GPU generator: %llu unique coordinates from %llu draws", (unsigned long long)nnz,
             (unsigned long long)ndraw);
    DevBufs K;
    if (kept) K.p.push_back(kept);
    // generating factors and the log1p table
    for (int n = 0; n < N; ++n) {
        std::vector<double> a(sp->dims[n] * (uint64_t)g.Rg);
        generating_factor(sp->seed, n, sp->dims[n], g.Rg, a.data());
        double* da;
        if ((rc = T.alloc(&da, a.size()))) return rc;
        HIP_TRY(hipMemcpy(da, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        g.A[n] = da;
    }
    {
        std::vector<double> tab(LOG1P_TABLE);
        for (uint32_t c = 0; c < LOG1P_TABLE; ++c) tab[c] = log1p_count(c);
        double* dt;
        if ((rc = T.alloc(&dt, tab.size()))) return rc;
        HIP_TRY(hipMemcpy(dt, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
        g.log1p_tab = dt;
    }
    // values and ALTO keys of the kept draws, then the encoder's sort
    const bool wide = L.total_bits > 64;
    uint64_t *lo = nullptr, *hi = nullptr, *idx = nullptr;
    double* vals = nullptr;
    int* d_bad;
    DevBufs V;
    if ((rc = V.alloc(&lo, nnz)) || (rc = V.alloc(&idx, nnz)) || (rc = V.alloc(&vals, nnz)) ||
        (rc = V.alloc(&d_bad, 1)))
        return rc;
    if (wide && (rc = V.alloc(&hi, nnz))) return rc;
    HIP_TRY(hipMemset(d_bad, 0, sizeof(int)));
    const KeyLayout kl = key_layout(L);
    if (nnz) {
        hipLaunchKernelGGL(k_values, dim3(grid_for(nnz)), dim3(WG), 0, st, g, kl, nnz, kept, lo, hi, idx, vals, d_bad);
        HIP_TRY(hipGetLastError());
    }
    int bad = 0;
    HIP_TRY(hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost));
    if (bad) return fail(BLCO_EHIP, "GPU generator: a Poisson draw did not terminate");
    K.release(kept);
    progress("GPU generator: values drawn");
    blco_host h;
    h.L = L;
    h.nnz = nnz;
    h.max_block_size = max_block_size;
    const uint64_t mbs = max_block_size ? max_block_size : (nnz ? nnz : 1);  // blco.hpp:264
    rc = encode_sorted_device(L, nnz, lo, hi, idx, vals, mbs, &h, st);
    if (rc) return rc;
    // lo holds the BLCO words, idx the values (f64)
    V.release(hi);
    V.release(vals);
    void* dvals = idx;
    if (value_dtype == BLCO_F32 && nnz) {
        float* f;
        if (hipMalloc(&f, nnz * sizeof(float)) != hipSuccess)
            return fail(BLCO_ENOMEM, "GPU generator: f32 values of %llu nonzeros", (unsigned long long)nnz);
        hipLaunchKernelGGL(k_to_f32, dim3(grid_for(nnz)), dim3(WG), 0, st, nnz, reinterpret_cast<const double*>(idx), f);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            (void)hipFree(f);
            return fail(BLCO_EHIP, "GPU generator: value conversion failed");
        }
        V.release(idx);
        dvals = f;
    }
    // the host-side tables of the device tensor come from a host copy of the words
    std::vector<uint64_t> hc;
    try {
        hc.resize(nnz);
    } catch (const std::bad_alloc&) {
        if (dvals != idx) (void)hipFree(dvals);
        return fail(BLCO_ENOMEM, "host copy of %llu BLCO words", (unsigned long long)nnz);
    }
    if (nnz && hipMemcpy(hc.data(), lo, nnz * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        if (dvals != idx) (void)hipFree(dvals);
        return fail(BLCO_EHIP, "GPU generator: download of the BLCO words failed");
    }
    // the tensor takes over lo and the values
    for (void*& q : V.p)
        if (q == lo || q == dvals) q = nullptr;
    progress("GPU generator: encoded, %zu block(s)", h.block_start.size());
    return create_device_tensor(N, L.dims, L.masks, L.pos, nnz, hc.data(), nullptr, h.block_start.size(),
                                h.block_start.data(), h.block_nnz.data(), h.block_coords.data(), value_dtype, 0, 0, 0,
                                nnz ? lo : nullptr, nnz ? dvals : nullptr, out);
}
