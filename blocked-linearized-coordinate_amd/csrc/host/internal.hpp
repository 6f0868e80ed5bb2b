// internal.hpp -- host-side objects behind the C ABI (blco_host.h).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "blco_host.h"

namespace blco {

// A line on stderr when BLCO_PROGRESS is set (long host phases of large
// tensors: generation, encoding, row-ordered copies).
void progress(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
// The block holding nonzero i (blocks are contiguous, in order; i < nnz).
uint64_t block_of(const uint64_t* block_start, uint64_t nblocks, uint64_t i);

// Sets the thread's last-error message and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// 64 - clz(d - 1); a mode of length 1 needs 0 bits (the reference evaluates
// clz(0) there, which is undefined: include/alto.hpp:613).
inline int mode_bits(uint64_t d) { return d <= 1 ? 0 : 64 - __builtin_clzll(d - 1); }

struct Layout {
    int nmodes = 0;
    uint64_t dims[BLCO_MAX_MODES] = {};
    int bits[BLCO_MAX_MODES] = {};
    int total_bits = 0;
    // ALTO interleave masks (setup_packed_alto, include/alto.hpp:597-695)
    uint64_t alto_lo[BLCO_MAX_MODES] = {};
    uint64_t alto_hi[BLCO_MAX_MODES] = {};
    // BLCO re-encoding (include/blco.hpp:222-233)
    int trunc[BLCO_MAX_MODES] = {};
    uint64_t masks[BLCO_MAX_MODES] = {};
    int pos[BLCO_MAX_MODES] = {};
};

// Computes the ALTO / BLCO layout of a shape; returns BLCO_OK or an error.
int make_layout(int nmodes, const uint64_t* dims, Layout* L);

// The order of a row-ordered copy (roworder.cpp): nonzeros grouped by
// (tile of the split mode's index, block) -- ntiles <= 1: by block --, by the
// output mode's index within a group and, with nouter > 0, by the outer modes'
// local indices within a row (the outer modes' mask bits must total <= 64).
struct RowOrderSpec {
    int mode = 0, nmodes = 0;
    const uint64_t* masks = nullptr;
    const int* pos = nullptr;
    uint64_t nblocks = 0;
    const uint64_t* block_start = nullptr;
    const uint64_t* block_nnz = nullptr;
    const uint64_t* bcoords = nullptr;  // nblocks x nmodes
    int split = -1, ntiles = 1;
    const uint8_t* tile_of = nullptr;   // tile of every split-mode index
    int nouter = 0;
    int outer[BLCO_MAX_MODES] = {};
};
// Stable permutation into that order; group_nnz[tile * nblocks + block]
// receives the group sizes.
void row_order(const RowOrderSpec& o, uint64_t nnz, const uint64_t* coords, uint64_t* perm, uint64_t* group_nnz);
// dst[i] = src[perm[i]] for esz-byte elements.
void gather_bytes(const void* src, size_t esz, const uint64_t* perm, uint64_t n, void* dst);

// Pieces of the synthetic-tensor generator (generate.cpp) that the GPU
// generator (gen_dev.hip) takes from the host so that both form the same
// doubles: the Zipf sampler's tables of a mode of length I (cdf[k] =
// sum_{j<=k} (j+1)^-s / sum_j (j+1)^-s, and guide[g] = the first k with
// cdf[k] > g / SAMPLER_GUIDE; both empty for s = 0, a uniform mode), the
// Exp(1) generating factor of a mode (rows x Rg), and log1p of a count.
constexpr uint32_t SAMPLER_GUIDE = 1u << 16;
void sampler_tables(uint64_t I, double s, std::vector<double>* cdf, std::vector<uint64_t>* guide);
void generating_factor(uint64_t seed, int mode, uint64_t rows, int Rg, double* out);
double log1p_count(uint64_t c);

// Host-side tables of a device tensor (analyze.cpp): fine chunks of at most
// fine_chunk nonzeros within each block (first nonzero, length, block), the
// hot rows of each mode (count >= max(2, nnz / 2048), at most hot_max, by
// decreasing count, and their counts), and per mode and fine chunk the row
// range [flo, fhi] of the non-hot rows (flo = ~0 when none).  Fails when an
// index decodes outside its mode.
struct TensorTables {
    std::vector<uint64_t> fbeg, flo, fhi;
    std::vector<uint32_t> flen, fblk;
    std::vector<std::vector<uint32_t>> hot;
    std::vector<std::vector<uint64_t>> hot_cnt;
};
int analyze_tensor(int nmodes, const uint64_t* dims, const uint64_t* masks, const int* pos, uint64_t nnz,
                   const uint64_t* coords, uint64_t nblocks, const uint64_t* block_start,
                   const uint64_t* block_nnz, const uint64_t* block_coords, uint32_t fine_chunk, uint32_t hot_max,
                   TensorTables* t);

// Kernel-selector features of every mode that depend on the tensor only
// (entries 0-3, 5, 7-22 of each row of `stride` floats; features.cpp).
void compute_static_features(int N, const uint64_t* dims, const uint64_t* masks, const int* pos, uint64_t nnz,
                             const uint64_t* coords, uint64_t nblocks, const uint64_t* block_start,
                             const uint64_t* block_nnz, const uint64_t* block_coords, int stride, float* out);

}  // namespace blco

struct blco_coo {
    int nmodes = 0;
    uint64_t dims[BLCO_MAX_MODES] = {};
    uint64_t nnz = 0;
    std::vector<uint64_t> cidx;  // mode-major
    std::vector<double> vals;
};

struct blco_host {
    blco::Layout L;
    uint64_t nnz = 0;
    uint64_t max_block_size = 0;
    uint64_t max_block_nnz = 0;
    std::vector<uint64_t> coords;
    std::vector<double> vals;
    std::vector<uint64_t> block_start;
    std::vector<uint64_t> block_nnz;
    std::vector<uint64_t> block_coords;  // nblocks x nmodes
};

namespace blco {
// Blocks of a BLCO tensor from the sorted keys' upper halves (hi[i] of
// sorted nonzero i, or nullptr for keys of at most 64 bits): runs of equal
// upper half split into pieces of at most mbs nonzeros (encode.cpp).
void build_blocks(const Layout& L, uint64_t nnz, const uint64_t* hi, uint64_t mbs, blco_host* h);
// The same from the runs themselves: run r starts at start[r] (ascending,
// start[0] = 0) and has upper half hival[r] (hival may be nullptr: 0).
void build_blocks_runs(const Layout& L, uint64_t nnz, uint64_t nruns, const uint64_t* start, const uint64_t* hival,
                       uint64_t mbs, blco_host* h);
}  // namespace blco
