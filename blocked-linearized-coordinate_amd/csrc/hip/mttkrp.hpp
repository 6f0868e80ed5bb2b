// mttkrp.hpp -- kernel arguments and launchers of mttkrp.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "blco_host.h"

namespace blco {

// A run of nonzeros of one block handled by one wave (begin is the global
// nonzero offset in ALTO order).
struct Chunk {
    uint64_t begin;
    uint32_t len;
    uint32_t block;
};

// A workgroup's share of a block: a run of nonzeros whose output rows, for
// each mode, lie in a known [lo, hi] (the row window).
struct SuperChunk {
    uint64_t begin;
    uint32_t len;
    uint32_t block;
};

struct MttkrpArgs {
    const uint64_t* coords;  // element e lives at coords[e - coord_base]
    const void* vals;        // same indexing, f64 or f32
    uint64_t coord_base;
    const Chunk* chunks;
    uint32_t chunk0;         // first chunk of this launch
    uint32_t nchunks;        // chunks in this launch
    const uint64_t* bcoords; // block coordinates, nblocks x nmodes
    uint64_t mask[BLCO_MAX_MODES];
    int pos[BLCO_MAX_MODES];
    int nmodes;
    int mode;
    int rank;
    const void* f[BLCO_MAX_MODES];
    void* out;
    // lvl3 only
    uint64_t out_rows;
    int nprtn;
    void* partials;
    uint32_t chunks_per_wg;
    // row-window kernel only
    const SuperChunk* sc;
    const uint64_t* sc_range;  // [entry][2]: lo, hi of the output mode (non-hot rows)
    uint32_t lds_rows;         // rows the LDS window holds
    const int32_t* hot_rank;   // output mode: frequency rank of a row, -1 if not a candidate
    uint32_t hot_rows;         // ranks < hot_rows are accumulated in the LDS hot area
    const uint64_t* sc_poff;   // [entry + 1]: the entry's window is partial rows [poff[e], poff[e+1])
    uint64_t hot_poff;         // first partial row of the hot areas (entry e: + e * hot_rows)
    void* part;                // partial rows (rank wide), reduced by k_window_reduce
    // fiber kernel only: the input mode gathered per nonzero
    int inner;
    // row-ordered kernels: chunks per wave (consecutive, or with
    // cpw_interleave one grid width apart), and for the first two input
    // modes (the inner and first outer mode with fibers) a window of rows
    // [hot_lo, hot_lo + hot_n) kept in LDS (hot_n = 0: none)
    uint32_t cpw;
    uint32_t cpw_interleave;
    uint32_t hot_lo[2];
    uint32_t hot_n[2];
};

// Launch plan of a row-ordered kernel for one (rank, factor type, form,
// chunk count) of a mode: chunks per wave and the hot-row windows.
struct RowsPlan {
    int rank, ftype;
    bool fib;
    uint32_t nchunks, cpw, interleave, hot_lo[2], hot_n[2];
    size_t lds;
};

enum { W_LVL1 = 0, W_LVL3 = 1, W_ATOMIC = 2, W_WINDOW = 3, W_WINDOW64 = 4, W_SORTED = 5, W_SORTED64 = 6,
       W_FIBER = 7, W_FIBER64 = 8, W_SORTED_LDS = 9, W_FIBER_LDS = 10 };

// The row-ordered and fiber kernels take 31-bit byte offsets into factors
// of fewer bytes than this (bit 31 marks an offset into an LDS window).
constexpr uint32_t FACTOR_BYTES_MAX = 0x80000000u;
// LDS for the hot-row windows of the row-ordered kernels (per workgroup)
constexpr uint32_t HOT_TABLE_BYTES = 24 * 1024;
constexpr uint32_t FINE_CHUNK = 1024;        // nonzeros per range-table entry
constexpr uint32_t WINDOW_LDS_BYTES = 40 * 1024;  // hot area + row window per workgroup
constexpr int WIN_WAVES = 8;                      // waves per row-window workgroup
constexpr int WIN_BATCH = 4;                      // positions whose gathers are issued together

// Marks a reduce segment whose output row is split over several segments.
constexpr uint64_t REDUCE_SHARED = 1ull << 63;

// out[rows[u]][:] += sum of the partial rows idx[ptr[u] .. ptr[u+1]), with
// an atomic add when rows[u] carries REDUCE_SHARED.
int launch_window_reduce(void* out, const void* part, const uint64_t* rows, const uint64_t* ptr,
                         const uint32_t* idx, uint64_t nrows, int rank, int otype, hipStream_t st);

const void* mttkrp_kernel(int which, int nmodes, int ftype, int rank, int* C, int* L);
int launch_mttkrp(int which, const MttkrpArgs& a, int ftype, uint32_t grid, size_t dyn_lds,
                  hipStream_t st);
int launch_reduce_partials(void* out, const void* part, uint64_t cells, int nprtn, int otype,
                           hipStream_t st);
int launch_fill(void* x, uint64_t n, int dtype, double v, hipStream_t st);

}  // namespace blco
