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
    void* hotpart;             // hot_parts x hot_rows x rank partial copies
    int hot_parts;
};

enum { W_LVL1 = 0, W_LVL3 = 1, W_ATOMIC = 2, W_WINDOW = 3 };

constexpr uint32_t FINE_CHUNK = 1024;        // nonzeros per range-table entry
constexpr uint32_t WINDOW_LDS_BYTES = 48 * 1024;
constexpr uint32_t HOT_CANDIDATES = 512;     // most frequent rows ranked per mode
constexpr int HOT_PARTS = 32;                // partial copies of the hot rows

int launch_hot_reduce(void* out, void* hotpart, const uint32_t* hot_list, uint32_t hot_rows,
                      int rank, int parts, int otype, hipStream_t st);

const void* mttkrp_kernel(int which, int nmodes, int ftype, int rank, int* C, int* L);
int launch_mttkrp(int which, const MttkrpArgs& a, int ftype, uint32_t grid, size_t dyn_lds,
                  hipStream_t st);
int launch_reduce_partials(void* out, const void* part, uint64_t cells, int nprtn, int otype,
                           hipStream_t st);
int launch_fill(void* x, uint64_t n, int dtype, double v, hipStream_t st);

}  // namespace blco
