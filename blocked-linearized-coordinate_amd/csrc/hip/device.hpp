// device.hpp -- the device tensor object behind blco_dev.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <vector>

#include "blco_hip.h"
#include "mttkrp.hpp"
#include "../host/internal.hpp"

namespace blco {

// A contiguous chunk range of one block that fits one streaming slot.
struct StreamUnit {
    uint32_t c0, c1;
    uint64_t begin;  // first nonzero
    uint64_t nnz;
};

// Super-chunks of 2^k fine chunks (within one block) with per-mode row
// ranges, for the row-window kernel.
struct WindowTab {
    int k = 0;
    SuperChunk* d_sc = nullptr;
    uint64_t* d_range = nullptr;  // [mode][entry][2]
    uint32_t n = 0;
    std::vector<uint32_t> first;  // first entry of each block (nblocks + 1)
    std::vector<uint64_t> h_range;
    std::vector<SuperChunk> h_sc;
};

// The row-window MTTKRP of one (mode, rank, accumulator size): window table,
// LDS split, where each workgroup writes its partial rows, and the
// partial-row lists that the reduce folds per output row.
struct WindowPlan {
    int mode = 0, rank = 0, acc = 0;
    const WindowTab* wt = nullptr;
    uint32_t lds_rows = 0, hot_rows = 0;
    uint64_t* d_poff = nullptr;  // [entry + 1]: window of entry e = partial rows [poff[e], poff[e+1])
    uint64_t hot_poff = 0;       // first partial row of the hot areas
    uint64_t prows = 0;          // partial rows in all
    uint64_t nred = 0;           // reduce segments (<= REDUCE_SEG partial rows of one output row)
    uint64_t* d_rows = nullptr;  // [nred] output row of each segment, | REDUCE_SHARED when split
    uint64_t* d_ptr = nullptr;   // [nred + 1] CSR into d_idx
    uint32_t* d_idx = nullptr;   // [prows] partial rows, grouped by output row
};

// A selector decision for (mode, rank, factor type); kernel 0 = none.
struct SelCache {
    int mode, rank, ftype, kernel, nprtn, tpw;
    uint64_t gen;  // selector generation the decision was made with
};

struct ChunkTab {
    int tpw = 0;
    Chunk* d = nullptr;
    uint32_t n = 0;
    std::vector<uint32_t> first;  // first chunk of each block (nblocks + 1)
    std::vector<StreamUnit> units;
};

// A contiguous run of a row-ordered copy: the nonzeros of one BLCO block
// whose split-mode index falls in one tile.
struct SGroup {
    uint64_t start, nnz;
    uint32_t block, tile;
};

}  // namespace blco

struct blco_dev {
    int nmodes = 0;
    uint64_t dims[BLCO_MAX_MODES] = {};
    uint64_t masks[BLCO_MAX_MODES] = {};
    int pos[BLCO_MAX_MODES] = {};
    uint64_t nnz = 0;
    int vdtype = BLCO_F64;
    int stream_data = 0;
    uint64_t slot_nnz = 0;
    uint64_t nblocks = 0;
    std::vector<uint64_t> bstart, bnnz;
    uint64_t* d_bcoords = nullptr;
    uint64_t* d_coords = nullptr;
    void* d_vals = nullptr;
    std::list<blco::ChunkTab> tabs;
    std::list<blco::WindowTab> wtabs;
    // fine chunks (FINE_CHUNK nonzeros, aligned per block) and their
    // per-mode output row ranges: fine_lo/hi[mode * nfine + f]
    std::vector<uint64_t> fine_begin;
    std::vector<uint32_t> fine_len, fine_block;
    std::vector<uint64_t> fine_lo, fine_hi;
    std::list<blco::WindowPlan> wplans;
    // most frequent rows per mode (count >= nnz / 2048, at most HOT_EXCLUDE),
    // in decreasing frequency; fine ranges exclude them
    std::vector<uint32_t> hot_list_h[BLCO_MAX_MODES];
    std::vector<uint64_t> hot_cnt_h[BLCO_MAX_MODES];  // their nonzero counts
    uint32_t* d_hot_list[BLCO_MAX_MODES] = {};
    int32_t* d_hot_rank[BLCO_MAX_MODES] = {};
    // row-ordered copies (kernel 14): BLCO words and values of every block
    // sorted by the mode's index, built on first use
    uint64_t* d_scoords[BLCO_MAX_MODES] = {};
    void* d_svals[BLCO_MAX_MODES] = {};
    bool sorted_failed[BLCO_MAX_MODES] = {};  // no device memory for the copy
    // groups of the copy in storage order (tile-major), the number of tiles
    // (1: untiled) and the positions where a run of equal (group, row)
    // starts, then nnz
    std::vector<blco::SGroup> sgrp_h[BLCO_MAX_MODES];
    int stiles[BLCO_MAX_MODES] = {};
    // fiber order of the copy: within a row by every input mode but sinner;
    // sfiber: the fiber kernel runs on it (fibers long enough)
    int sinner[BLCO_MAX_MODES] = {};
    bool sfiber[BLCO_MAX_MODES] = {};
    std::vector<uint64_t> srun_h[BLCO_MAX_MODES];
    std::list<blco::ChunkTab> stabs[BLCO_MAX_MODES];  // row-aware chunks of the copy, per tpw
    // nonzeros per index of every mode, as prefix sums (dims + 1 entries),
    // for the hot-row windows of the row-ordered kernels (built on first use)
    std::vector<uint64_t> idx_prefix[BLCO_MAX_MODES];
    std::vector<blco::RowsPlan> rows_plans[BLCO_MAX_MODES];
    void* part = nullptr;  // partial rows of the row-window MTTKRP
    size_t part_bytes = 0;
    // selector features (rank- and type-independent entries) and decisions
    float sfeat[BLCO_MAX_MODES][BLCO_SEL_NFEAT] = {};
    std::vector<blco::SelCache> sel_cache;
    // streaming: the whole tensor in pinned host memory; its first res_nnz
    // nonzeros also stay on the device (the rest pass through the slots)
    uint64_t res_nnz = 0;
    uint64_t budget = 0;  // device bytes for nonzeros, slots and copies (0: no cap)
    uint64_t* h_coords = nullptr;
    void* h_vals = nullptr;
    uint64_t* slot_coords[2] = {nullptr, nullptr};
    void* slot_vals[2] = {nullptr, nullptr};
    hipEvent_t ev_copied[2] = {nullptr, nullptr};
    hipEvent_t ev_done[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    // lvl3 partial matrices
    void* partials = nullptr;
    size_t partials_bytes = 0;
    hipStream_t stream = nullptr;
    uint64_t bytes = 0;
    int device = 0;
};

namespace blco {

const ChunkTab* chunk_table(blco_dev* d, int tpw);
// Row-ordered copies built on the GPU (roworder_dev.hip): nonzeros per
// split-mode index; per output row the tiles (bit t for tile t) its nonzeros
// meet; and the copy itself in the order of row_order (csrc/host), with its
// groups in storage order, the run starts of (group, row) followed by nnz,
// and the fiber count when fibers is non-null.  BLCO_EUNSUPPORTED when the
// sort key would exceed 64 bits.
int dev_split_counts(const blco_dev* d, int split, std::vector<uint64_t>* cnt);
int dev_row_tiles(const blco_dev* d, int mode, int split, const std::vector<uint8_t>& tile_of,
                  std::vector<uint8_t>* met);
int dev_row_copy(const blco_dev* d, const RowOrderSpec& o, std::vector<SGroup>* groups, std::vector<uint64_t>* runs,
                 uint64_t* fibers, uint64_t** sc, void** sv);
const WindowTab* window_table(blco_dev* d, int k);
const WindowPlan* window_plan(blco_dev* d, int mode, int rank, int acc_bytes);
int auto_kernel(const blco_dev* d, int mode, int rank, int ftype);
bool lvl3_output_fits(const blco_dev* d, int mode, int rank, int ftype);
int mode_features(const blco_dev* d, int mode, int rank, int ftype, float* out);
bool select_kernel(blco_dev* d, int mode, int rank, int ftype, int* kernel, int* nprtn, int* tpw);
int mttkrp(blco_dev* d, int mode, int rank, int ftype, void* const* factors, void* out, int kernel,
           int nprtn, int tpw, hipStream_t st, int* used_kernel);
// The host-side and device-side set-up that mttkrp() of (mode, rank, ftype,
// kernel) would do on its first call (kernel 14's row-ordered copy), without
// a launch; a copy that does not fit leaves the mode to the row window.
int mttkrp_prepare(blco_dev* d, int mode, int rank, int ftype, int kernel);
size_t rows_kernel_plan(blco_dev* d, int mode, int rank, int ftype, bool fib, MttkrpArgs* a);

}  // namespace blco
