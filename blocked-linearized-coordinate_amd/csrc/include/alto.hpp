// alto.hpp -- the ALTO tensor record that a BLCO block extends
// (include/alto.hpp:22-51 of the reference).  Only the record is kept: the
// CPU ALTO routines of that header are not part of this library (its CPU
// MTTKRP lives in the test oracle).  The fields keep the reference's names
// and types so that code reading a block's nmode, nnz, dims, mode_masks,
// mode_pos, idx or vals compiles unchanged.
#pragma once

#include "common.hpp"

typedef struct Interval_ {
    IType start;
    IType stop;
} Interval;

template <typename LIT>
struct AltoTensor {
    int nmode = 0;                    // number of modes
    int nprtn = 0;                    // number of partitions (CPU ALTO; unused here)
    IType nnz = 0;                    // number of nonzeros
    IType* dims = nullptr;            // mode lengths (the parent tensor's)
    LIT alto_mask = 0;
    LIT* mode_masks = nullptr;        // BLCO field masks of the lower key bits (nmode)
    int* mode_pos = nullptr;          // shift of each mode's field (nmode)
    LIT* idx = nullptr;               // BLCO coordinates of the nonzeros
    FType* vals = nullptr;            // values
    IType* prtn_ptr = nullptr;        // CPU ALTO partitions (unused here)
    Interval* prtn_intervals = nullptr;
    LIT alto_cr_mask = 0;
    LIT* cr_masks = nullptr;
};
