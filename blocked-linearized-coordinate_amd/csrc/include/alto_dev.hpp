// alto_dev.hpp -- GPU MTTKRP API of the reference (include/alto_dev.hpp:18-27).
#pragma once

#include "blco.hpp"
#include "kruskal_model.hpp"

// Masks travel with the device tensor; kept for the reference call sequence.
void send_masks_over(blcotensor* at_host);
// Partial matrices are allocated by the library on first use; returns null.
FType** create_da_mem_dev(int kernel, int target_mode, IType rank, int nprtn, blcotensor* at);

// MTTKRP benchmark driver (src/alto_dev.cu:1856-1952): runs `iters` MTTKRPs of
// target_mode (all modes for -1), prints "Total time for MTTKRP: ..." and
// copies the last result into M->U[target_mode] (plus the initial factor, as
// the reference accumulates into it).
template <typename LIT>
void mttkrp_alto_dev(blcotensor* at_host, KruskalModel* M, int kernel, int iters, int target_mode,
                     IType thread_cf, bool stream_data, bool do_batching, int nprtn);
void mttkrp_alto_dev_any(blcotensor* at_host, KruskalModel* M, int kernel, int iters,
                         int target_mode, IType thread_cf, bool stream_data, bool do_batching,
                         int nprtn);
template <typename LIT>
inline void mttkrp_alto_dev(blcotensor* at_host, KruskalModel* M, int kernel, int iters,
                            int target_mode, IType thread_cf, bool stream_data, bool do_batching,
                            int nprtn)
{
    mttkrp_alto_dev_any(at_host, M, kernel, iters, target_mode, thread_cf, stream_data,
                        do_batching, nprtn);
}

// One MTTKRP into device `output` (dims[target_mode] x rank f64, accumulated)
// from device factors mats_dev (f64) (src/alto_dev.cu:1954-2032).
void mttkrp_alto_dev_onemode_any(blcotensor* at_dev, IType rank, FType* output, int kernel,
                                 int target_mode, IType thread_cf, int nprtn, FType** mats_dev);
