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

// mttkrp_alto_dev_onemode (include/alto_dev.hpp:26-27,
// src/alto_dev.cu:1954-2032): output (device, modes[target_mode] x rank f64)
// += the mode-target_mode MTTKRP of at_dev with the device factors
// mats_staging_ptr[m] (a host array of device pointers, as the reference's
// M_dev->U).  The rank is M->rank; stream_data follows the device tensor
// (it was fixed when at_dev was made).  mats_dev and the partial matrices
// are accepted for the reference's call sequence: the library's kernels
// take the factor pointers by value and keep their partial matrices with
// the device tensor.
void mttkrp_alto_dev_onemode_ref(blcotensor* at_host, blcotensor* at_dev, KruskalModel* M,
                                 FType* output, int kernel, int target_mode, IType thread_cf,
                                 bool stream_data, int nprtn, FType** mats_staging_ptr,
                                 FType** mats_dev, FType** partial_matrices,
                                 FType** partial_matrices_dev);
template <typename LIT>
inline void mttkrp_alto_dev_onemode(blcotensor* at_host, blcotensor* at_dev, KruskalModel* M,
                                    FType* output, int kernel, int target_mode, IType thread_cf,
                                    bool stream_data, int nprtn, FType** mats_staging_ptr,
                                    FType** mats_dev, FType** partial_matrices,
                                    FType** partial_matrices_dev)
{
    mttkrp_alto_dev_onemode_ref(at_host, at_dev, M, output, kernel, target_mode, thread_cf,
                                stream_data, nprtn, mats_staging_ptr, mats_dev, partial_matrices,
                                partial_matrices_dev);
}

// The same from explicit arguments: rank and device factors (f64).
void mttkrp_alto_dev_onemode_any(blcotensor* at_dev, IType rank, FType* output, int kernel,
                                 int target_mode, IType thread_cf, int nprtn, FType** mats_dev);
