// blco.hpp -- the BLCO tensor API of the reference (include/blco.hpp:11-193)
// on top of the C ABI in blco_host.h / blco_hip.h.
#pragma once

#include <hip/hip_runtime_api.h>

#include "alto.hpp"
#include "common.hpp"
#include "sptensor.hpp"

struct blcotensor;
struct blco_host;
struct blco_dev;

// One block: nonzeros sharing the upper key bits (blco.hpp:11-20).  As in
// the reference it extends the ALTO record: nmode, nnz, idx, vals, and the
// parent's dims, mode_masks and mode_pos.  Host blocks point into the
// parent's arrays; the blocks of a device tensor point at its device arrays
// when the tensor is resident (idx and vals are null for a streamed tensor,
// whose blocks pass through device slots).
struct blco_block : public AltoTensor<IType> {
    IType* block_coords = nullptr;  // nmode upper-bit coordinates (host)
    // lvl3 partial matrices of the reference's kernels; this library keeps
    // its own with the device tensor, so these stay null
    FType** pmatrices_staging_ptr = nullptr;
    FType** pmatrices = nullptr;
    blcotensor* parent = nullptr;
};

struct blcotensor {
    IType N = 0;
    IType* modes = nullptr;
    IType* mode_masks = nullptr;
    int* mode_pos = nullptr;
    IType total_nnz = 0;
    IType max_nnz = 0;
    IType block_count = 0;
    IType max_block_size = 0;
    blco_block** blocks = nullptr;
    // Device copies of the block records for the reference's batched
    // kernels (src/blco.cu:12-58) and their tile / warp tables; this
    // library's kernels take the device tensor handle (dev) and build their
    // own chunk tables, so these stay null.
    blco_block** blocks_dev_staging = nullptr;
    blco_block** blocks_dev_ptr = nullptr;
    IType* tile_info = nullptr;
    IType* coords = nullptr;  // total_nnz BLCO coordinates (host)
    FType* values = nullptr;  // total_nnz values (host)
    hipStream_t* streams = nullptr;
    IType* warp_info = nullptr;
    IType* warp_info_gpu = nullptr;
    IType warp_info_length = 0;
    // implementation handles
    blco_host* host = nullptr;  // host tensors
    blco_dev* dev = nullptr;    // device tensors
    bool on_device = false;
};

// gen_blcotensor_host<LIT> (blco.hpp:196-331).  LIT selects the reference's
// key width (unsigned long long for cpd64, unsigned __int128 for cpd128);
// keys wider than 64 bits are accepted with either.
template <typename LIT>
blcotensor* gen_blcotensor_host(SparseTensor* tensor, IType max_block_size);
blcotensor* gen_blcotensor_host_any(SparseTensor* tensor, IType max_block_size);
template <typename LIT>
inline blcotensor* gen_blcotensor_host(SparseTensor* tensor, IType max_block_size)
{
    return gen_blcotensor_host_any(tensor, max_block_size);
}

// gen_blcotensor_device (src/blco.cu:12-58): MB == 0 keeps every block
// resident; MB > 0 streams blocks through device slots of MB nonzeros.
// The device copy holds values as f64 (value_dtype 0) or f32 (1).
blcotensor* gen_blcotensor_device(blcotensor* X_host, IType block_count, IType MB,
                                  bool do_batching, int value_dtype = 0);
void delete_blcotensor_host(blcotensor* tensor);
void delete_blcotensor_device(blcotensor* tensor);
// The upload happens in gen_blcotensor_device; these keep the reference's
// call sequence valid (they only wait for the copies).
void send_blcotensor_over(hipStream_t stream, blcotensor* X_gpu, blcotensor* X);
void send_block_over(hipStream_t stream, blco_block* gpu_block, blco_block* cpu_block);
blco_block* gen_block_host(IType N, IType nnz, blcotensor* parent);
blco_block* gen_block_device(IType N, IType nnz, blcotensor* parent);
void delete_block_host(blco_block* block);
void delete_block_device(blco_block* block);
