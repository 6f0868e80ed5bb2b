// cpd_gpu.hpp -- GPU CP-ALS API of the reference (include/cpd_gpu.hpp:22-140)
// with its cuBLAS / cuSOLVER handle parameters replaced by a HIP stream.
#pragma once

#include <hip/hip_runtime_api.h>

#include "blco.hpp"
#include "kruskal_model.hpp"

// Precision of the device factors for cpals_blco_dev: 0 = f64 (plain form),
// 1 = f32, 2 = bf16 storage (f32 arithmetic); solve, lambda and fit are f64.
void cpals_set_precision(int precision);
// Mode order: 0 = 0..N-1 (CPU reference), 1 = longest first (GPU reference).
void cpals_set_mode_order(int order);
// Capture the ALS sweep in a hipGraph and replay it (1) or launch eagerly (0).
void cpals_set_graph(int use_graph);

FType cpals_blco_dev(blcotensor* X, KruskalModel* A, int maxiter, FType tolerance, int kernel,
                     bool stream_data, bool do_batching, int thread_cf, int nprtn);

// Device-array forms of the ALS helpers (row-major f64 matrices).
FType kruskal_fit_gpu(hipStream_t stream, FType** grams, KruskalModel* A_dev, IType nmodes,
                      FType* last_mttkrp, IType last_rows, FType X_norm, IType mode);
void vector_hadamard(hipStream_t stream, FType* x, FType* y, IType n);
void value_fill(hipStream_t stream, FType* x, IType n, FType val);
void vecsum(hipStream_t stream, FType* x, IType n);
// U = M * V^+ (Cholesky solve, pseudo-inverse fallback) for V = the Hadamard
// product of grams other than `mode`; returns 1 when the fallback was used.
int pseudoinverse_gpu(hipStream_t stream, FType** grams, IType nmodes, IType mode, IType rank,
                      FType* M, FType* U, IType rows);
void columnwise_normscal(hipStream_t stream, FType* mat, IType rank, IType mode_length,
                         FType* lambda, mat_norm_type which = MAT_NORM_2);
void ata_gpu(hipStream_t stream, const FType* A, FType* C, IType m, IType n);
