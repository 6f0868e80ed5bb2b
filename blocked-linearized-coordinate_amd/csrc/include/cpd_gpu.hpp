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
// Device budget (bytes) for the nonzeros of every device tensor that
// cpals_blco_dev and the MTTKRP drivers create: a tensor that fits stays
// resident, a larger one keeps what fits resident and streams the rest
// through two slots of slot_nnz nonzeros (0: no budget).
void cpals_set_device_budget(uint64_t bytes, uint64_t slot_nnz);

FType cpals_blco_dev(blcotensor* X, KruskalModel* A, int maxiter, FType tolerance, int kernel,
                     bool stream_data, bool do_batching, int thread_cf, int nprtn);

// The reference's ALS helpers (include/cpd_gpu.hpp:25-140) with the same
// arguments, minus the cuBLAS / cuSOLVER handles.  Matrices are row-major
// f64 device arrays; grams is a host array of N device pointers.

// kruskal_fit_gpu (:48): fit = 1 - sqrt(|X_norm^2 + ||M||^2 - 2 <X, M>|) /
// X_norm from the Grams, lambda and the last mode's factor A->U[mode]
// against its MTTKRP result last_mttkrp (src/cpd_gpu.cu:333-384).  A is the
// device model (make_device_copy: U[n] and lambda are device arrays), X the
// host tensor (mode lengths).  work and blocks are accepted for the
// reference's call sequence; the library sizes its own grid and scratch.
// Synchronises the stream.
FType kruskal_fit_gpu(hipStream_t v_stream, FType** grams, KruskalModel* A, blcotensor* X,
                      FType* work, FType* last_mttkrp, FType X_norm, IType mode, IType blocks);
void vector_hadamard(hipStream_t stream, FType* x, FType* y, IType n);
void value_fill(hipStream_t stream, FType* x, IType n, FType val);
void vecsum(hipStream_t stream, FType* x, IType n);
// pseudoinverse_gpu (:109-110): A (n x n, device) is overwritten with its
// Moore-Penrose pseudo-inverse (singular values at or below
// n * (nextafter(s_max) - s_max) dropped, src/cpd_gpu.cu:487-489), by a
// one-sided Jacobi SVD in one workgroup (n <= 64).  info (device int, may be
// null): 0 when the SVD converged.  work / lwork are accepted for the
// reference's call sequence and not used.  Synchronises the stream.
void pseudoinverse_gpu(hipStream_t stream, FType* A, IType n, FType* work, IType lwork, int* info);
// columnwise_normscal (:124): lambda_r = ||mat[:, r]||_2 (overwritten, as
// the reference documents; its kernel adds onto lambda's previous content),
// columns scaled by 1 / lambda_r (zero columns stay zero).  blocks is
// accepted for the reference's call sequence.
void columnwise_normscal(hipStream_t stream, IType blocks, FType* mat, IType rank, IType mode_length,
                         FType* lambda);
// ata_gpu (:140): C (n x n) = A^T A for A (m x n) on MFMA.
void ata_gpu(hipStream_t stream, const FType* A, FType* C, IType m, IType n);

// The CP-ALS step helpers this library adds beside them:
// U = M V^{-1} (Cholesky solve, pseudo-inverse fallback) for V = the
// Hadamard product of the grams other than `mode`; returns 1 when the
// fallback was used (blco_solve).
int normal_solve_gpu(hipStream_t stream, FType** grams, IType nmodes, IType mode, IType rank, FType* M,
                     FType* U, IType rows);
// columnwise_normscal with the norm kind of KruskalModelNorm (L2 or max).
void columnwise_normscal_which(hipStream_t stream, FType* mat, IType rank, IType mode_length,
                               FType* lambda, mat_norm_type which);
// kruskal_fit_gpu from explicit sizes (no blcotensor).
FType kruskal_fit_gpu_rows(hipStream_t stream, FType** grams, KruskalModel* A_dev, IType nmodes,
                           FType* last_mttkrp, IType last_rows, FType X_norm, IType mode);
