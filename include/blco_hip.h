/*
 * blco_hip.h -- C ABI of the MI355X (gfx950) side of the BLCO CP-ALS path:
 * device tensors, MTTKRP kernels, the ALS step and the CP-ALS driver.
 *
 * Status codes, blco_last_error() and the host tensor handle are declared in
 * blco_host.h.  Device pointers are plain `void*` / `double*` into device
 * memory; `stream` is a hipStream_t passed as void* (NULL = the library's own
 * stream for that object).  Every entry point checks its arguments on the
 * host before any kernel launch and returns an error instead of launching.
 */
#ifndef BLCO_HIP_H_
#define BLCO_HIP_H_

#include <stdint.h>

#include "blco_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of values / factors. */
enum blco_dtype { BLCO_F64 = 0, BLCO_F32 = 1, BLCO_BF16 = 2 };

/* Kernel ids, the reference's -k values (src/main.cpp:88,
 * src/alto_dev.cu:1954-2032). */
enum blco_kernel {
    BLCO_K_LVL1 = 1,       /* wave64 segmented reduction, one launch per block */
    BLCO_K_LVL2 = 2,       /* row window (LDS window + hot rows, partial rows
                              folded by a reduce), one launch per block; takes
                              the place of the reference's lvl2 stash kernel */
    BLCO_K_LVL3 = 3,       /* LDS-privatised output + nprtn partial matrices;
                              modes too long for LDS run kernel 12 */
    BLCO_K_AUTO = 10,      /* the reference rule: I_n <= 100 (and fits) -> 3,
                              else 14 (12 when streamed or when the copy
                              does not fit in device memory) */
    BLCO_K_BATCHED = 11,   /* segmented lvl1, all blocks in one launch */
    BLCO_K_BATCHED2 = 12,  /* row window, all blocks in one launch */
    BLCO_K_ATOMIC = 13,    /* one HBM atomic per (nonzero, column) */
    BLCO_K_SORTED = 14,    /* segmented reduction over a row-ordered copy of
                              the tensor (each block's nonzeros sorted by the
                              output index; built on first use per mode,
                              nnz * (8 + value size) more bytes per mode);
                              streamed tensors run kernel 12 */
    BLCO_K_SELECT = 20     /* per-mode choice by the installed selector
                              (the rule of 10 when none is installed) */
};

/* Number of visible GPUs (0 when none); never initialises a context on a
 * machine without /dev/kfd. */
int blco_device_count(void);

/* Device memory for callers without their own allocator (plain hipMalloc /
 * hipMemcpy on the null stream; the copies wait for all prior device work). */
int blco_malloc(void** ptr, uint64_t bytes);
int blco_free(void* ptr);
int blco_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int blco_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy (the runtime's copy kernel; its HBM counters for a
 * known size calibrate the traffic figures of the MTTKRP kernels). */
int blco_memcpy_dtod(void* dst, const void* src, uint64_t bytes);
int blco_memset_zero(void* ptr, uint64_t bytes);
int blco_synchronize(void);

/* ------------------------------------------------------- device encoder */

/* gen_blcotensor_host<LIT> (include/blco.hpp:196-331) computed on the GPU:
 * the same blco_host as blco_host_encode, bit for bit (ALTO keys formed and
 * radix-sorted on the device, BLCO words re-encoded there; blocks built on
 * the host from the sorted keys' upper halves).  Arguments as for
 * blco_host_encode: host arrays, cidx mode-major (nmodes x nnz). */
int blco_device_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                       const uint64_t* cidx, const double* vals,
                       uint64_t max_block_size, blco_host** out);

/* ---------------------------------------------------------- device tensor */

typedef struct blco_dev blco_dev;

/* gen_blcotensor_device + send_blcotensor_over + send_masks_over
 * (src/blco.cu:12-58, :120-138; src/alto_dev.cu:1849-1853).  value_dtype is
 * BLCO_F64 or BLCO_F32 (values converted on upload).  stream_data = 0 keeps
 * every block resident in HBM; 1 keeps the tensor in pinned host memory and
 * streams blocks of at most slot_nnz nonzeros through two device slots per
 * MTTKRP (double-buffered copies on a side stream). */
int blco_dev_create(const blco_host* h, int value_dtype, int stream_data,
                    uint64_t slot_nnz, blco_dev** out);
/* Same from plain arrays (a BLCO tensor built elsewhere). */
int blco_dev_create_arrays(int nmodes, const uint64_t* dims,
                           const uint64_t* masks, const int* pos, uint64_t nnz,
                           const uint64_t* coords, const double* vals,
                           uint64_t nblocks, const uint64_t* block_start,
                           const uint64_t* block_nnz,
                           const uint64_t* block_coords, int value_dtype,
                           int stream_data, uint64_t slot_nnz, blco_dev** out);
/* gen_blcotensor_device under a device-memory budget (no counterpart
 * flag in the reference, whose --stream-data streams every block through
 * NUM_STREAMS slots, src/alto_dev.cu:1869-1874, 1987).  device_budget caps
 * the device bytes that hold nonzeros: a tensor whose nonzeros fit
 * (nnz * (8 + value size) <= device_budget) is resident as with
 * blco_dev_create(h, value_dtype, 0, 0, out); a larger one keeps the whole
 * tensor in pinned host memory, keeps its leading nonzeros (ALTO order) on
 * the device as far as the budget allows beside two slots of slot_nnz
 * nonzeros (0: 2^24), and streams the rest through those slots.  The
 * budget also caps kernel 14's row-ordered copies (each nnz * (8 + value
 * size) bytes): a copy that would exceed it is not built and the mode runs
 * the row window instead.  BLCO_EINVAL when the two slots alone exceed
 * the budget. */
int blco_dev_create_budget(const blco_host* h, int value_dtype, uint64_t device_budget,
                           uint64_t slot_nnz, blco_dev** out);
/* Nonzeros held on the device and streamed from host memory per MTTKRP,
 * and the device bytes the tensor holds (blco_dev_bytes). */
int blco_dev_residency(const blco_dev* d, uint64_t* resident_nnz, uint64_t* streamed_nnz,
                       uint64_t* device_bytes);
/* Launch plan of kernel 14 (the row-ordered MTTKRP) for a mode after it has
 * run there at this rank and factor dtype (no counterpart in the reference,
 * whose kernels have no such plan): chunks of the copy per wave, whether a
 * wave's chunks lie one grid width apart (1) or next to each other (0), and
 * the hot factor rows the workgroups keep in LDS.  BLCO_EINVAL before the
 * first such run. */
int blco_dev_rows_plan(const blco_dev* d, int mode, int rank, int factor_dtype,
                       uint32_t* chunks_per_wave, int* interleaved, uint64_t* lds_rows);
/* delete_blcotensor_device, src/blco.cu:74-95 */
void blco_dev_free(blco_dev* d);
/* Bytes of device memory held by the tensor. */
uint64_t blco_dev_bytes(const blco_dev* d);
/* Device arrays of a resident tensor (the blocks of gen_blcotensor_device
 * point into them, src/blco.cu:12-58): BLCO coordinates (u64[nnz], block
 * after block), values (value_dtype elements) and the block coordinates
 * (u64[nblocks * nmodes]).  All three are NULL for a streamed tensor, whose
 * blocks stay in pinned host memory. */
int blco_dev_arrays(const blco_dev* d, uint64_t** coords, void** vals, int* value_dtype,
                    uint64_t** block_coords);

/* The synthetic Poisson tensor of blco_generate (blco_host.h) generated,
 * encoded (gen_blcotensor_host semantics, blocks of at most max_block_size
 * nonzeros per upper key, 0 = unbounded) and made resident on the GPU
 * without a host copy of the tensor: the same tensor as blco_generate +
 * blco_host_encode + blco_dev_create(value_dtype, resident) (no reference
 * counterpart: the reference generates on the host,
 * src/poisson_generator.cpp).  Keys of at most 72 bits. */
int blco_generate_device(const blco_gen_spec* spec, uint64_t max_block_size,
                         int value_dtype, blco_dev** out);
/* Shape of a device tensor: nmodes, dims (nmodes entries), nonzeros and
 * blocks (any pointer may be NULL). */
int blco_dev_info(const blco_dev* d, int* nmodes, uint64_t* dims, uint64_t* nnz, uint64_t* nblocks);
/* Host copy of a resident device tensor (values as f64). */
int blco_dev_to_host(const blco_dev* d, blco_host** out);
/* Every step-th nonzero (in BLCO order) of a resident tensor, decoded to COO
 * on the host: a uniform sample for CPU-side baselines. */
int blco_dev_sample(const blco_dev* d, uint64_t step, blco_coo** out);
/* Converts the resident values to value_dtype (BLCO_F64 or BLCO_F32) in
 * place; row-ordered copies are dropped (rebuilt on their next use). */
int blco_dev_set_value_dtype(blco_dev* d, int value_dtype);

/* ----------------------------------------------------------------- MTTKRP */

/* mttkrp_alto_dev_onemode, src/alto_dev.cu:1954-2032:
 * out[i_mode, :] += sum over nonzeros of val * prod_{m != mode} U_m[i_m, :].
 * factors: nmodes device pointers (row-major dims[m] x rank) of factor_dtype
 * (BLCO_F64, BLCO_F32 or BLCO_BF16); out: device dims[mode] x rank, f64 for
 * BLCO_F64 factors, f32 otherwise; it is accumulated into (the caller
 * zeroes it, src/cpd_gpu.cu:155).  nprtn: lvl3 partial-matrix count (0 =
 * default); thread_cf: tiles of 64 nonzeros per wave (0 = default: 4, and
 * for kernel 14 16, or 2/4/8 for tensors of up to 2^20/2^21/2^22
 * nonzeros, or 32 for FP64 factors at rank >= 32 from 2^26 nonzeros). */
int blco_mttkrp(blco_dev* d, int mode, int rank, int factor_dtype,
                void* const* factors, void* out, int kernel, int nprtn,
                int thread_cf, void* stream);

/* The row-ordered copy behind kernel 14 for (mode, rank, factor_dtype),
 * built if it is not (on the GPU; with BLCO_ROWORDER=host in the
 * environment on the host), for inspection: info[5] = tiles, inner mode (-1:
 * ALTO order within rows), fiber kernel on (0/1), groups, run starts (with
 * the final nnz); then (each may be NULL) its BLCO words (nnz), values
 * (nnz, the tensor's value type), run starts (info[4]) and groups (4 per
 * group: first nonzero, nonzeros, block, tile). */
int blco_dev_row_copy(blco_dev* d, int mode, int rank, int factor_dtype, uint64_t* info,
                      uint64_t* coords, void* vals, uint64_t* runs, uint64_t* groups);

/* Kernel the AUTO heuristic would run for (mode, rank)
 * (src/alto_dev.cu:1975,1996-2001). */
int blco_auto_kernel(const blco_dev* d, int mode, int rank);

/* --------------------------------------------------------------- ALS step */

/* ata_gpu, src/cpd_gpu.cu:547-568: G = U^T U (rank x rank, f64, row-major)
 * on MFMA from a rows x rank factor of dtype (f64 via f64 MFMA, f32/bf16 via
 * f32 MFMA with f64 partial sums). */
int blco_gram(const void* U, int dtype, uint64_t rows, int rank, double* G,
              void* stream);
/* vector_hadamard, src/cpd_gpu.cu:397-404: x <- x .* y (f64, n entries). */
int blco_hadamard(double* x, const double* y, uint64_t n, void* stream);
/* value_fill, src/cpd_gpu.cu:408-415 */
int blco_value_fill(double* x, uint64_t n, double val, void* stream);
/* vecsum, src/cpd_gpu.cu:419-448: x[0] <- sum(x[0..n)); rest is scratch. */
int blco_vecsum(double* x, uint64_t n, void* stream);

/* Normal-equation solve of one ALS mode update, replacing the reference's
 * pseudoinverse_gpu + Dgemm (src/cpd_gpu.cu:160-177, :452-532) with the CPU
 * reference's Cholesky / minimum-norm fallback (include/cpd.hpp:547-648):
 * V = Hadamard of grams[m] for m != mode (nmodes device pointers, f64
 * rank x rank); U = M V^{-1} row by row (Cholesky, forward/backward
 * substitution), or M V^+ when a Cholesky pivot falls to
 * rank * eps * max V_ii or below (symmetric pseudo-inverse, tolerance
 * rank * ulp(max |eig|)), eps and ulp of f64 for every dtype.  M
 * (rows x rank, f64 for dtype BLCO_F64, else f32) is read, U (dtype) written.
 * rank <= 64.  *used_fallback (host, may be NULL) receives 1 when the
 * pseudo-inverse was used (this reads back one flag and synchronises the
 * stream). */
int blco_solve(double* const* grams, int nmodes, int mode, int rank,
               const void* M, void* U, int dtype, uint64_t rows,
               int* used_fallback, void* stream);

/* blco_solve without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results as blco_solve; the solve's scratch is a
 * file-scope device array of blco_solve_noalloc_scratch_bytes() bytes per
 * GPU (held until the process ends), which bounds rank to 64.  Asynchronous
 * on `stream` unless used_fallback is non-NULL (then the flag is read back
 * and the stream synchronised).  Calls that overlap on one GPU share that
 * array and race: serialise them. */
int blco_solve_noalloc(double* const* grams, int nmodes, int mode, int rank,
                       const void* M, void* U, int dtype, uint64_t rows,
                       int* used_fallback, void* stream);
uint64_t blco_solve_noalloc_scratch_bytes(void);

/* pseudoinverse_gpu, src/cpd_gpu.cu:452-532: A (device, n x n row-major
 * f64, n <= 64) is overwritten with its Moore-Penrose pseudo-inverse; the
 * singular values at or below n * (nextafter(s_max, +inf) - s_max) are
 * dropped (src/cpd_gpu.cu:487-489).  One-sided Jacobi SVD in one
 * workgroup.  info (device int, may be NULL): 0 when the SVD converged,
 * else the sweep count.  Synchronises the stream. */
int blco_pinv(double* A, uint64_t n, int* info, void* stream);

/* Column normalisation of one factor (KruskalModelNorm,
 * src/kruskal_model.cpp:142-259; columnwise_normscal, src/cpd_gpu.cu:244-304
 * without the lambda carry-over): which 0 = L2 (lambda_r = ||U_:r||, zero
 * columns left at zero), 1 = max (lambda_r = max(1, max_i U_ir)).  U scaled
 * in place; lambda (device f64, rank) overwritten. */
int blco_colnorm(void* U, int dtype, uint64_t rows, int rank, int which,
                 double* lambda, void* stream);

/* kruskal_fit_gpu, src/cpd_gpu.cu:333-384 with the CPU reference's residual
 * rule (include/cpd.hpp:190-272): fit = 1 - r / ||X||, r^2 = ||X||^2 +
 * |sum_ij lambda_i lambda_j prod_n G_n(i,j)| - 2 sum_r lambda_r
 * sum_i U_ir M_ir (r taken as is when r^2 <= 0).  U_last / M_last are the
 * last updated factor and its MTTKRP (rows x rank, dtype).  Result to *fit
 * (host; synchronises the stream). */
int blco_fit(double* const* grams, int nmodes, int rank, const double* lambda,
             const void* U_last, const void* M_last, int dtype, uint64_t rows,
             double normX2, double* fit, void* stream);

/* ---------------------------------------------------------------- CP-ALS */

/* Options of cpals_blco_dev (include/cpd_gpu.hpp:22-23). */
typedef struct blco_cpals_opts {
    int rank;
    int max_iters;
    double tolerance;   /* stop when i > 0 and |fit - fit_old| < tolerance */
    int kernel;         /* enum blco_kernel */
    int precision;      /* BLCO_F64 (plain form) or BLCO_F32 / BLCO_BF16 factors */
    int nprtn;          /* lvl3 partial matrices (0 = default) */
    int thread_cf;      /* tiles per wave (0 = default) */
    int mode_order;     /* 0: modes 0..N-1 (CPU reference), 1: longest first (GPU reference) */
    int use_graph;      /* 1: capture one ALS sweep in a hipGraph and replay it */
    int verbose;        /* 1: print the reference's per-iteration lines */
    int solve_floor;    /* 0 (default): the solve's pivot floor and pseudo-inverse
                           tolerance at f64 precision for every form (the FP64
                           oracle's decisions); 1: at the factors' precision
                           (f32's for BLCO_F32 / BLCO_BF16), so a V singular to
                           that precision takes the pseudo-inverse whatever the
                           MTTKRP's summation order */
} blco_cpals_opts;

typedef struct blco_cpals_state blco_cpals_state; /* device-resident solver state */

/* Uploads the initial factors U[n] (host f64, dims[n] x rank) and prepares
 * the solver: grams, work buffers, ||X||^2. */
int blco_cpals_create(blco_dev* d, const blco_cpals_opts* opts,
                      const double* const* U, blco_cpals_state** out);
/* Runs up to n_iters ALS iterations (stopping early on the tolerance test),
 * appending one fit per iteration to fits (host, may be NULL); returns the
 * number of iterations run in *done. */
int blco_cpals_iterate(blco_cpals_state* s, int n_iters, double* fits, int* done);
/* Factors (dims[n] x rank, converted to f64) and lambda back to the host. */
int blco_cpals_get(blco_cpals_state* s, double* const* U, double* lambda);
/* Accumulated device time of the MTTKRP launches per mode (ms, nmodes
 * entries, measured with HIP events on the solver stream), of whole
 * iterations, the iteration count, and the kernel chosen per mode. */
int blco_cpals_times(const blco_cpals_state* s, double* mttkrp_ms, double* iter_ms,
                     int* iters, int* kernels);
void blco_cpals_reset_times(blco_cpals_state* s);
void blco_cpals_free(blco_cpals_state* s);

/* cpals_blco_dev, src/cpd_gpu.cu:21-242, in one call: create, iterate up to
 * max_iters, copy back.  fit_trace (host, max_iters entries, may be NULL);
 * returns iterations run in *iters and the final fit in *final_fit. */
int blco_cpals(blco_dev* d, const blco_cpals_opts* opts, double* const* U,
               double* lambda, double* fit_trace, int* iters,
               double* final_fit);

/* ---- many small tensors decomposed together ----
 * (no reference counterpart: the reference decomposes one tensor per run,
 * src/main.cpp:322-390).  A batch refers to resident device tensors owned
 * by the caller (freed after the batch).  Mode n of the batch is mode n of
 * every tensor that has it.  Factor and output tables are indexed by batch
 * position: factors[t * BLCO_MAX_MODES + m], outs[t]. */
typedef struct blco_batch blco_batch;
int blco_batch_create(int ntensors, blco_dev* const* devs, blco_batch** out);
void blco_batch_free(blco_batch* b);
/* Times the persistent MTTKRP of `mode` over every (nonzeros per chunk /
 * 64, waves per workgroup, workgroups per CU) in {1,2,4,8}^3 (at most 32
 * waves per CU) and keeps the fastest for (mode, rank, factor type).  The
 * outputs are overwritten with accumulated results. */
int blco_batch_tune(blco_batch* b, int mode, int rank, int factor_dtype,
                    void* const* factors, void* const* outs, int* tpw,
                    int* waves, int* wgs_per_cu, float* ms);
/* One persistent launch: outs[t] += MTTKRP(mode) of every tensor t that has
 * the mode.  A zero shape argument takes the tuned shape (else 4, 4, 4).
 * Synchronous. */
int blco_batch_mttkrp(blco_batch* b, int mode, int rank, int factor_dtype,
                      void* const* factors, void* const* outs, int tpw,
                      int waves, int wgs_per_cu);
/* CP-ALS of every tensor (cpals_blco_dev semantics per tensor, modes
 * 0..N-1): per mode one persistent MTTKRP for the whole batch, the ALS steps
 * of all tensors on their own streams, and after iteration 0 the sweep
 * replayed as one captured hipGraph.  U (host f64, U[t * BLCO_MAX_MODES + m],
 * dims x rank) holds the initial factors and receives the results; lambda
 * (t * rank + r), fits[t] and iters[t] (may be NULL) receive each tensor's
 * result at its own stopping iteration.  *iter_ms: mean device time of one
 * sweep of the whole batch after iteration 0. */
int blco_batch_cpals(blco_batch* b, const blco_cpals_opts* opts, double* const* U,
                     double* lambda, double* fits, int* iters, double* iter_ms);

/* Event-timed MTTKRP: one untimed call (builds tables and scratch), then
 * `reps` calls timed with HIP events on the tensor's own stream; *ms is the
 * mean time of one call.  The output accumulates reps + 1 results.  The
 * reference's -p benchmark (src/main.cpp:336-378, src/alto_dev.cu:1856-1952)
 * times the same way; this is also the timing source of the selector's
 * training data. */
int blco_mttkrp_time(blco_dev* d, int mode, int rank, int factor_dtype,
                     void* const* factors, void* out, int kernel, int nprtn,
                     int thread_cf, int reps, float* ms, int* used_kernel);

/* ---- a small MLP run and trained with the library's own HIP kernels ----
 * (no reference counterpart: the reference's kernel choice is the fixed
 * rule of src/alto_dev.cu:1975,1996-2001).  Layers are dense, ReLU between
 * hidden layers and a linear output; widths[0] inputs, widths[nlayers]
 * outputs.  Parameters are flat f32, layer by layer: W (out x in, row-major)
 * then b.  Host arrays are row-major, one row per example. */
#define BLCO_MLP_MAX_LAYERS 4
#define BLCO_MLP_MAX_WIDTH 256
typedef struct blco_mlp blco_mlp;
/* He-uniform weights from the counter generator keyed by seed, zero biases. */
int blco_mlp_create(int nlayers, const int* widths, uint64_t seed, blco_mlp** out);
void blco_mlp_free(blco_mlp* h);
uint64_t blco_mlp_nparams(const blco_mlp* h);
/* Copies parameters and/or the last gradients (either may be NULL). */
int blco_mlp_get_params(const blco_mlp* h, float* params, float* grads);
/* Sets the parameters and resets the Adam state. */
int blco_mlp_set_params(blco_mlp* h, const float* params);
int blco_mlp_forward(blco_mlp* h, const float* X, int batch, float* Y);
/* Masked mean squared error sum(M (Y - T)^2) / sum(M) of the batch and its
 * gradient with respect to every parameter (kept on the device). */
int blco_mlp_loss_grad(blco_mlp* h, const float* X, const float* T,
                       const float* M, int batch, float* loss);
/* One Adam step (bias-corrected) with the last gradients. */
int blco_mlp_adam_step(blco_mlp* h, float lr, float beta1, float beta2, float eps);
/* Mini-batch Adam (0.9, 0.999, 1e-8) over n examples for `epochs` epochs,
 * shuffled by the counter generator keyed by seed; loss_trace (epochs
 * entries, may be NULL) gets each epoch's mean batch loss. */
int blco_mlp_train(blco_mlp* h, const float* X, const float* T, const float* M,
                   int n, int epochs, int batch, float lr, uint64_t seed,
                   float* loss_trace);

/* ---- learned MTTKRP kernel selector (kernel id 20) ----
 * Candidates c = 0..BLCO_SEL_NCAND-1: 0 row window (12), 1 segmented lvl1
 * (11), 2 atomic (13), 3..9 lvl3 with 1, 2, 4, 8, 16, 32, 64 partitions,
 * 10 row-ordered (14), and the thread coarsening of the reference's
 * --thread-cf (tiles of 64 nonzeros per wave, src/main.cpp:127) as
 * candidates of their own: 11, 12 segmented lvl1 with 1 and 16 tiles per
 * wave (4 otherwise), 13, 14 row-ordered with 8 and 32 (16 otherwise).
 * The network maps BLCO_SEL_NFEAT features (normalised by the installed mean
 * and scale) to predicted log2 times of every candidate; the smallest
 * prediction among the candidates that can run is taken.  With no network
 * installed, kernel 20 is kernel 10.  A selector file next to the library
 * (lib/../selector/selector.bin, or $BLCO_SELECTOR) is installed on first
 * use. */
#define BLCO_SEL_NFEAT 24
#define BLCO_SEL_NCAND 15
/* Candidate c -> kernel id, lvl3 partitions, tiles per wave (0: the
 * kernel's default). */
int blco_sel_candidate(int c, int* kernel, int* nprtn, int* thread_cf);
/* 1 when candidate c can run for (mode, rank, factor type), else 0. */
int blco_sel_valid(const blco_dev* d, int mode, int rank, int factor_dtype, int c);
int blco_mode_features(const blco_dev* d, int mode, int rank, int factor_dtype,
                       float* feat);
/* nlayers = 0 uninstalls. */
int blco_selector_install(int nlayers, const int* widths, const float* params,
                          const float* xmean, const float* xstd);
int blco_selector_load(const char* path);
int blco_selector_installed(void);
/* The kernel id 20 would run for (mode, rank, factor type), its lvl3
 * partition count and tiles per wave (0 = the caller's). */
int blco_select_kernel(blco_dev* d, int mode, int rank, int factor_dtype,
                       int* kernel, int* nprtn, int* thread_cf);

#ifdef __cplusplus
}
#endif
#endif
