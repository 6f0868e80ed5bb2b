/*
 * blco_host.h -- C ABI of the host side of the BLCO CP-ALS path:
 * tensor generation, ALTO/BLCO encoding and factor initialisation.
 *
 * All functions return BLCO_OK (0) or a negative status; blco_last_error()
 * gives the message of the last failure on the calling thread.  Arrays are
 * plain pointers with explicit sizes; coordinate arrays are mode-major
 * (cidx[n * nnz + i] is the mode-n index of nonzero i), factor matrices are
 * row-major (rows x rank).
 */
#ifndef BLCO_HOST_H_
#define BLCO_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLCO_MAX_MODES 5

enum blco_status {
    BLCO_OK = 0,
    BLCO_EINVAL = -1,       /* bad argument */
    BLCO_EHIP = -2,         /* HIP runtime error */
    BLCO_EUNSUPPORTED = -3, /* valid request this build does not support */
    BLCO_ENOMEM = -4,       /* host or device allocation failed */
    BLCO_ECAPACITY = -5,    /* an output array is too small */
    BLCO_ENODEV = -6        /* no GPU visible */
};

/* Message of the last failing call on this thread ("" if none). */
const char* blco_last_error(void);

/* ------------------------------------------------------------------ COO */

typedef struct blco_coo blco_coo; /* owned sparse tensor in COO form */

/* CreateSparseTensor, src/sptensor.cpp:251-297 (copying constructor; the
 * reference takes nonzero-major cidx, this one mode-major). */
int blco_coo_create(int nmodes, const uint64_t* dims, uint64_t nnz,
                    const uint64_t* cidx, const double* vals, blco_coo** out);
/* ImportSparseTensor, src/sptensor.cpp:162-249 (format 0 = .tns text with
 * 1-based indices, 1 = .bin: u64 nmodes, u64 dims[nmodes], u64 nnz, then
 * cidx[mode][nnz] u64 and vals f64). */
int blco_coo_import(const char* path, int format, blco_coo** out);
/* ExportSparseTensor, src/sptensor.cpp:22-71 (.tns values use %g). */
int blco_coo_export(const blco_coo* t, const char* path, int format);
int blco_coo_info(const blco_coo* t, int* nmodes, uint64_t* dims,
                  uint64_t* nnz);
const uint64_t* blco_coo_cidx(const blco_coo* t); /* mode-major */
const double* blco_coo_vals(const blco_coo* t);
void blco_coo_free(blco_coo* t); /* DestroySparseTensor, sptensor.cpp:11-20 */

/* ------------------------------------------------------------ generator */

/* Synthetic Poisson tensor (replaces PoissonGeneratorRun,
 * src/poisson_generator.cpp:336-376, with a thread-count independent
 * counter-based generator).  Indices per mode are uniform (zipf_s[n] == 0) or
 * Zipf(s) over ranks 1..dims[n] mapped to row rank-1; duplicates are merged
 * keeping the first draw.  With topup, more draws are made until target_nnz
 * unique coordinates exist (the first target_nnz in draw order are kept).
 * Values are zero-truncated Poisson counts whose rate is
 * sum_r prod_n A_n(i_n, r) with Gamma(1,1) entries in A_n (dims[n] x
 * gen_rank); value_kind 0 stores the count, 1 stores log1p(count). */
typedef struct blco_gen_spec {
    int nmodes;
    uint64_t dims[BLCO_MAX_MODES];
    double zipf_s[BLCO_MAX_MODES];
    uint64_t target_nnz;
    int topup;
    int gen_rank;
    int value_kind;
    uint64_t seed;
} blco_gen_spec;

int blco_generate(const blco_gen_spec* spec, blco_coo** out);

/* Counter-based uniform [0,1) factor initialisation keyed by (seed, mode,
 * row, column), lambda = 1 (replaces KruskalModelRandomInit,
 * src/kruskal_model.cpp:93-110, whose output depends on the thread count). */
int blco_init_factor(uint64_t seed, int mode, uint64_t rows, int rank,
                     double* U);

/* ------------------------------------------------------------ encoding */

typedef struct blco_host blco_host; /* BLCO tensor in host memory */

/* gen_blcotensor_host<LIT>, include/blco.hpp:196-331 (with gen_alto,
 * include/alto.hpp:107-186).  lit_bits 64 (cpd64) or 128 (cpd128); a key of
 * more than 64 bits is accepted with lit_bits 64 too, encoded through a
 * 128-bit intermediate.  max_block_size 0 = one block per upper-key value. */
int blco_host_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                     const uint64_t* cidx, const double* vals,
                     uint64_t max_block_size, blco_host** out);
int blco_host_encode_coo(const blco_coo* t, uint64_t max_block_size,
                         blco_host** out);

/* Shape: modes, lengths, nonzeros, block count, largest block. */
int blco_host_info(const blco_host* h, int* nmodes, uint64_t* dims,
                   uint64_t* nnz, uint64_t* nblocks, uint64_t* max_block_nnz);
/* BLCO masks/positions (blcotensor::mode_masks/mode_pos, blco.hpp:31-35)
 * and the ALTO interleaving masks they were derived from. */
int blco_host_layout(const blco_host* h, uint64_t* masks, int* pos,
                     uint64_t* alto_masks_lo, uint64_t* alto_masks_hi);
/* Block table: start offset, nonzero count and block_coords[nmodes] of each
 * block (blco_block, blco.hpp:11-20). */
int blco_host_blocks(const blco_host* h, uint64_t* block_start,
                     uint64_t* block_nnz, uint64_t* block_coords);
const uint64_t* blco_host_coords(const blco_host* h); /* blcotensor::coords */
const double* blco_host_values(const blco_host* h);   /* blcotensor::values */
/* delete_blcotensor_host, src/blco.cu:62-71 */
void blco_host_free(blco_host* h);

#ifdef __cplusplus
}
#endif
#endif
