/*
 * oracle.h -- CPU restatement of the reference BLCO / ALTO CP-ALS path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (blocked-linearized-coordinate_amd/) never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it restates.  How each function
 * is pinned (golden vectors, the compiled reference under oracle/_ref, closed
 * forms) is written in the header of oracle.c.
 */
#ifndef BLCO_ORACLE_H_
#define BLCO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_MODES 5

/* Software pdep/pext, include/bitops.hpp:17-42 (64-bit) and :84-106 (128-bit
 * as lo/hi halves). */
uint64_t orc_pdep64(uint64_t x, uint64_t mask);
uint64_t orc_pext64(uint64_t x, uint64_t mask);
void orc_pdep128(uint64_t x, uint64_t mask_lo, uint64_t mask_hi,
                 uint64_t* out_lo, uint64_t* out_hi);
uint64_t orc_pext128(uint64_t x_lo, uint64_t x_hi, uint64_t mask_lo,
                     uint64_t mask_hi);

/* setup_packed_alto(LSB_FIRST, SHORT_FIRST), include/alto.hpp:597-695.
 * lit_bits is 64 (cpd64) or 128 (cpd128).  Returns 0, or -1 when the
 * interleaved key needs more than lit_bits bits (alto.hpp:634). */
int orc_alto_setup(int nmodes, const uint64_t* dims, int lit_bits,
                   uint64_t* masks_lo, uint64_t* masks_hi, int* mode_pos,
                   int* mode_bits);

/* gen_alto + sort_alto, include/alto.hpp:107-186, 697-754: key_i = OR_n
 * pdep(cidx[n][i], mask_n), then (key, val) sorted ascending by key.  Ties
 * (duplicate coordinates) keep input order.  cidx is mode-major:
 * cidx[n*nnz + i].  perm[i] is the input index of sorted entry i. */
int orc_alto_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                    const uint64_t* cidx, const double* vals, int lit_bits,
                    uint64_t* keys_lo, uint64_t* keys_hi, double* vals_out,
                    uint64_t* perm);

/* gen_blcotensor_host<LIT>, include/blco.hpp:196-331.  Outputs the
 * relinearised coordinates (in ALTO order), the values in the same order, the
 * BLCO masks/positions and the split block list.  block_cap bounds the block
 * arrays; *nblocks receives the block count (return -2 if it exceeds cap). */
int orc_blco_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                    const uint64_t* cidx, const double* vals,
                    uint64_t max_block_size, int lit_bits, uint64_t* coords,
                    double* vals_out, uint64_t* blco_masks, int* blco_pos,
                    uint64_t* nblocks, uint64_t* block_start,
                    uint64_t* block_nnz, uint64_t* block_coords,
                    uint64_t block_cap);

/* Sequential COO MTTKRP, include/cpd.hpp:510-544.  out += MTTKRP (the caller
 * zeroes or pre-loads it).  U[m] are row-major dims[m] x rank. */
void orc_mttkrp_coo(int nmodes, uint64_t nnz, const uint64_t* cidx,
                    const double* vals, int rank, double* const* U, int mode,
                    double* out);

/* ALTO tensor ready for the CPU MTTKRP: create_alto (gen_alto, re_encode_alto,
 * prtn_alto), include/alto.hpp:188-235, 756-800, 802-897. */
typedef struct orc_alto orc_alto;
orc_alto* orc_alto_create(int nmodes, const uint64_t* dims, uint64_t nnz,
                          const uint64_t* cidx, const double* vals,
                          int nprtn);
void orc_alto_free(orc_alto* at);
uint64_t orc_alto_nnz(const orc_alto* at);

/* mttkrp_alto_par, include/alto.hpp:544-550, dispatching to
 * mttkrp_alto_da_mem_pull (:1216-1307) when nnz/I_n > 4, otherwise
 * mttkrp_alto_atomic (:1019-1075).  out += MTTKRP(mode). */
int orc_alto_mttkrp(orc_alto* at, int mode, int rank, double* const* U,
                    double* out);

/* CP-ALS with the semantics of cpd_alto, include/cpd.hpp:44-187:
 * modes 0..N-1, MTTKRP, Cholesky solve (pseudo-inverse fallback), L2 column
 * norm at iteration 0 and max-norm afterwards, Gram update, fit per
 * cpd_fit_alto (:190-272).  U (in/out, row-major dims[n] x rank) and lambda
 * (out, rank).  fit_trace gets one fit per executed iteration; returns the
 * number of iterations executed.  iter_seconds / mttkrp_seconds (optional)
 * get wall times per iteration. */
int orc_cpd_alto(orc_alto* at, int rank, double** U, double* lambda,
                 int max_iters, double epsilon, double* fit_trace,
                 double* iter_seconds, double* mttkrp_seconds, int verbose);

/* orc_cpd_alto with the GPU library's solve rule: pivot_eps > 0 makes a
 * Cholesky pivot at or below R * pivot_eps * max V_ii fail (as a pivot <= 0
 * does) and sets the pseudo-inverse tolerance to R * ulp in the precision of
 * pivot_eps; pivot_eps = 0 is orc_cpd_alto. */
int orc_cpd_alto_ex(orc_alto* at, int rank, double** U, double* lambda,
                    int max_iters, double epsilon, double* fit_trace,
                    double* iter_seconds, double* mttkrp_seconds, int verbose,
                    double pivot_eps);

/* Normal-equation pieces used by orc_cpd_alto, exported for unit tests. */
/* Lower Cholesky V = L L^T (LAPACK dpotf2 order).  Returns 0 or j+1 for the
 * first non-positive pivot.  V is row-major R x R; L written to the lower
 * triangle of Lout (upper set to 0). */
int orc_cholesky(int R, const double* V, double* Lout);
/* Solve rows in place: each row b of B (nrows x R) -> V^{-1} b using L
 * (forward then backward substitution, as LAPACK dpotrs). */
void orc_cholesky_solve_rows(int R, const double* L, uint64_t nrows,
                             double* B);
/* Moore-Penrose pseudo-inverse of symmetric V by cyclic Jacobi
 * eigendecomposition, tolerance R * ulp(max |eig|) (the reference's GPU rule,
 * src/cpd_gpu.cu:487-489). */
void orc_pinv_sym(int R, const double* V, double* Vpinv);
/* Column norms, src/kruskal_model.cpp:142-259.  which=0: L2, which=1: max. */
void orc_kruskal_norm(uint64_t rows, int rank, double* U, double* lambda,
                      int which);
/* Fit, include/cpd.hpp:190-272 (grams full symmetric, row-major). */
double orc_cpd_fit(int nmodes, uint64_t last_rows, int rank,
                   const double* U_last, const double* mttkrp_last,
                   const double* lambda, double* const* grams, double normX2);

#ifdef __cplusplus
}
#endif
#endif
