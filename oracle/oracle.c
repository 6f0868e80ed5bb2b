/*
 * oracle.c -- CPU restatement of the reference BLCO / ALTO CP-ALS path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Plain C11 + OpenMP, built by oracle/Makefile.
 *
 * What pins each function (checked in tests/test_oracle*.py):
 *   pdep/pext 64/128      brute-force per-bit loop on random words; round trip
 *                         pext(pdep(x,m),m) == x & (2^popcount(m)-1); and the
 *                         compiled reference bitops (oracle/_ref).
 *   orc_alto_setup        hand-worked masks for the 2000x1500x1000,
 *                         12092x9184x28818 and 4.8Mx1.8Mx1.8M shapes
 *                         (tests/golden/alto_masks.json); invariants (disjoint
 *                         masks, union = low sum-of-bits, popcounts); the
 *                         compiled reference setup_packed_alto (oracle/_ref).
 *   orc_alto_encode       compiled reference gen_alto/sort_alto (oracle/_ref)
 *                         key-for-key; strictly increasing keys on unique
 *                         coordinates; pext round trip.
 *   orc_blco_encode       decode(coords, block_coords) == input coordinate for
 *                         every nonzero; concatenated blocks == ALTO order; one
 *                         upper-key value per block; block sizes <= max;
 *                         masks/pos from the hand-worked shapes (65-bit case
 *                         included).  The reference BLCO header cannot be
 *                         compiled here (it includes cublas_v2.h).
 *   orc_mttkrp_coo        dense brute-force einsum (numpy) on tiny tensors;
 *                         all-ones factors give slice sums; rank-1 separable
 *                         tensors.
 *   orc_alto_*            equals orc_mttkrp_coo to 1e-13 relative; compiled
 *                         reference mttkrp_alto_par (oracle/_ref) on the same
 *                         inputs for both the da_mem-pull and atomic variants.
 *   orc_cholesky/solve    L L^T == V; V x == b against numpy.linalg.solve.
 *   orc_pinv_sym          Moore-Penrose identities V V+ V == V, V+ V V+ == V+,
 *                         symmetric V V+; numpy.linalg.pinv on rank-deficient
 *                         matrices.
 *   orc_kruskal_norm      compiled reference KruskalModelNorm (oracle/_ref);
 *                         closed forms (unit L2 columns; max-norm <= 1).
 *   orc_cpd_fit / orc_cpd_alto
 *                         fit == 1 - ||X - [[lambda; U]]||_F / ||X||_F computed
 *                         densely on tiny tensors; an exact low-rank tensor
 *                         converges to fit -> 1; one ALS sweep equals a numpy
 *                         normal-equation restatement.
 *   orc_cpd_alto_ex       pivot_eps = 0 is orc_cpd_alto bit for bit; with a
 *                         floor, a planted near-duplicate column switches the
 *                         solve to the pseudo-inverse and one sweep equals
 *                         the numpy restatement with numpy.linalg.pinv.
 * The reference cpd_alto / pseudo_inverse / SYRK cannot be compiled here (they
 * need cblas.h / lapacke.h), so the ALS step is pinned by the closed forms
 * above, not by the reference binary.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- bitops */

/* include/bitops.hpp:17-29 */
uint64_t orc_pdep64(uint64_t x, uint64_t y)
{
    uint64_t res = 0;
    for (uint64_t b = 1; y != 0; b += b) {
        if (x & b) res |= y & (0 - y);
        y &= (y - 1);
    }
    return res;
}

/* include/bitops.hpp:31-42 */
uint64_t orc_pext64(uint64_t x, uint64_t y)
{
    uint64_t res = 0;
    for (uint64_t b = 1; y != 0; b += b) {
        if (x & y & (0 - y)) res |= b;
        y &= (y - 1);
    }
    return res;
}

static inline uint64_t shr64(uint64_t x, int s) { return s >= 64 ? 0 : (x >> s); }
static inline uint64_t shl64(uint64_t x, int s) { return s >= 64 ? 0 : (x << s); }

/* include/bitops.hpp:84-93: low half deposits the low popcount(mask_lo)
 * bits, the high half the rest.  A shift by 64 (whole low half in one mode)
 * is undefined in the reference; here it yields 0. */
void orc_pdep128(uint64_t x, uint64_t mask_lo, uint64_t mask_hi,
                 uint64_t* out_lo, uint64_t* out_hi)
{
    int shift = __builtin_popcountll(mask_lo);
    *out_hi = orc_pdep64(shr64(x, shift), mask_hi);
    *out_lo = orc_pdep64(x, mask_lo);
}

/* include/bitops.hpp:95-104 */
uint64_t orc_pext128(uint64_t x_lo, uint64_t x_hi, uint64_t mask_lo,
                     uint64_t mask_hi)
{
    int shift = __builtin_popcountll(mask_lo);
    return shl64(orc_pext64(x_hi, mask_hi), shift) | orc_pext64(x_lo, mask_lo);
}

/* ------------------------------------------------------- ALTO linearisation */

/* Bits needed by a mode of length d: 64 - clz(d - 1) (alto.hpp:613).  The
 * reference evaluates clz(0) for d == 1, which is undefined; here 0 bits. */
static int mode_bits_of(uint64_t d)
{
    if (d <= 1) return 0;
    return 64 - __builtin_clzll(d - 1);
}

/* include/alto.hpp:597-695 with po = LSB_FIRST, mo = SHORT_FIRST. */
int orc_alto_setup(int nmodes, const uint64_t* dims, int lit_bits,
                   uint64_t* masks_lo, uint64_t* masks_hi, int* mode_pos,
                   int* mode_bits)
{
    if (nmodes < 1 || nmodes > ORC_MAX_MODES) return -1;
    int bits[ORC_MAX_MODES], order[ORC_MAX_MODES];
    int total = 0, maxb = 0;
    for (int n = 0; n < nmodes; ++n) {
        bits[n] = mode_bits_of(dims[n]);
        total += bits[n];
        if (bits[n] > maxb) maxb = bits[n];
        order[n] = n;
        if (mode_bits) mode_bits[n] = bits[n];
    }
    /* mode_pos: prefix sum in natural mode order (alto.hpp:621-627). */
    mode_pos[0] = 0;
    for (int n = 1; n < nmodes; ++n) mode_pos[n] = mode_pos[n - 1] + bits[n - 1];
    if (total > lit_bits) return -1; /* alto.hpp:634 */
    /* Stable ascending sort by bit count: std::sort on <= 5 elements is an
     * insertion sort, which keeps ties in mode order. */
    for (int i = 1; i < nmodes; ++i) {
        int v = order[i], j = i - 1;
        while (j >= 0 && bits[order[j]] > bits[v]) { order[j + 1] = order[j]; --j; }
        order[j + 1] = v;
    }
    u128 m[ORC_MAX_MODES];
    for (int n = 0; n < nmodes; ++n) m[n] = 0;
    int shift = 0;
    for (int level = 0; level < maxb; ++level) {
        for (int k = 0; k < nmodes; ++k) {
            int n = order[k];
            if (level < bits[n]) {
                m[n] |= ((u128)1) << shift;
                ++shift;
            }
        }
    }
    for (int n = 0; n < nmodes; ++n) {
        masks_lo[n] = (uint64_t)m[n];
        masks_hi[n] = (uint64_t)(m[n] >> 64);
    }
    return 0;
}

/* Stable LSD radix sort of (key, index) pairs over the low `bits` key bits.
 * Used for sort_alto (alto.hpp:697-754); stability fixes the tie order. */
static void radix_sort_keys(uint64_t n, u128* keys, uint64_t* idx, int bits)
{
    if (n < 2 || bits <= 0) return;
    u128* k2 = (u128*)malloc(n * sizeof(u128));
    uint64_t* i2 = (uint64_t*)malloc(n * sizeof(uint64_t));
    int nt = omp_get_max_threads();
    uint64_t* hist = (uint64_t*)malloc((size_t)nt * 256 * sizeof(uint64_t));
    u128* src_k = keys; uint64_t* src_i = idx;
    u128* dst_k = k2;   uint64_t* dst_i = i2;
    int passes = (bits + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        int sh = 8 * p;
        memset(hist, 0, (size_t)nt * 256 * sizeof(uint64_t));
        #pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num();
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t* h = hist + (size_t)t * 256;
            for (uint64_t i = lo; i < hi; ++i) h[(unsigned)(src_k[i] >> sh) & 255u]++;
        }
        uint64_t run = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t) {
                uint64_t c = hist[(size_t)t * 256 + d];
                hist[(size_t)t * 256 + d] = run;
                run += c;
            }
        #pragma omp parallel num_threads(nt)
        {
            int t = omp_get_thread_num();
            uint64_t lo = n * (uint64_t)t / nt, hi = n * (uint64_t)(t + 1) / nt;
            uint64_t* h = hist + (size_t)t * 256;
            for (uint64_t i = lo; i < hi; ++i) {
                uint64_t o = h[(unsigned)(src_k[i] >> sh) & 255u]++;
                dst_k[o] = src_k[i];
                dst_i[o] = src_i[i];
            }
        }
        u128* tk = src_k; src_k = dst_k; dst_k = tk;
        uint64_t* ti = src_i; src_i = dst_i; dst_i = ti;
    }
    if (src_k != keys) {
        memcpy(keys, src_k, n * sizeof(u128));
        memcpy(idx, src_i, n * sizeof(uint64_t));
    }
    free(hist);
    free(k2);
    free(i2);
}

/* Linearise and sort; keys are full 128-bit values (upper half 0 for
 * lit_bits == 64).  Returns total key bits or -1. */
static int alto_keys_sorted(int nmodes, const uint64_t* dims, uint64_t nnz,
                            const uint64_t* cidx, int lit_bits, u128* keys,
                            uint64_t* perm, uint64_t* mlo, uint64_t* mhi,
                            int* pos, int* bits)
{
    if (orc_alto_setup(nmodes, dims, lit_bits, mlo, mhi, pos, bits) != 0) return -1;
    int total = 0;
    for (int n = 0; n < nmodes; ++n) total += bits[n];
    /* gen_alto linearisation loop, alto.hpp:157-167 */
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        u128 k = 0;
        for (int n = 0; n < nmodes; ++n) {
            uint64_t lo, hi;
            orc_pdep128(cidx[(uint64_t)n * nnz + i], mlo[n], mhi[n], &lo, &hi);
            k |= ((u128)hi << 64) | lo;
        }
        keys[i] = k;
        perm[i] = i;
    }
    radix_sort_keys(nnz, keys, perm, total);
    return total;
}

int orc_alto_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                    const uint64_t* cidx, const double* vals, int lit_bits,
                    uint64_t* keys_lo, uint64_t* keys_hi, double* vals_out,
                    uint64_t* perm)
{
    uint64_t mlo[ORC_MAX_MODES], mhi[ORC_MAX_MODES];
    int pos[ORC_MAX_MODES], bits[ORC_MAX_MODES];
    u128* keys = (u128*)malloc((nnz ? nnz : 1) * sizeof(u128));
    if (!keys) return -3;
    int total = alto_keys_sorted(nmodes, dims, nnz, cidx, lit_bits, keys, perm,
                                 mlo, mhi, pos, bits);
    if (total < 0) { free(keys); return -1; }
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        keys_lo[i] = (uint64_t)keys[i];
        keys_hi[i] = (uint64_t)(keys[i] >> 64);
        vals_out[i] = vals[perm[i]];
    }
    free(keys);
    return 0;
}

/* ------------------------------------------------------------ BLCO encode */

/* include/blco.hpp:196-331. */
int orc_blco_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                    const uint64_t* cidx, const double* vals,
                    uint64_t max_block_size, int lit_bits, uint64_t* coords,
                    double* vals_out, uint64_t* blco_masks, int* blco_pos,
                    uint64_t* nblocks, uint64_t* block_start,
                    uint64_t* block_nnz, uint64_t* block_coords,
                    uint64_t block_cap)
{
    uint64_t mlo[ORC_MAX_MODES], mhi[ORC_MAX_MODES];
    int pos[ORC_MAX_MODES], bits[ORC_MAX_MODES];
    u128* keys = (u128*)malloc((nnz ? nnz : 1) * sizeof(u128));
    uint64_t* perm = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    if (!keys || !perm) { free(keys); free(perm); return -3; }
    int total = alto_keys_sorted(nmodes, dims, nnz, cidx, lit_bits, keys, perm,
                                 mlo, mhi, pos, bits);
    if (total < 0) { free(keys); free(perm); return -1; }

    /* truncated bit counts, masks and positions, blco.hpp:222-233 */
    int trunc[ORC_MAX_MODES];
    for (int n = 0; n < nmodes; ++n) {
        trunc[n] = __builtin_popcountll(mlo[n]);
        blco_masks[n] = trunc[n] >= 64 ? ~0ull : ((1ull << trunc[n]) - 1);
        blco_pos[n] = (n == 0) ? 0 : blco_pos[n - 1] + trunc[n - 1];
    }
    /* relinearise, blco.hpp:274-287: for each mode the low-half bits of the
     * mode index, packed at blco_pos[n]. */
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t klo = (uint64_t)keys[i];
        uint64_t c = 0;
        for (int n = 0; n < nmodes; ++n)
            c |= shl64(orc_pext64(klo, mlo[n]) & blco_masks[n], blco_pos[n]);
        coords[i] = c;
        vals_out[i] = vals[perm[i]];
    }
    /* blocks: runs of equal upper key half (the reference's histogram over
     * uhalf(key), blco.hpp:251-272, visited in ascending block id), each split
     * into pieces of at most max_block_size (blco.hpp:296-321). */
    uint64_t mbs = max_block_size ? max_block_size : nnz; /* blco.hpp:264 */
    uint64_t nb = 0;
    uint64_t i = 0;
    while (i < nnz) {
        uint64_t bid = (uint64_t)(keys[i] >> 64);
        uint64_t j = i + 1;
        while (j < nnz && (uint64_t)(keys[j] >> 64) == bid) ++j;
        for (uint64_t s = i; s < j; s += mbs) {
            uint64_t cnt = (j - s) < mbs ? (j - s) : mbs;
            if (nb < block_cap) {
                block_start[nb] = s;
                block_nnz[nb] = cnt;
                for (int n = 0; n < nmodes; ++n)
                    block_coords[nb * (uint64_t)nmodes + n] =
                        shl64(orc_pext64(bid, mhi[n]), trunc[n]);
            }
            ++nb;
        }
        i = j;
    }
    *nblocks = nb;
    free(keys);
    free(perm);
    return nb > block_cap ? -2 : 0;
}

/* --------------------------------------------------------- COO MTTKRP */

/* include/cpd.hpp:510-544 */
void orc_mttkrp_coo(int nmodes, uint64_t nnz, const uint64_t* cidx,
                    const double* vals, int rank, double* const* U, int mode,
                    double* out)
{
    double* row = (double*)malloc(sizeof(double) * (size_t)rank);
    for (uint64_t i = 0; i < nnz; ++i) {
        for (int r = 0; r < rank; ++r) row[r] = vals[i];
        for (int m = 0; m < nmodes; ++m) {
            if (m == mode) continue;
            uint64_t rid = cidx[(uint64_t)m * nnz + i];
            const double* f = U[m] + rid * (uint64_t)rank;
            for (int r = 0; r < rank; ++r) row[r] *= f[r];
        }
        uint64_t rid = cidx[(uint64_t)mode * nnz + i];
        double* o = out + rid * (uint64_t)rank;
        for (int r = 0; r < rank; ++r) o[r] += row[r];
    }
    free(row);
}

/* ------------------------------------------------------ ALTO CPU MTTKRP */

typedef struct { uint64_t start, stop; } orc_interval;

struct orc_alto {
    int nmode;
    int nprtn;
    uint64_t nnz;
    uint64_t dims[ORC_MAX_MODES];
    int wide;                      /* 1: keys need 128 bits */
    uint64_t* idx64;               /* re-encoded keys, 64-bit form */
    u128* idx128;                  /* re-encoded keys, 128-bit form */
    double* vals;
    uint64_t mode_masks[ORC_MAX_MODES]; /* (1 << bits) - 1 after re-encode */
    int mode_pos[ORC_MAX_MODES];
    uint64_t* prtn_ptr;            /* nprtn + 1 */
    orc_interval* intervals;       /* nprtn * nmode */
    double** ofibs;                /* da_mem buffers, one per partition */
    uint64_t* ofib_rows;           /* capacity (rows x rank cells) of each buffer */
};

uint64_t orc_alto_nnz(const orc_alto* at) { return at->nnz; }

static inline uint64_t alto_field(const orc_alto* at, uint64_t i, int n)
{
    /* pext(x, mask, pos) = (x >> pos) & mask, bitops.hpp:44-48 */
    if (at->wide) return (uint64_t)(at->idx128[i] >> at->mode_pos[n]) & at->mode_masks[n];
    return (at->idx64[i] >> at->mode_pos[n]) & at->mode_masks[n];
}

/* create_alto, include/alto.hpp:188-235: gen_alto, re_encode_alto
 * (:756-800) and prtn_alto (:802-897). */
orc_alto* orc_alto_create(int nmodes, const uint64_t* dims, uint64_t nnz,
                          const uint64_t* cidx, const double* vals, int nprtn)
{
    if (nprtn < 1) nprtn = 1;
    orc_alto* at = (orc_alto*)calloc(1, sizeof(orc_alto));
    uint64_t mlo[ORC_MAX_MODES], mhi[ORC_MAX_MODES];
    int bits[ORC_MAX_MODES];
    u128* keys = (u128*)malloc((nnz ? nnz : 1) * sizeof(u128));
    uint64_t* perm = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    int total = alto_keys_sorted(nmodes, dims, nnz, cidx, 128, keys, perm, mlo,
                                 mhi, at->mode_pos, bits);
    if (total < 0) { free(keys); free(perm); free(at); return NULL; }
    at->nmode = nmodes;
    at->nprtn = nprtn;
    at->nnz = nnz;
    at->wide = total > 64;
    for (int n = 0; n < nmodes; ++n) {
        at->dims[n] = dims[n];
        at->mode_masks[n] = bits[n] >= 64 ? ~0ull : ((1ull << bits[n]) - 1);
    }
    at->vals = (double*)malloc((nnz ? nnz : 1) * sizeof(double));
    if (at->wide) at->idx128 = (u128*)malloc((nnz ? nnz : 1) * sizeof(u128));
    else at->idx64 = (uint64_t*)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
    /* re_encode_alto: new_index = OR_n pext(key, mask_n) << mode_pos[n] */
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t klo = (uint64_t)keys[i], khi = (uint64_t)(keys[i] >> 64);
        u128 ni = 0;
        for (int n = 0; n < nmodes; ++n)
            ni |= ((u128)orc_pext128(klo, khi, mlo[n], mhi[n])) << at->mode_pos[n];
        if (at->wide) at->idx128[i] = ni;
        else at->idx64[i] = (uint64_t)ni;
        at->vals[i] = vals[perm[i]];
    }
    free(keys);
    free(perm);
    /* prtn_alto: equal nnz ranges and per-mode index intervals */
    at->prtn_ptr = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nprtn + 1));
    at->intervals = (orc_interval*)malloc(sizeof(orc_interval) * (size_t)nprtn * nmodes);
    uint64_t per = (nnz + nprtn - 1) / nprtn;
    at->prtn_ptr[0] = 0;
    for (int p = 0; p < nprtn; ++p) {
        uint64_t s = (uint64_t)p * per, e = (uint64_t)(p + 1) * per;
        if (e > nnz) e = nnz;
        if (s > e) s = e;
        (void)s;
        at->prtn_ptr[p + 1] = e;
    }
    #pragma omp parallel for schedule(static, 1)
    for (int p = 0; p < nprtn; ++p) {
        uint64_t lo[ORC_MAX_MODES], hi[ORC_MAX_MODES];
        for (int n = 0; n < nmodes; ++n) { lo[n] = dims[n]; hi[n] = 0; }
        for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i)
            for (int n = 0; n < nmodes; ++n) {
                uint64_t v = alto_field(at, i, n);
                if (v < lo[n]) lo[n] = v;
                if (v > hi[n]) hi[n] = v;
            }
        for (int n = 0; n < nmodes; ++n) {
            at->intervals[(size_t)p * nmodes + n].start = lo[n];
            at->intervals[(size_t)p * nmodes + n].stop = hi[n];
        }
    }
    at->ofibs = (double**)calloc((size_t)nprtn, sizeof(double*));
    at->ofib_rows = (uint64_t*)calloc((size_t)nprtn, sizeof(uint64_t));
    return at;
}

void orc_alto_free(orc_alto* at)
{
    if (!at) return;
    for (int p = 0; p < at->nprtn; ++p) free(at->ofibs[p]);
    free(at->ofibs);
    free(at->ofib_rows);
    free(at->idx64);
    free(at->idx128);
    free(at->vals);
    free(at->prtn_ptr);
    free(at->intervals);
    free(at);
}

#define MIN_FIBER_REUSE 4 /* alto.hpp:89 */

/* mttkrp_alto_atomic, alto.hpp:1019-1075 */
static void alto_mttkrp_atomic(orc_alto* at, int mode, int rank,
                               double* const* U, double* out)
{
    int nm = at->nmode;
    #pragma omp parallel for schedule(static, 1)
    for (int p = 0; p < at->nprtn; ++p) {
        double row[512];
        double* rowp = rank <= 512 ? row : (double*)malloc(sizeof(double) * rank);
        for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
            double v = at->vals[i];
            for (int r = 0; r < rank; ++r) rowp[r] = v;
            for (int m = 0; m < nm; ++m) {
                if (m == mode) continue;
                const double* f = U[m] + alto_field(at, i, m) * (uint64_t)rank;
                for (int r = 0; r < rank; ++r) rowp[r] *= f[r];
            }
            double* o = out + alto_field(at, i, mode) * (uint64_t)rank;
            for (int r = 0; r < rank; ++r) {
                #pragma omp atomic update
                o[r] += rowp[r];
            }
        }
        if (rowp != row) free(rowp);
    }
}

/* mttkrp_alto_da_mem_pull, alto.hpp:1216-1307: every partition accumulates
 * into a private dense buffer covering its index interval, then rows are
 * pulled from all overlapping partitions in partition order. */
static void alto_mttkrp_pull(orc_alto* at, int mode, int rank,
                             double* const* U, double* out)
{
    int nm = at->nmode, np = at->nprtn;
    for (int p = 0; p < np; ++p) {
        orc_interval iv = at->intervals[(size_t)p * nm + mode];
        uint64_t rows = (iv.stop >= iv.start) ? iv.stop - iv.start + 1 : 0;
        uint64_t cells = rows * (uint64_t)rank;  /* capacity is kept in cells */
        if (cells > at->ofib_rows[p]) {
            free(at->ofibs[p]);
            at->ofibs[p] = (double*)malloc(sizeof(double) * cells);
            at->ofib_rows[p] = cells;
        }
    }
    uint64_t I = at->dims[mode];
    #pragma omp parallel
    {
        double row[512];
        double* rowp = rank <= 512 ? row : (double*)malloc(sizeof(double) * rank);
        #pragma omp for schedule(static, 1)
        for (int p = 0; p < np; ++p) {
            orc_interval iv = at->intervals[(size_t)p * nm + mode];
            if (iv.stop < iv.start) continue;
            double* buf = at->ofibs[p];
            memset(buf, 0, sizeof(double) * (iv.stop - iv.start + 1) * (uint64_t)rank);
            for (uint64_t i = at->prtn_ptr[p]; i < at->prtn_ptr[p + 1]; ++i) {
                double v = at->vals[i];
                for (int r = 0; r < rank; ++r) rowp[r] = v;
                for (int m = 0; m < nm; ++m) {
                    if (m == mode) continue;
                    const double* f = U[m] + alto_field(at, i, m) * (uint64_t)rank;
                    for (int r = 0; r < rank; ++r) rowp[r] *= f[r];
                }
                double* o = buf + (alto_field(at, i, mode) - iv.start) * (uint64_t)rank;
                for (int r = 0; r < rank; ++r) o[r] += rowp[r];
            }
        }
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < I; ++i) {
            for (int p = 0; p < np; ++p) {
                orc_interval iv = at->intervals[(size_t)p * nm + mode];
                if (iv.stop < iv.start || i < iv.start || i > iv.stop) continue;
                const double* b = at->ofibs[p] + (i - iv.start) * (uint64_t)rank;
                double* o = out + i * (uint64_t)rank;
                for (int r = 0; r < rank; ++r) o[r] += b[r];
            }
        }
        if (rowp != row) free(rowp);
    }
}

/* mttkrp_alto_par, alto.hpp:544-550 (rank_specializer dispatch :487-496) */
int orc_alto_mttkrp(orc_alto* at, int mode, int rank, double* const* U,
                    double* out)
{
    if (!at || mode < 0 || mode >= at->nmode || rank < 1) return -1;
    uint64_t reuse = at->nnz / at->dims[mode];
    if (reuse <= MIN_FIBER_REUSE) alto_mttkrp_atomic(at, mode, rank, U, out);
    else alto_mttkrp_pull(at, mode, rank, U, out);
    return 0;
}

/* ------------------------------------------------- normal equations */

/* Unblocked lower Cholesky in LAPACK dpotf2 order (the reference calls
 * dpotrf_ with uplo 'L', cpd.hpp:594).  A pivot at or below `floor` fails
 * (dpotrf: floor 0). */
static int cholesky_floor(int R, const double* V, double* L, double floor)
{
    for (int i = 0; i < R * R; ++i) L[i] = 0.0;
    for (int j = 0; j < R; ++j) {
        double ajj = V[j * R + j];
        for (int k = 0; k < j; ++k) ajj -= L[j * R + k] * L[j * R + k];
        if (!(ajj > floor)) return j + 1;
        ajj = sqrt(ajj);
        L[j * R + j] = ajj;
        for (int i = j + 1; i < R; ++i) {
            double s = V[i * R + j];
            for (int k = 0; k < j; ++k) s -= L[i * R + k] * L[j * R + k];
            L[i * R + j] = s / ajj;
        }
    }
    return 0;
}

int orc_cholesky(int R, const double* V, double* L) { return cholesky_floor(R, V, L, 0.0); }

/* dpotrs_ with one right-hand side per row of B, cpd.hpp:597-599. */
void orc_cholesky_solve_rows(int R, const double* L, uint64_t nrows, double* B)
{
    #pragma omp parallel for schedule(static)
    for (uint64_t row = 0; row < nrows; ++row) {
        double* b = B + row * (uint64_t)R;
        for (int i = 0; i < R; ++i) { /* L y = b */
            double s = b[i];
            for (int k = 0; k < i; ++k) s -= L[i * R + k] * b[k];
            b[i] = s / L[i * R + i];
        }
        for (int i = R - 1; i >= 0; --i) { /* L^T x = y */
            double s = b[i];
            for (int k = i + 1; k < R; ++k) s -= L[k * R + i] * b[k];
            b[i] = s / L[i * R + i];
        }
    }
}

/* Cyclic Jacobi eigendecomposition of symmetric A (R x R, row-major);
 * A is destroyed (diagonal -> eigenvalues), Q receives eigenvectors as
 * columns.  Rotation formula of Golub & Van Loan, Algorithm 8.4.2. */
static void jacobi_eig(int R, double* A, double* Q)
{
    for (int i = 0; i < R * R; ++i) Q[i] = 0.0;
    for (int i = 0; i < R; ++i) Q[i * R + i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < R; ++i)
            for (int j = 0; j < R; ++j) {
                if (i != j) off += A[i * R + j] * A[i * R + j];
                else diag += A[i * R + j] * A[i * R + j];
            }
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < R - 1; ++p)
            for (int q = p + 1; q < R; ++q) {
                double apq = A[p * R + q];
                if (apq == 0.0) continue;
                double app = A[p * R + p], aqq = A[q * R + q];
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < R; ++k) { /* A <- J^T A J: columns */
                    double akp = A[k * R + p], akq = A[k * R + q];
                    A[k * R + p] = c * akp - s * akq;
                    A[k * R + q] = s * akp + c * akq;
                }
                for (int k = 0; k < R; ++k) { /* rows */
                    double apk = A[p * R + k], aqk = A[q * R + k];
                    A[p * R + k] = c * apk - s * aqk;
                    A[q * R + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < R; ++k) {
                    double qkp = Q[k * R + p], qkq = Q[k * R + q];
                    Q[k * R + p] = c * qkp - s * qkq;
                    Q[k * R + q] = s * qkp + c * qkq;
                }
            }
    }
}

/* tol = R * ulp(emax) in the precision of machine epsilon eps: for
 * eps = DBL_EPSILON, R * (nextafter(emax, emax+1) - emax) of
 * src/cpd_gpu.cu:487-489. */
static void pinv_sym_eps(int R, const double* V, double* P, double eps)
{
    double* A = (double*)malloc(sizeof(double) * R * R);
    double* Q = (double*)malloc(sizeof(double) * R * R);
    double* w = (double*)malloc(sizeof(double) * R);
    memcpy(A, V, sizeof(double) * R * R);
    jacobi_eig(R, A, Q);
    double emax = 0.0;
    for (int i = 0; i < R; ++i) {
        w[i] = A[i * R + i];
        if (fabs(w[i]) > emax) emax = fabs(w[i]);
    }
    double tol = (eps == DBL_EPSILON) ? R * (nextafter(emax, emax + 1.0) - emax)
                 : (emax > 0.0 ? R * ldexp(eps, ilogb(emax)) : 0.0);
    for (int i = 0; i < R; ++i) w[i] = (fabs(w[i]) > tol) ? 1.0 / w[i] : 0.0;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < R; ++j) {
            double s = 0.0;
            for (int k = 0; k < R; ++k) s += Q[i * R + k] * w[k] * Q[j * R + k];
            P[i * R + j] = s;
        }
    free(A);
    free(Q);
    free(w);
}

void orc_pinv_sym(int R, const double* V, double* P) { pinv_sym_eps(R, V, P, DBL_EPSILON); }

/* KruskalModelNorm, src/kruskal_model.cpp:142-259.  A zero L2 column is
 * left at zero with lambda 0 (the reference divides by zero on the CPU and
 * guards on the GPU, src/cpd_gpu.cu:276-279; the guard is kept). */
void orc_kruskal_norm(uint64_t rows, int rank, double* U, double* lambda,
                      int which)
{
    for (int r = 0; r < rank; ++r) lambda[r] = (which == 0) ? 0.0 : 1.0;
    if (which == 0) {
        for (uint64_t i = 0; i < rows; ++i)
            for (int r = 0; r < rank; ++r) lambda[r] += U[i * rank + r] * U[i * rank + r];
        for (int r = 0; r < rank; ++r) lambda[r] = sqrt(lambda[r]);
    } else {
        /* lambda_r = max(1, max_i U_ir) (kruskal_model.cpp:183-223) */
        for (uint64_t i = 0; i < rows; ++i)
            for (int r = 0; r < rank; ++r)
                if (U[i * rank + r] > lambda[r]) lambda[r] = U[i * rank + r];
    }
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < rows; ++i)
        for (int r = 0; r < rank; ++r)
            if (lambda[r] != 0.0) U[i * rank + r] /= lambda[r];
}

static void gram_full(uint64_t rows, int R, const double* U, double* G)
{
    /* SYRK of the reference (cpd.hpp:676-689), both triangles. */
    for (int i = 0; i < R * R; ++i) G[i] = 0.0;
    int nt = omp_get_max_threads();
    double* part = (double*)calloc((size_t)nt * R * R, sizeof(double));
    #pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        double* g = part + (size_t)t * R * R;
        #pragma omp for schedule(static)
        for (uint64_t i = 0; i < rows; ++i) {
            const double* u = U + i * (uint64_t)R;
            for (int a = 0; a < R; ++a)
                for (int b = 0; b <= a; ++b) g[a * R + b] += u[a] * u[b];
        }
    }
    for (int t = 0; t < nt; ++t)
        for (int a = 0; a < R; ++a)
            for (int b = 0; b <= a; ++b) G[a * R + b] += part[(size_t)t * R * R + a * R + b];
    for (int a = 0; a < R; ++a)
        for (int b = a + 1; b < R; ++b) G[a * R + b] = G[b * R + a];
    free(part);
}

/* cpd_fit_alto, include/cpd.hpp:190-272 */
double orc_cpd_fit(int nmodes, uint64_t last_rows, int rank,
                   const double* U_last, const double* mttkrp_last,
                   const double* lambda, double* const* grams, double normX2)
{
    double* acc = (double*)calloc((size_t)rank, sizeof(double));
    for (uint64_t i = 0; i < last_rows; ++i)
        for (int r = 0; r < rank; ++r) acc[r] += U_last[i * rank + r] * mttkrp_last[i * rank + r];
    double inner = 0.0;
    for (int r = 0; r < rank; ++r) inner += acc[r] * lambda[r];
    double normU = 0.0, diag = 0.0;
    for (int i = 0; i < rank; ++i)
        for (int j = 0; j <= i; ++j) {
            double t = lambda[i] * lambda[j];
            for (int n = 0; n < nmodes; ++n) t *= grams[n][i * rank + j];
            if (j < i) normU += 2.0 * t;
            else diag += t;
        }
    normU = fabs(normU + diag);
    double res = normX2 + normU - 2.0 * inner;
    if (res > 0.0) res = sqrt(res);
    free(acc);
    return 1.0 - res / sqrt(normX2);
}

/* cpd_alto, include/cpd.hpp:44-187.  pivot_eps = 0: the reference's solve
 * (dpotrf fails on a pivot <= 0, then the pseudo-inverse with tolerance
 * R * ulp_f64); pivot_eps > 0: a pivot at or below R * pivot_eps * max V_ii
 * also fails and the pseudo-inverse tolerance is R * ulp in the precision of
 * pivot_eps (the rule of the GPU library's factorisation). */
int orc_cpd_alto_ex(orc_alto* at, int rank, double** U, double* lambda,
                    int max_iters, double epsilon, double* fit_trace,
                    double* iter_seconds, double* mttkrp_seconds, int verbose,
                    double pivot_eps)
{
    int N = at->nmode, R = rank;
    double normX2 = 0.0;
    #pragma omp parallel for reduction(+ : normX2) schedule(static)
    for (uint64_t i = 0; i < at->nnz; ++i) normX2 += at->vals[i] * at->vals[i];
    double* grams[ORC_MAX_MODES];
    for (int n = 0; n < N; ++n) {
        grams[n] = (double*)malloc(sizeof(double) * R * R);
        gram_full(at->dims[n], R, U[n], grams[n]);
    }
    double* V = (double*)malloc(sizeof(double) * R * R);
    double* L = (double*)malloc(sizeof(double) * R * R);
    double* P = (double*)malloc(sizeof(double) * R * R);
    uint64_t last = at->dims[N - 1];
    double* scratch = (double*)malloc(sizeof(double) * last * R);
    double* tmp = NULL;
    uint64_t tmp_rows = 0;
    double prev = 0.0, fit = 0.0;
    int done = 0;
    for (int it = 0; it < max_iters; ++it) {
        double t0 = omp_get_wtime(), tm = 0.0;
        for (int j = 0; j < N; ++j) {
            uint64_t I = at->dims[j];
            memset(U[j], 0, sizeof(double) * I * R);
            double t1 = omp_get_wtime();
            orc_alto_mttkrp(at, j, R, U, U[j]);
            tm += omp_get_wtime() - t1;
            if (j == N - 1) memcpy(scratch, U[j], sizeof(double) * I * R);
            /* pseudo_inverse, cpd.hpp:547-648: V = Hadamard of other grams */
            for (int k = 0; k < R * R; ++k) V[k] = 1.0;
            for (int m = 0; m < N; ++m)
                if (m != j)
                    for (int k = 0; k < R * R; ++k) V[k] *= grams[m][k];
            double floor = 0.0;
            if (pivot_eps > 0.0) {
                double dmax = 0.0;
                for (int k = 0; k < R; ++k) dmax = fmax(dmax, V[k * R + k]);
                floor = R * pivot_eps * dmax;
            }
            if (cholesky_floor(R, V, L, floor) == 0) {
                orc_cholesky_solve_rows(R, L, I, U[j]);
            } else {
                /* rank-deficient fallback: minimum-norm solution V^+ b */
                pinv_sym_eps(R, V, P, pivot_eps > 0.0 ? pivot_eps : DBL_EPSILON);
                if (tmp_rows < I) { free(tmp); tmp = (double*)malloc(sizeof(double) * I * R); tmp_rows = I; }
                memcpy(tmp, U[j], sizeof(double) * I * R);
                #pragma omp parallel for schedule(static)
                for (uint64_t i = 0; i < I; ++i)
                    for (int a = 0; a < R; ++a) {
                        double s = 0.0;
                        for (int b = 0; b < R; ++b) s += P[a * R + b] * tmp[i * R + b];
                        U[j][i * R + a] = s;
                    }
                if (verbose) fprintf(stderr, "\t Mode %d Min Norm Solve\n", j);
            }
            orc_kruskal_norm(I, R, U[j], lambda, it == 0 ? 0 : 1);
            gram_full(I, R, U[j], grams[j]);
        }
        fit = orc_cpd_fit(N, last, R, U[N - 1], scratch, lambda, grams, normX2);
        double dt = omp_get_wtime() - t0;
        if (fit_trace) fit_trace[it] = fit;
        if (iter_seconds) iter_seconds[it] = dt;
        if (mttkrp_seconds) mttkrp_seconds[it] = tm;
        if (verbose)
            printf("it: %d\t fit: %g\t fit-delta: %g\ttime(for MTTKRP): %.4f s (%.4f s)\n",
                   it, fit, fabs(prev - fit), dt, tm);
        done = it + 1;
        if (it > 0 && fabs(prev - fit) < epsilon) break;
        prev = fit;
    }
    for (int n = 0; n < N; ++n) free(grams[n]);
    free(V); free(L); free(P); free(scratch); free(tmp);
    return done;
}

int orc_cpd_alto(orc_alto* at, int rank, double** U, double* lambda,
                 int max_iters, double epsilon, double* fit_trace,
                 double* iter_seconds, double* mttkrp_seconds, int verbose)
{
    return orc_cpd_alto_ex(at, rank, U, lambda, max_iters, epsilon, fit_trace,
                           iter_seconds, mttkrp_seconds, verbose, 0.0);
}
