// ref_harness.cpp -- extern "C" entry points around the reference's own CPU
// ALTO sources, so the oracle's restatement can be checked against them.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile.ref against the
// headers and sources where they lie under /root/reference (include/alto.hpp,
// include/bitops.hpp, include/common.hpp, include/sptensor.hpp,
// include/kruskal_model.hpp, src/kruskal_model.cpp) into oracle/_ref/.  No
// reference text is copied here; this file only builds their structs from
// plain arrays and calls their functions.
#include <stdint.h>
#include <string.h>
#include "alto.hpp"
#include "kruskal_model.hpp"

template <typename LIT>
static AltoTensor<LIT>* shell(int nmodes, const uint64_t* dims)
{
    AltoTensor<LIT>* at = new AltoTensor<LIT>();
    at->nmode = nmodes;
    at->nnz = 0;
    at->dims = (IType*)malloc(sizeof(IType) * nmodes);
    for (int n = 0; n < nmodes; ++n) at->dims[n] = dims[n];
    at->mode_masks = (LIT*)calloc(nmodes, sizeof(LIT));
    at->mode_pos = (int*)calloc(nmodes, sizeof(int));
    return at;
}

static SparseTensor* make_spt(int nmodes, const uint64_t* dims, uint64_t nnz,
                              const uint64_t* cidx, const double* vals)
{
    SparseTensor* X = (SparseTensor*)malloc(sizeof(SparseTensor));
    X->nmodes = nmodes;
    X->nnz = nnz;
    X->dims = (IType*)malloc(sizeof(IType) * nmodes);
    X->cidx = (IType**)malloc(sizeof(IType*) * nmodes);
    for (int n = 0; n < nmodes; ++n) {
        X->dims[n] = dims[n];
        X->cidx[n] = (IType*)malloc(sizeof(IType) * (nnz ? nnz : 1));
        memcpy(X->cidx[n], cidx + (uint64_t)n * nnz, sizeof(IType) * nnz);
    }
    X->vals = (FType*)malloc(sizeof(FType) * (nnz ? nnz : 1));
    memcpy(X->vals, vals, sizeof(FType) * nnz);
    return X;
}

static void free_spt(SparseTensor* X)
{
    for (int n = 0; n < X->nmodes; ++n) free(X->cidx[n]);
    free(X->cidx); free(X->dims); free(X->vals); free(X);
}

// Mean seconds of one mttkrp_alto_par call of `mode` (omp_get_wtime around
// each of `reps` calls after one untimed call; ALTO construction and the
// partial matrices excluded, as src/main.cpp:538-552 times it).  Keys of more
// than 64 bits take the reference's 128-bit ALTO index.
template <typename LIT>
static double time_mttkrp(SparseTensor* X, int nprtn, int rank, double* const* U, int mode, double* out, int reps)
{
    AltoTensor<LIT>* at;
    create_alto(X, &at, nprtn);
    FType** ofibs = NULL;
    create_da_mem(mode, (IType)rank, at, &ofibs);
    FType* factors[MAX_NUM_MODES];
    for (int n = 0; n < X->nmodes; ++n) factors[n] = U[n];
    factors[mode] = out;
    mttkrp_alto_par(mode, factors, (IType)rank, at, NULL, ofibs);
    double t = 0.0;
    for (int r = 0; r < reps; ++r) {
        const double t0 = omp_get_wtime();
        mttkrp_alto_par(mode, factors, (IType)rank, at, NULL, ofibs);
        t += omp_get_wtime() - t0;
    }
    destroy_da_mem(at, ofibs, (IType)rank, mode);
    destroy_alto(at);
    return reps > 0 ? t / reps : 0.0;
}

extern "C" {

int ref_alto_setup(int nmodes, const uint64_t* dims, int lit_bits,
                   uint64_t* masks_lo, uint64_t* masks_hi, int* pos)
{
    if (lit_bits == 64) {
        AltoTensor<unsigned long long>* at = shell<unsigned long long>(nmodes, dims);
        setup_packed_alto(at, LSB_FIRST, SHORT_FIRST);
        for (int n = 0; n < nmodes; ++n) { masks_lo[n] = at->mode_masks[n]; masks_hi[n] = 0; pos[n] = at->mode_pos[n]; }
        return 0;
    }
    AltoTensor<unsigned __int128>* at = shell<unsigned __int128>(nmodes, dims);
    setup_packed_alto(at, LSB_FIRST, SHORT_FIRST);
    for (int n = 0; n < nmodes; ++n) {
        masks_lo[n] = (uint64_t)at->mode_masks[n];
        masks_hi[n] = (uint64_t)(at->mode_masks[n] >> 64);
        pos[n] = at->mode_pos[n];
    }
    return 0;
}

int ref_gen_alto(int nmodes, const uint64_t* dims, uint64_t nnz,
                 const uint64_t* cidx, const double* vals, int lit_bits,
                 uint64_t* keys_lo, uint64_t* keys_hi, double* vals_out)
{
    SparseTensor* X = make_spt(nmodes, dims, nnz, cidx, vals);
    if (lit_bits == 64) {
        AltoTensor<unsigned long long>* at;
        gen_alto(X, &at);
        for (uint64_t i = 0; i < nnz; ++i) { keys_lo[i] = at->idx[i]; keys_hi[i] = 0; vals_out[i] = at->vals[i]; }
    } else {
        AltoTensor<unsigned __int128>* at;
        gen_alto(X, &at);
        for (uint64_t i = 0; i < nnz; ++i) {
            keys_lo[i] = (uint64_t)at->idx[i];
            keys_hi[i] = (uint64_t)(at->idx[i] >> 64);
            vals_out[i] = at->vals[i];
        }
    }
    free_spt(X);
    return 0;
}

// out must hold dims[mode]*rank values (the caller zeroes it); U[m] for m !=
// mode are the input factors.
int ref_mttkrp_alto_par(int nmodes, const uint64_t* dims, uint64_t nnz,
                        const uint64_t* cidx, const double* vals, int nprtn,
                        int rank, double* const* U, int mode, double* out)
{
    SparseTensor* X = make_spt(nmodes, dims, nnz, cidx, vals);
    AltoTensor<unsigned long long>* at;
    create_alto(X, &at, nprtn);
    FType** ofibs = NULL;
    create_da_mem(mode, (IType)rank, at, &ofibs);
    FType* factors[MAX_NUM_MODES];
    for (int n = 0; n < nmodes; ++n) factors[n] = U[n];
    factors[mode] = out;
    mttkrp_alto_par(mode, factors, (IType)rank, at, NULL, ofibs);
    destroy_da_mem(at, ofibs, (IType)rank, mode);
    destroy_alto(at);
    free_spt(X);
    return 0;
}

double ref_mttkrp_alto_par_time(int nmodes, const uint64_t* dims, uint64_t nnz, const uint64_t* cidx,
                                const double* vals, int nprtn, int rank, double* const* U, int mode, double* out,
                                int reps)
{
    SparseTensor* X = make_spt(nmodes, dims, nnz, cidx, vals);
    int bits = 0;
    for (int n = 0; n < nmodes; ++n)
        for (uint64_t d = dims[n] > 1 ? dims[n] - 1 : 0; d; d >>= 1) ++bits;
    const double t = bits > 64 ? time_mttkrp<unsigned __int128>(X, nprtn, rank, U, mode, out, reps)
                               : time_mttkrp<unsigned long long>(X, nprtn, rank, U, mode, out, reps);
    free_spt(X);
    return t;
}

// KruskalModelNorm on one factor (which: 0 = MAT_NORM_2, 1 = MAT_NORM_MAX).
void ref_kruskal_norm(uint64_t rows, int rank, double* U, double* lambda, int which)
{
    IType dims[1] = {rows};
    KruskalModel* M;
    CreateKruskalModel(1, dims, (IType)rank, &M);
    memcpy(M->U[0], U, sizeof(double) * rows * rank);
    int nt = omp_get_max_threads();
    FType** sp = (FType**)malloc(sizeof(FType*) * nt);
    for (int t = 0; t < nt; ++t) sp[t] = (FType*)malloc(sizeof(FType) * rank);
    KruskalModelNorm(M, 0, which == 0 ? MAT_NORM_2 : MAT_NORM_MAX, sp);
    memcpy(U, M->U[0], sizeof(double) * rows * rank);
    memcpy(lambda, M->lambda, sizeof(double) * rank);
    for (int t = 0; t < nt; ++t) free(sp[t]);
    free(sp);
    DestroyKruskalModel(M);
}

}  // extern "C"
