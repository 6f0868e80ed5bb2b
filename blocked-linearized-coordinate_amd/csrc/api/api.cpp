// api.cpp -- the reference's C++ API (sptensor.hpp, kruskal_model.hpp,
// blco.hpp, alto_dev.hpp, cpd_gpu.hpp) implemented on the C ABI of
// blco_host.h / blco_hip.h.  Errors print the library message and exit, as
// the reference's check_cuda does (src/utils.cu:131-182).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "alto_dev.hpp"
#include "blco.hpp"
#include "blco_hip.h"
#include "blco_host.h"
#include "cpd_gpu.hpp"
#include "kruskal_model.hpp"
#include "sptensor.hpp"

static void die_if(int rc, const char* what)
{
    if (rc != BLCO_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, blco_last_error());
        exit(EXIT_FAILURE);
    }
}

void* AlignedMalloc(size_t size)
{
    void* p = nullptr;
    if (posix_memalign(&p, 64, size ? size : 64) != 0) return nullptr;
    return p;
}

void AlignedFree(void* addr) { free(addr); }

// ------------------------------------------------------------- sptensor

static SparseTensor* from_coo(const blco_coo* t)
{
    int nm = 0;
    uint64_t dims[BLCO_MAX_MODES], nnz = 0;
    die_if(blco_coo_info(t, &nm, dims, &nnz), "blco_coo_info");
    SparseTensor* X = (SparseTensor*)AlignedMalloc(sizeof(SparseTensor));
    X->nmodes = nm;
    X->nnz = nnz;
    X->dims = (IType*)AlignedMalloc(sizeof(IType) * nm);
    X->cidx = (IType**)AlignedMalloc(sizeof(IType*) * nm);
    const uint64_t* c = blco_coo_cidx(t);
    for (int n = 0; n < nm; ++n) {
        X->dims[n] = dims[n];
        X->cidx[n] = (IType*)AlignedMalloc(sizeof(IType) * (nnz ? nnz : 1));
        std::memcpy(X->cidx[n], c + (uint64_t)n * nnz, sizeof(IType) * nnz);
    }
    X->vals = (FType*)AlignedMalloc(sizeof(FType) * (nnz ? nnz : 1));
    std::memcpy(X->vals, blco_coo_vals(t), sizeof(FType) * nnz);
    return X;
}

static blco_coo* to_coo(const SparseTensor* X)
{
    std::vector<uint64_t> c((uint64_t)X->nmodes * X->nnz);
    for (int n = 0; n < X->nmodes; ++n)
        std::memcpy(c.data() + (uint64_t)n * X->nnz, X->cidx[n], sizeof(IType) * X->nnz);
    blco_coo* t = nullptr;
    die_if(blco_coo_create(X->nmodes, reinterpret_cast<const uint64_t*>(X->dims), X->nnz, c.data(), X->vals, &t),
           "blco_coo_create");
    return t;
}

void ImportSparseTensor(const char* file_path, FileFormat f, SparseTensor** X_)
{
    printf("Reading in tensor...\n");
    blco_coo* t = nullptr;
    die_if(blco_coo_import(file_path, (int)f, &t), "ImportSparseTensor");
    *X_ = from_coo(t);
    blco_coo_free(t);
}

void ExportSparseTensor(const char* file_path, FileFormat f, SparseTensor* X)
{
    char fn[1024];
    if (!file_path) {
        snprintf(fn, sizeof fn, "%dD_%llu.%s", X->nmodes, X->nnz, f == TEXT_FORMAT ? "tns" : "bin");
        file_path = fn;
    }
    blco_coo* t = to_coo(X);
    die_if(blco_coo_export(t, file_path, (int)f), "ExportSparseTensor");
    blco_coo_free(t);
}

void DestroySparseTensor(SparseTensor* X)
{
    if (!X) return;
    for (int n = 0; n < X->nmodes; ++n) AlignedFree(X->cidx[n]);
    AlignedFree(X->cidx);
    AlignedFree(X->dims);
    AlignedFree(X->vals);
    AlignedFree(X);
}

void CreateSparseTensor(IType nmodes, IType* dims, IType nnz, IType* cidx, FType* vals,
                        SparseTensor** X_)
{
    std::vector<uint64_t> c(nmodes * nnz);
    for (IType i = 0; i < nnz; ++i)
        for (IType n = 0; n < nmodes; ++n) c[n * nnz + i] = cidx[i * nmodes + n];
    blco_coo* t = nullptr;
    die_if(blco_coo_create((int)nmodes, reinterpret_cast<const uint64_t*>(dims), nnz, c.data(), vals, &t),
           "CreateSparseTensor");
    *X_ = from_coo(t);
    blco_coo_free(t);
}

void GenerateSparseTensor(const blco_gen_spec* spec, SparseTensor** X_)
{
    blco_coo* t = nullptr;
    die_if(blco_generate(spec, &t), "GenerateSparseTensor");
    *X_ = from_coo(t);
    blco_coo_free(t);
}

// -------------------------------------------------------- Kruskal model

void CreateKruskalModel(int mode, IType* dims, IType rank, KruskalModel** M_)
{
    KruskalModel* M = (KruskalModel*)AlignedMalloc(sizeof(KruskalModel));
    M->mode = mode;
    M->rank = rank;
    M->dims = (IType*)AlignedMalloc(sizeof(IType) * mode);
    std::memcpy(M->dims, dims, sizeof(IType) * mode);
    M->U = (FType**)AlignedMalloc(sizeof(FType*) * mode);
    for (int n = 0; n < mode; ++n) M->U[n] = (FType*)AlignedMalloc(sizeof(FType) * dims[n] * rank);
    M->lambda = (FType*)AlignedMalloc(sizeof(FType) * rank);
    M->U_dev = nullptr;
    *M_ = M;
}

void KruskalModelRandomInit(KruskalModel* M, unsigned int seed)
{
    for (IType r = 0; r < M->rank; ++r) M->lambda[r] = 1.0;
    for (int n = 0; n < M->mode; ++n)
        die_if(blco_init_factor(seed, n, M->dims[n], (int)M->rank, M->U[n]), "KruskalModelRandomInit");
}

// L1 column normalisation absorbed into lambda (src/kruskal_model.cpp:113-140).
void KruskalModelNormalize(KruskalModel* M)
{
    for (int n = 0; n < M->mode; ++n)
        for (IType j = 0; j < M->rank; ++j) {
            FType t = 0.0;
            for (IType k = 0; k < M->dims[n]; ++k) t += std::fabs(M->U[n][k * M->rank + j]);
            for (IType k = 0; k < M->dims[n]; ++k) M->U[n][k * M->rank + j] /= t;
            M->lambda[j] *= t;
        }
}

// L2 or max(1, max) column norms (src/kruskal_model.cpp:142-259); zero L2
// columns stay zero.
void KruskalModelNorm(KruskalModel* M, IType mode, mat_norm_type which, FType** /*scratch*/)
{
    const IType R = M->rank, I = M->dims[mode];
    FType* U = M->U[mode];
    for (IType r = 0; r < R; ++r) M->lambda[r] = which == MAT_NORM_2 ? 0.0 : 1.0;
    for (IType i = 0; i < I; ++i)
        for (IType r = 0; r < R; ++r) {
            FType v = U[i * R + r];
            if (which == MAT_NORM_2) M->lambda[r] += v * v;
            else M->lambda[r] = std::max(M->lambda[r], v);
        }
    if (which == MAT_NORM_2)
        for (IType r = 0; r < R; ++r) M->lambda[r] = std::sqrt(M->lambda[r]);
    for (IType i = 0; i < I; ++i)
        for (IType r = 0; r < R; ++r)
            if (M->lambda[r] != 0.0) U[i * R + r] /= M->lambda[r];
}

// Column by column, 20 decimals (src/kruskal_model.cpp:11-42).
void ExportKruskalModel(KruskalModel* M, const char* file_path)
{
    for (int n = 0; n < M->mode; ++n) {
        char str[1100];
        snprintf(str, sizeof str, "%s.%d.out", file_path, n);
        FILE* fp = fopen(str, "w");
        if (!fp) { fprintf(stderr, "cannot open %s\n", str); exit(EXIT_FAILURE); }
        fprintf(fp, "matrix\n2\n%llu %llu\n", M->dims[n], M->rank);
        for (IType i = 0; i < M->rank; ++i)
            for (IType j = 0; j < M->dims[n]; ++j) fprintf(fp, "%.20lf\n", M->U[n][j * M->rank + i]);
        fclose(fp);
    }
    char str[1100];
    snprintf(str, sizeof str, "%s.lambda.out", file_path);
    FILE* fp = fopen(str, "w");
    if (!fp) { fprintf(stderr, "cannot open %s\n", str); exit(EXIT_FAILURE); }
    fprintf(fp, "vector\n1\n%llu\n", M->rank);
    for (IType i = 0; i < M->rank; ++i) fprintf(fp, "%.20lf\n", M->lambda[i]);
    fclose(fp);
}

void DestroyKruskalModel(KruskalModel* M)
{
    if (!M) return;
    for (int n = 0; n < M->mode; ++n) AlignedFree(M->U[n]);
    AlignedFree(M->U);
    AlignedFree(M->dims);
    AlignedFree(M->lambda);
    AlignedFree(M);
}

void PrintKruskalModel(KruskalModel* M)
{
    for (int n = 0; n < M->mode; ++n) {
        printf("mode %d:\n", n);
        for (IType j = 0; j < M->dims[n]; ++j) {
            for (IType i = 0; i < M->rank; ++i) printf("%g ", M->U[n][j * M->rank + i]);
            printf("\n");
        }
    }
    printf("lambda:");
    for (IType r = 0; r < M->rank; ++r) printf("%g ", M->lambda[r]);
    printf("\n");
}

void RedistributeLambda(KruskalModel* M, int n)
{
    for (IType r = 0; r < M->rank; ++r) {
        for (IType i = 0; i < M->dims[n]; ++i) M->U[n][i * M->rank + r] *= M->lambda[r];
        M->lambda[r] = 1.0;
    }
}

double KruskalTensorFit() { return 0.0; }

KruskalModel* make_device_copy(KruskalModel* M_host)
{
    KruskalModel* D = new KruskalModel;
    D->mode = M_host->mode;
    D->rank = M_host->rank;
    D->dims = M_host->dims;
    D->U = new FType*[M_host->mode];
    for (int n = 0; n < M_host->mode; ++n) {
        uint64_t bytes = sizeof(FType) * M_host->dims[n] * M_host->rank;
        die_if(blco_malloc((void**)&D->U[n], bytes), "make_device_copy");
        die_if(blco_memcpy_htod(D->U[n], M_host->U[n], bytes), "make_device_copy");
    }
    die_if(blco_malloc((void**)&D->lambda, sizeof(FType) * M_host->rank), "make_device_copy");
    die_if(blco_memcpy_htod(D->lambda, M_host->lambda, sizeof(FType) * M_host->rank), "make_device_copy");
    D->U_dev = D->U;  // the C ABI takes a host array of device pointers
    return D;
}

void destroy_kruskal_model_dev(KruskalModel* D)
{
    if (!D) return;
    for (int n = 0; n < D->mode; ++n) blco_free(D->U[n]);
    blco_free(D->lambda);
    delete[] D->U;
    delete D;
}

// ----------------------------------------------------------------- BLCO

blcotensor* gen_blcotensor_host_any(SparseTensor* X, IType max_block_size)
{
    double t0 = omp_get_wtime();
    blco_coo* t = to_coo(X);
    blcotensor* B = new blcotensor;
    die_if(blco_host_encode_coo(t, max_block_size, &B->host), "gen_blcotensor_host");
    blco_coo_free(t);
    int nm = 0;
    uint64_t dims[BLCO_MAX_MODES], nnz = 0, nb = 0, mx = 0;
    die_if(blco_host_info(B->host, &nm, dims, &nnz, &nb, &mx), "blco_host_info");
    B->N = nm;
    B->modes = new IType[nm];
    B->mode_masks = new IType[nm];
    B->mode_pos = new int[nm];
    std::memcpy(B->modes, dims, sizeof(IType) * nm);
    blco_host_layout(B->host, reinterpret_cast<uint64_t*>(B->mode_masks), B->mode_pos, nullptr, nullptr);
    B->total_nnz = nnz;
    B->max_nnz = mx;
    B->block_count = nb;
    B->max_block_size = max_block_size;
    B->coords = reinterpret_cast<IType*>(const_cast<uint64_t*>(blco_host_coords(B->host)));
    B->values = const_cast<FType*>(blco_host_values(B->host));
    std::vector<uint64_t> bs(nb), bn(nb), bc(nb * nm);
    blco_host_blocks(B->host, bs.data(), bn.data(), bc.data());
    B->blocks = new blco_block*[nb ? nb : 1];
    for (uint64_t b = 0; b < nb; ++b) {
        blco_block* k = gen_block_host(nm, bn[b], B);
        k->idx = B->coords + bs[b];
        k->vals = B->values + bs[b];
        k->dims = B->modes;
        k->mode_masks = B->mode_masks;
        k->mode_pos = B->mode_pos;
        for (int n = 0; n < nm; ++n) k->block_coords[n] = bc[b * nm + n];
        B->blocks[b] = k;
    }
    printf("BLCO: construction time = %f (s)\nTotal blocks: %llu\n\n", omp_get_wtime() - t0,
           (unsigned long long)nb);
    return B;
}

static uint64_t g_budget = 0, g_budget_slot = 0;
void cpals_set_device_budget(uint64_t bytes, uint64_t slot_nnz)
{
    g_budget = bytes;
    g_budget_slot = slot_nnz;
}

// The device tensor of the drivers: under a device budget the budget decides
// what is resident (slots of the streaming block size when streaming was
// asked for, else of the budget's own slot size).
static int create_dev(const blco_host* h, int value_dtype, bool stream_data, uint64_t slot_nnz, blco_dev** out)
{
    if (g_budget) return blco_dev_create_budget(h, value_dtype, g_budget, stream_data ? slot_nnz : g_budget_slot, out);
    return blco_dev_create(h, value_dtype, stream_data ? 1 : 0, stream_data ? slot_nnz : 0, out);
}

blcotensor* gen_blcotensor_device(blcotensor* X, IType /*block_count*/, IType MB,
                                  bool /*do_batching*/, int value_dtype)
{
    blcotensor* D = new blcotensor;
    die_if(create_dev(X->host, value_dtype, MB > 0, MB, &D->dev), "gen_blcotensor_device");
    D->N = X->N;
    D->modes = new IType[X->N];
    std::memcpy(D->modes, X->modes, sizeof(IType) * X->N);
    D->total_nnz = X->total_nnz;
    D->max_nnz = X->max_nnz;
    D->block_count = X->block_count;
    D->max_block_size = MB;
    D->mode_masks = new IType[X->N];
    D->mode_pos = new int[X->N];
    std::memcpy(D->mode_masks, X->mode_masks, sizeof(IType) * X->N);
    std::memcpy(D->mode_pos, X->mode_pos, sizeof(int) * X->N);
    D->on_device = true;
    // block records of the device tensor: device coordinates (and f64
    // values) of each resident block, host block coordinates
    uint64_t* dc = nullptr;
    void* dv = nullptr;
    int vdt = 0;
    die_if(blco_dev_arrays(D->dev, &dc, &dv, &vdt, nullptr), "gen_blcotensor_device");
    D->blocks = new blco_block*[X->block_count ? X->block_count : 1];
    uint64_t off = 0;
    for (IType b = 0; b < X->block_count; ++b) {
        const blco_block* hb = X->blocks[b];
        blco_block* k = gen_block_device(X->N, hb->nnz, D);
        std::memcpy(k->block_coords, hb->block_coords, sizeof(IType) * X->N);
        k->dims = D->modes;
        k->mode_masks = D->mode_masks;
        k->mode_pos = D->mode_pos;
        if (dc) k->idx = reinterpret_cast<IType*>(dc) + off;
        if (dv && vdt == BLCO_F64) k->vals = static_cast<FType*>(dv) + off;
        off += hb->nnz;
        D->blocks[b] = k;
    }
    return D;
}

void delete_blcotensor_host(blcotensor* B)
{
    if (!B) return;
    for (IType b = 0; b < B->block_count; ++b) delete_block_host(B->blocks[b]);
    delete[] B->blocks;
    delete[] B->modes;
    delete[] B->mode_masks;
    delete[] B->mode_pos;
    blco_host_free(B->host);
    delete B;
}

void delete_blcotensor_device(blcotensor* D)
{
    if (!D) return;
    if (D->blocks)
        for (IType b = 0; b < D->block_count; ++b) delete_block_device(D->blocks[b]);
    delete[] D->blocks;
    blco_dev_free(D->dev);
    delete[] D->modes;
    delete[] D->mode_masks;
    delete[] D->mode_pos;
    delete D;
}

void send_blcotensor_over(hipStream_t, blcotensor*, blcotensor*) { blco_synchronize(); }
void send_block_over(hipStream_t, blco_block*, blco_block*) { blco_synchronize(); }

blco_block* gen_block_host(IType N, IType nnz, blcotensor* parent)
{
    blco_block* b = new blco_block;
    b->nmode = (int)N;
    b->nnz = nnz;
    b->parent = parent;
    b->block_coords = new IType[N];
    return b;
}

blco_block* gen_block_device(IType N, IType nnz, blcotensor* parent) { return gen_block_host(N, nnz, parent); }

void delete_block_host(blco_block* b)
{
    if (!b) return;
    delete[] b->block_coords;
    delete b;
}

void delete_block_device(blco_block* b) { delete_block_host(b); }

// ------------------------------------------------------------ MTTKRP

void send_masks_over(blcotensor*) {}

FType** create_da_mem_dev(int, int, IType, int, blcotensor*) { return nullptr; }

void mttkrp_alto_dev_onemode_any(blcotensor* at_dev, IType rank, FType* output, int kernel,
                                 int target_mode, IType thread_cf, int nprtn, FType** mats_dev)
{
    die_if(blco_mttkrp(at_dev->dev, target_mode, (int)rank, BLCO_F64, (void* const*)mats_dev, output,
                       kernel, nprtn, (int)thread_cf, nullptr),
           "mttkrp_alto_dev_onemode");
}

void mttkrp_alto_dev_onemode_ref(blcotensor* /*at_host*/, blcotensor* at_dev, KruskalModel* M, FType* output,
                                 int kernel, int target_mode, IType thread_cf, bool /*stream_data*/, int nprtn,
                                 FType** mats_staging_ptr, FType** /*mats_dev*/, FType** /*partial_matrices*/,
                                 FType** /*partial_matrices_dev*/)
{
    if (!at_dev || !at_dev->dev || !M || !mats_staging_ptr) {
        fprintf(stderr, "mttkrp_alto_dev_onemode: null device tensor, model or factors\n");
        exit(EXIT_FAILURE);
    }
    mttkrp_alto_dev_onemode_any(at_dev, M->rank, output, kernel, target_mode, thread_cf, nprtn, mats_staging_ptr);
}

void mttkrp_alto_dev_any(blcotensor* at_host, KruskalModel* M, int kernel, int iters, int target_mode,
                         IType thread_cf, bool stream_data, bool do_batching, int nprtn)
{
    printf("\nCopying to GPU device\n--> Init BLCO on GPU\n");
    blcotensor* D = gen_blcotensor_device(at_host, at_host->block_count,
                                          stream_data ? at_host->max_block_size : 0, do_batching);
    printf("--> Copy over Kruskal model (factor matrices)\n");
    KruskalModel* Md = make_device_copy(M);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int first = target_mode < 0 ? 0 : target_mode;
    const int last = target_mode < 0 ? (int)at_host->N - 1 : target_mode;
    for (int n = first; n <= last; ++n) {
        // the reference accumulates straight into U[n] (src/alto_dev.cu:1917)
        float etime = 0.f;
        printf("--> Start MTTKRP mode %d (%d iters)\n", n, iters);
        for (int i = 0; i < iters; ++i) {
            (void)hipDeviceSynchronize();
            (void)hipEventRecord(e0, 0);
            mttkrp_alto_dev_onemode_any(D, M->rank, Md->U[n], kernel, n, thread_cf, nprtn, Md->U);
            (void)hipDeviceSynchronize();
            (void)hipEventRecord(e1, 0);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            etime += ms;
        }
        printf("Total time for MTTKRP: %0.3f (time per iter: %0.3f ms)\n", etime, etime / std::max(iters, 1));
        die_if(blco_memcpy_dtoh(M->U[n], Md->U[n], sizeof(FType) * M->dims[n] * M->rank), "copy back");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    destroy_kruskal_model_dev(Md);
    delete_blcotensor_device(D);
}

// ------------------------------------------------------------ CP-ALS

static int g_precision = 0, g_mode_order = 0, g_graph = 0;
void cpals_set_precision(int p) { g_precision = p; }
void cpals_set_mode_order(int o) { g_mode_order = o; }
void cpals_set_graph(int g) { g_graph = g; }

FType cpals_blco_dev(blcotensor* X, KruskalModel* A, int maxiter, FType tolerance, int kernel,
                     bool stream_data, bool /*do_batching*/, int thread_cf, int nprtn)
{
    printf("\nCopying to GPU device\n--> Init BLCO on GPU\n");
    double t0 = omp_get_wtime();
    blco_dev* d = nullptr;
    die_if(create_dev(X->host, g_precision == 0 ? BLCO_F64 : BLCO_F32, stream_data, X->max_block_size, &d),
           "gen_blcotensor_device");
    printf("--> Total initialization = %0.3f (s)\n\n", omp_get_wtime() - t0);
    blco_cpals_opts o;
    std::memset(&o, 0, sizeof o);
    o.rank = (int)A->rank;
    o.max_iters = maxiter;
    o.tolerance = tolerance;
    o.kernel = kernel;
    o.precision = g_precision;
    o.nprtn = nprtn;
    o.thread_cf = thread_cf;
    o.mode_order = g_mode_order;
    o.use_graph = g_graph;
    o.verbose = 1;
    std::vector<double> fits(maxiter > 0 ? maxiter : 1);
    int iters = 0;
    double fit = 0.0;
    die_if(blco_cpals(d, &o, A->U, A->lambda, fits.data(), &iters, &fit), "cpals_blco_dev");
    blco_dev_free(d);
    return fit;
}

FType kruskal_fit_gpu(hipStream_t v_stream, FType** grams, KruskalModel* A, blcotensor* X, FType* /*work*/,
                      FType* last_mttkrp, FType X_norm, IType mode, IType /*blocks*/)
{
    if (!A || !X || mode >= (IType)A->mode) {
        fprintf(stderr, "kruskal_fit_gpu: bad model, tensor or mode\n");
        exit(EXIT_FAILURE);
    }
    return kruskal_fit_gpu_rows(v_stream, grams, A, (IType)A->mode, last_mttkrp, X->modes[mode], X_norm, mode);
}

FType kruskal_fit_gpu_rows(hipStream_t stream, FType** grams, KruskalModel* A_dev, IType nmodes,
                           FType* last_mttkrp, IType last_rows, FType X_norm, IType mode)
{
    double fit = 0.0;
    die_if(blco_fit(grams, (int)nmodes, (int)A_dev->rank, A_dev->lambda, A_dev->U[mode], last_mttkrp,
                    BLCO_F64, last_rows, X_norm * X_norm, &fit, stream),
           "kruskal_fit_gpu");
    return fit;
}

void vector_hadamard(hipStream_t stream, FType* x, FType* y, IType n)
{
    die_if(blco_hadamard(x, y, n, stream), "vector_hadamard");
}

void value_fill(hipStream_t stream, FType* x, IType n, FType val)
{
    die_if(blco_value_fill(x, n, val, stream), "value_fill");
}

void vecsum(hipStream_t stream, FType* x, IType n) { die_if(blco_vecsum(x, n, stream), "vecsum"); }

void pseudoinverse_gpu(hipStream_t stream, FType* A, IType n, FType* /*work*/, IType /*lwork*/, int* info)
{
    die_if(blco_pinv(A, n, info, stream), "pseudoinverse_gpu");
}

int normal_solve_gpu(hipStream_t stream, FType** grams, IType nmodes, IType mode, IType rank, FType* M,
                     FType* U, IType rows)
{
    int fb = 0;
    die_if(blco_solve(grams, (int)nmodes, (int)mode, (int)rank, M, U, BLCO_F64, rows, &fb, stream),
           "normal_solve_gpu");
    return fb;
}

void columnwise_normscal(hipStream_t stream, IType /*blocks*/, FType* mat, IType rank, IType mode_length,
                         FType* lambda)
{
    columnwise_normscal_which(stream, mat, rank, mode_length, lambda, MAT_NORM_2);
}

void columnwise_normscal_which(hipStream_t stream, FType* mat, IType rank, IType mode_length, FType* lambda,
                               mat_norm_type which)
{
    die_if(blco_colnorm(mat, BLCO_F64, mode_length, (int)rank, which == MAT_NORM_2 ? 0 : 1, lambda, stream),
           "columnwise_normscal");
}

void ata_gpu(hipStream_t stream, const FType* A, FType* C, IType m, IType n)
{
    die_if(blco_gram(A, BLCO_F64, m, (int)n, C, stream), "ata_gpu");
}
