// kruskal_model.hpp -- Kruskal (CP) model; same struct and functions as the
// reference include/kruskal_model.hpp:8-49.  U[n] is row-major dims[n] x
// rank.  U_dev holds device copies made by make_device_copy().
#pragma once

#include "common.hpp"

typedef struct KruskalModel_ {
    int mode;
    IType* dims;
    IType rank;
    FType** U;      // factors, dims[n] x rank, row-major
    FType* lambda;  // rank
    FType** U_dev;  // device factors (make_device_copy)
} KruskalModel;

typedef enum { MAT_NORM_2, MAT_NORM_MAX } mat_norm_type;

void CreateKruskalModel(int mode, IType* dim, IType rank, KruskalModel** M_);
void KruskalModelNormalize(KruskalModel* M);
void KruskalModelNorm(KruskalModel* M, IType mode, mat_norm_type which, FType** scratchpad);
// Uniform [0,1) entries from a counter-based generator keyed by (seed, mode,
// row, column); lambda = 1.  Independent of the thread count.
void KruskalModelRandomInit(KruskalModel* M, unsigned int seed);
void ExportKruskalModel(KruskalModel* M, const char* file_path);
void DestroyKruskalModel(KruskalModel* M);
void PrintKruskalModel(KruskalModel* M);
void RedistributeLambda(KruskalModel* M, int n);
double KruskalTensorFit();

// Replicates the model on the device (f64 factors) and destroys the copy.
KruskalModel* make_device_copy(KruskalModel* M_host);
void destroy_kruskal_model_dev(KruskalModel* M_dev);
