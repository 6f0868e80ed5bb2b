// sptensor.hpp -- COO sparse tensor and its file formats; same struct and
// functions as the reference include/sptensor.hpp:8-42.
#pragma once

#include "blco_host.h"
#include "common.hpp"

struct sptensor_struct {
    int nmodes;    // number of modes
    IType* dims;   // length of each mode
    IType nnz;     // number of nonzeros
    IType** cidx;  // cidx[mode][nnz], 0-based
    FType* vals;   // nonzero values
};
typedef struct sptensor_struct SparseTensor;

typedef enum FileFormat_ { TEXT_FORMAT = 0, BINARY_FORMAT = 1 } FileFormat;

void ExportSparseTensor(const char* file_path, FileFormat f, SparseTensor* X);
void ImportSparseTensor(const char* file_path, FileFormat f, SparseTensor** X_);
void DestroySparseTensor(SparseTensor* X);
// cidx is nonzero-major here as in the reference: cidx[i * nmodes + n].
void CreateSparseTensor(IType mode, IType* dims, IType nnz, IType* cidx, FType* vals,
                        SparseTensor** X_);
// Synthetic Poisson tensor (the -d / -s generator of the command line).
void GenerateSparseTensor(const blco_gen_spec* spec, SparseTensor** X_);
