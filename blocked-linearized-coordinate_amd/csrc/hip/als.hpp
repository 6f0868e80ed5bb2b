// als.hpp -- launchers of als.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "blco_host.h"

namespace blco {

struct GramSet {
    const double* g[BLCO_MAX_MODES];
};

constexpr int GRAM_MAX_PARTS = 2048;  // partial Gram blocks (workgroups) at most
int gram_parts(uint64_t rows, int* nparts, uint64_t* rows_per_wg);
// G = U^T U (f64, R x R) from dtype U; with M, ip[col] = sum_i U_ic M_ic.
// part needs gram_parts(rows) * (R*R + R) doubles.
int launch_gram(const void* U, int dtype, uint64_t rows, int R, const void* M,
                double* part, double* G, double* ip, hipStream_t st);
// Machine epsilon of the precision the Grams of factor_dtype factors are
// accumulated in (f64 for f64 factors, f32 for f32 and bf16 ones).
double gram_eps(int factor_dtype);
// fac: R x R (V^{-1}, or V^+ with status 1); work: 2 R^2 + R doubles.  eps
// scales the Cholesky pivot floor and the pseudo-inverse tolerance.  With
// dinv (nmodes x R) colscale[k] receives prod_{m != mode} dinv[m][k].
int launch_factorize(const GramSet& gs, int nmodes, int mode, int R, double eps, double* fac,
                     double* work, int* status, hipStream_t st, const double* dinv = nullptr,
                     double* colscale = nullptr);
// In-place pseudo-inverse of a general n x n matrix (n <= 64); info
// (device, may be null): 0 converged, else the sweep count.
int launch_pinv_general(double* A, int n, int* info, hipStream_t st);
uint32_t solve_parts(uint64_t rows);
// U = M F^T row by row; stats: solve_parts(rows) * R doubles.
int launch_solve_rows(const void* M, void* U, int dtype, uint64_t rows, int R, const double* fac,
                      int which, double* stats, hipStream_t st);
// A mode's whole ALS update after its MTTKRP, in one pass over M and U on
// MFMA: U = M diag(colscale) F^T (colscale may be null), lambda, inv and
// dinv (may be null) from U's column statistics, G = diag(inv) U^T U
// diag(inv), and with want_ip ip = inv o colscale o sum_i U_i M_i.  stats:
// solve_gram_parts(rows) * R doubles; part: as for launch_gram.
uint32_t solve_gram_parts(uint64_t rows);
int launch_solve_gram(const void* M, void* U, int dtype, uint64_t rows, int R, const double* fac,
                      const double* colscale, int which, bool want_ip, double* stats, double* part, double* lambda,
                      double* inv, double* dinv, double* G, double* ip, hipStream_t st);
// lambda, 1/lambda (and into dinv when non-null) from the row solve's
// statistics (folded in place).
int launch_lambda(double* stats, int nparts, int R, int which, double* lambda, double* inv, hipStream_t st,
                  double* dinv = nullptr);
// out[0] = fit, out[1] = inner product, out[2] = ||M||^2.
int launch_fit(const GramSet& gs, int nmodes, int R, const double* lambda, const double* ip,
               double normX2, double* out, hipStream_t st);
int launch_sumsq(const void* x, int dtype, uint64_t n, double* part, int nparts, double* out,
                 hipStream_t st);
int launch_hadamard(double* x, const double* y, uint64_t n, hipStream_t st);

int launch_colstats(const void* U, int dtype, uint64_t rows, int R, int which, double* stats,
                    int nparts, hipStream_t st);
int launch_scale(void* U, int dtype, uint64_t rows, int R, const double* inv, hipStream_t st);
int launch_vecsum(double* x, uint64_t n, hipStream_t st);

}  // namespace blco
