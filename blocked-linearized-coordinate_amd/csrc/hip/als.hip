// als.hip -- the ALS step of CP-ALS on gfx950 (src/cpd_gpu.cu:21-242 of the
// reference does it with cuBLAS / cuSOLVER; the semantics kept here are the
// CPU reference's, include/cpd.hpp:44-272).
//
//   gram       G = U^T U on MFMA: v_mfma_f64_16x16x4f64 for f64 factors,
//              v_mfma_f32_16x16x4f32 for f32/bf16 factors.  One wave walks
//              4-row steps of U; lane l holds U[r0 + l/16][16 t + l%16], which
//              is both the A operand (A[i][k] = U[r0+k][16 ti + i]) and the B
//              operand of tile (ti, tj).  The accumulator layouts differ: the
//              f32 form leaves D[4*(l/16) + v][l%16] in register v of lane l,
//              the f64 form D[l/16 + 4*v][l%16] (checked on the hardware).  Per-workgroup partial Grams are
//              summed in a fixed order (deterministic).  Optionally the same
//              pass forms the fit inner products sum_i U_ir M_ir.
//   factorize  one workgroup: V = prod_{m != n} G_m (Hadamard), lower
//              Cholesky in LAPACK dpotf2 order in LDS; when a pivot falls to
//              R * eps * max V_ii or below, cyclic Jacobi eigendecomposition
//              of V and the pseudo-inverse V^+ = Q diag(1/w) Q^T with
//              tolerance R * ulp(max |w|) (src/cpd_gpu.cu:487-489), where eps
//              and ulp are those of the precision the Grams were accumulated
//              in (f64, or f32 for the f32 / bf16 factor forms).
//              On success the explicit inverse V^{-1} = L^{-T} L^{-1} is
//              formed (so both outcomes leave one R x R matrix F).
//   solve rows a group of R lanes per row: u = F m with the row of M staged
//              in LDS; the same pass forms per-workgroup column statistics
//              for the norm (sum of squares, or max).
//   solve+gram the CP-ALS update of a mode in one pass on MFMA (k_solve_gram):
//              the solve's result registers feed the Gram tiles directly.
//   lambda     lambda_r = sqrt(sum) (L2) or max(1, max) (max norm,
//              src/kruskal_model.cpp:183-223), and 1/lambda (0 for a zero
//              column), from the statistics folded in two passes.
//   fit        one workgroup: fit = 1 - r / ||X|| with
//              r^2 = ||X||^2 + |sum_ij l_i l_j prod_n G_n(i,j)| - 2 sum_r l_r ip_r
//              (include/cpd.hpp:190-272).
#include "als.hpp"
#include "devcommon.hpp"

namespace blco {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> __device__ __forceinline__ float as_f32(T x) { return (float)x; }
template <> __device__ __forceinline__ float as_f32<bf16>(bf16 x) { return bf16_to_f32(x.bits); }
template <typename T> __device__ __forceinline__ double as_f64(T x) { return (double)x; }
template <> __device__ __forceinline__ double as_f64<bf16>(bf16 x) { return (double)bf16_to_f32(x.bits); }
template <typename T> __device__ __forceinline__ bool is_zero(T x) { return x == (T)0; }

// ------------------------------------------------------------------ gram

// Partial Gram (and optional scaling / inner products) of rows
// [g * rows_per_wg, ...) for workgroup g.  T tiles of 16 columns.
template <typename UT, int T>
__global__ __launch_bounds__(WG) void k_gram(const UT* __restrict__ U, uint64_t rows, int R,
                                             uint64_t rows_per_wg,
                                             const void* __restrict__ M,
                                             double* __restrict__ part)
{
    constexpr bool F64 = std::is_same<UT, double>::value;
    using AT = typename std::conditional<F64, double, float>::type;
    using VT = typename std::conditional<F64, d4, f4>::type;
    constexpr int NT = T * (T + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i16 = lane & 15, k4 = lane >> 4;
    const uint64_t rb = (uint64_t)blockIdx.x * rows_per_wg;
    const uint64_t re = min(rb + rows_per_wg, rows);
    VT acc[NT];
    #pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = VT{0, 0, 0, 0};
    double ip[T];
    #pragma unroll
    for (int t = 0; t < T; ++t) ip[t] = 0.0;
    for (uint64_t r0 = rb + (uint64_t)w * 4; r0 < re; r0 += WAVES_PER_WG * 4) {
        const uint64_t row = r0 + k4;
        AT a[T];
        #pragma unroll
        for (int t = 0; t < T; ++t) {
            const int col = t * 16 + i16;
            AT x = 0;
            if (row < re && col < R) {
                x = (AT)as_f64(U[row * (uint64_t)R + col]);
                if (M) {
                    double mv = F64 ? reinterpret_cast<const double*>(M)[row * (uint64_t)R + col]
                                    : (double)reinterpret_cast<const float*>(M)[row * (uint64_t)R + col];
                    ip[t] += (double)x * mv;
                }
            }
            a[t] = x;
        }
        int q = 0;
        #pragma unroll
        for (int ti = 0; ti < T; ++ti)
            #pragma unroll
            for (int tj = ti; tj < T; ++tj) {
                if constexpr (F64)
                    acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], a[tj], acc[q], 0, 0, 0);
                else
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ti], a[tj], acc[q], 0, 0, 0);
                ++q;
            }
    }
    // combine the 4 waves in LDS, then write this workgroup's partial
    __shared__ double s_g[WAVES_PER_WG][T * 16][T * 16 + 1];
    __shared__ double s_ip[WAVES_PER_WG][T * 16];
    int q = 0;
    #pragma unroll
    for (int ti = 0; ti < T; ++ti)
        #pragma unroll
        for (int tj = ti; tj < T; ++tj) {
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                // D register v of lane l holds row 4*(l/16) + v (f32 form) or
                // row l/16 + 4*v (f64 form), column l % 16 of the tile
                int i = ti * 16 + (F64 ? k4 + 4 * v : 4 * k4 + v), j = tj * 16 + i16;
                double x = (double)acc[q][v];
                s_g[w][i][j] = x;
                s_g[w][j][i] = x;
            }
            ++q;
        }
    #pragma unroll
    for (int t = 0; t < T; ++t) {
        // lanes with equal i16 hold partial sums of the same column
        double x = ip[t];
        x += __shfl_xor(x, 16);
        x += __shfl_xor(x, 32);
        if (k4 == 0) s_ip[w][t * 16 + i16] = x;
    }
    __syncthreads();
    double* out = part + (uint64_t)blockIdx.x * (uint64_t)(R * R + R);
    for (int e = threadIdx.x; e < R * R; e += WG) {
        int i = e / R, j = e % R;
        out[e] = ((s_g[0][i][j] + s_g[1][i][j]) + s_g[2][i][j]) + s_g[3][i][j];
    }
    for (int e = threadIdx.x; e < R; e += WG)
        out[R * R + e] = ((s_ip[0][e] + s_ip[1][e]) + s_ip[2][e]) + s_ip[3][e];
}

// Deterministic reduction of nparts rows of W doubles (part[p][e], row
// major) over p, in two passes that both fill the GPU.  Pass 1: workgroup
// (x, y) takes entries [64 x, 64 x + 64) of the rows p = y (mod gridDim.y);
// wave w of it folds the rows y + P2 (w + 4 j), j = 0, 1, ..., with two
// accumulators, the four waves' results are combined in wave order and
// written over row y (only this workgroup reads row y's entries, and it
// has read them all by then).  Pass 2 folds rows 0..P2-1 in order.  which:
// 0 sum (seed 0), 1 max seeded with 1.
constexpr int RED_SPLIT = 16;

__device__ __forceinline__ double fold_op(int which, double a, double b) { return which == 0 ? a + b : fmax(a, b); }

__global__ __launch_bounds__(WG) void k_fold_parts(double* __restrict__ part, int nparts, int W, int which)
{
    __shared__ double s_sum[WAVES_PER_WG][WAVE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = blockIdx.x * WAVE + lane, y = blockIdx.y, P2 = gridDim.y;
    const double seed = which == 0 ? 0.0 : 1.0;
    double a0 = seed, a1 = seed;
    if (e < W) {
        int p = y + P2 * w;
        for (; p + WAVES_PER_WG * P2 < nparts; p += 2 * WAVES_PER_WG * P2) {
            a0 = fold_op(which, a0, part[(uint64_t)p * W + e]);
            a1 = fold_op(which, a1, part[(uint64_t)(p + WAVES_PER_WG * P2) * W + e]);
        }
        if (p < nparts) a0 = fold_op(which, a0, part[(uint64_t)p * W + e]);
    }
    s_sum[w][lane] = fold_op(which, a0, a1);
    __syncthreads();
    if (w == 0 && e < W)
        part[(uint64_t)y * W + e] =
            fold_op(which, fold_op(which, fold_op(which, s_sum[0][lane], s_sum[1][lane]), s_sum[2][lane]), s_sum[3][lane]);
}

static int fold_parts(double* part, int nparts, int W, int which, int* rows_left, hipStream_t st)
{
    const int p2 = nparts < RED_SPLIT ? (nparts < 1 ? 1 : nparts) : RED_SPLIT;
    hipLaunchKernelGGL(k_fold_parts, dim3((W + WAVE - 1) / WAVE, p2), dim3(WG), 0, st, part, nparts, W, which);
    HIP_TRY(hipGetLastError());
    *rows_left = p2;
    return BLCO_OK;
}

// Pass 2 of the Gram reduction: G (R*R) and ip (R) from the P2 folded rows
// of (R*R + R).  With inv (the new lambda's inverse), G is that of the
// factor scaled by inv, G_ij = inv_i inv_j sum, and ip_r = inv_r c_r sum,
// where c (colscale, may be null: ones) is the column scale the MTTKRP
// result carries relative to that of the normalised factors.
__global__ __launch_bounds__(WAVE) void k_gram_finish(const double* __restrict__ part, int p2, int R,
                                                      const double* __restrict__ inv,
                                                      const double* __restrict__ colscale,
                                                      double* __restrict__ G, double* __restrict__ ip)
{
    const int e = blockIdx.x * WAVE + threadIdx.x, W = R * R + R;
    if (e >= W) return;
    double s = part[e];
    for (int y = 1; y < p2; ++y) s += part[(uint64_t)y * W + e];
    if (e < R * R) {
        if (inv) s *= inv[e / R] * inv[e % R];
        G[e] = s;
    } else if (ip) {
        const int r = e - R * R;
        if (inv) s *= inv[r] * (colscale ? colscale[r] : 1.0);
        ip[r] = s;
    }
}

// ------------------------------------------------------------- factorize

// status: 0 V^{-1} in fac (by Cholesky), 1 pseudo-inverse in fac.
// The Cholesky factor is formed right-looking (all trailing entries updated
// in parallel after each column), which subtracts the terms of every entry
// in the same order k = 0, 1, ... as dpotf2's column dot products.  L^{-1}
// is formed row by row with the same parallel pattern, then
// V^{-1} = L^{-T} L^{-1}.
__global__ __launch_bounds__(WG) void k_factorize(GramSet gs, int nmodes, int mode, int R,
                                                  double eps, double* __restrict__ fac,
                                                  double* __restrict__ work,
                                                  int* __restrict__ status,
                                                  const double* __restrict__ dinv,
                                                  double* __restrict__ colscale)
{
    extern __shared__ double s_a[];  // R x R: V, then L (lower); then R x R X = L^{-1}
    double* s_x = s_a + R * R;
    const int tid = threadIdx.x;
    // the column scale of this mode's MTTKRP result: the product of the other
    // factors' pending inverse scales (their stored columns carry lambda)
    if (colscale)
        for (int k = tid; k < R; k += WG) {
            double c = 1.0;
            for (int m = 0; m < nmodes; ++m)
                if (m != mode) c *= dinv[m * R + k];
            colscale[k] = c;
        }
    for (int e = tid; e < R * R; e += WG) {
        double v = 1.0;
        for (int m = 0; m < nmodes; ++m)
            if (m != mode) v *= gs.g[m][e];
        s_a[e] = v;
        work[e] = v;  // keep V for the fallback
        s_x[e] = (e / R == e % R) ? 1.0 : 0.0;
    }
    __syncthreads();
    // A pivot at or below R * eps * max_i V_ii means V is singular to working
    // precision its Grams carry: its Cholesky factor would carry pivots of
    // rounding noise (and the solve factors of that size), so the
    // pseudo-inverse takes over -- the threshold the Jacobi fallback uses for
    // eigenvalues.  eps is that of the Grams' accumulation (2^-52 for f64
    // factors, 2^-23 for f32 / bf16 ones: an f32 V that is singular to f32
    // precision clears an f64 floor or not depending on the summation order).
    // dpotrf's own test (a pivot <= 0) is the special case of an exactly
    // singular V.
    double dmax = 0.0;
    for (int i = 0; i < R; ++i) dmax = fmax(dmax, s_a[i * R + i]);
    const double floor_piv = R * eps * dmax;
    bool ok = true;
    for (int j = 0; j < R; ++j) {
        const double ajj = s_a[j * R + j];
        if (!(ajj > floor_piv)) {  // every thread reads the same value: a uniform exit
            ok = false;
            break;
        }
        const double piv = sqrt(ajj);
        __syncthreads();  // all have read a_jj
        for (int i = j + 1 + tid; i < R; i += WG) s_a[i * R + j] /= piv;
        if (tid == 0) s_a[j * R + j] = piv;
        __syncthreads();
        const int n = R - 1 - j;
        for (int t = tid; t < n * n; t += WG) {
            const int i = j + 1 + t / n, k = j + 1 + t % n;
            if (k <= i) s_a[i * R + k] -= s_a[i * R + j] * s_a[k * R + j];
        }
        __syncthreads();
    }
    if (ok) {
        for (int k = 0; k < R; ++k) {
            const double lkk = s_a[k * R + k];
            for (int c = tid; c <= k; c += WG) s_x[k * R + c] /= lkk;
            __syncthreads();
            const int n = R - 1 - k, w = k + 1;
            for (int t = tid; t < n * w; t += WG) {
                const int i = k + 1 + t / w, c = t % w;
                s_x[i * R + c] -= s_a[i * R + k] * s_x[k * R + c];
            }
            __syncthreads();
        }
        for (int e = tid; e < R * R; e += WG) {
            int a = e / R, b = e % R;
            int k0 = a > b ? a : b;
            double t = 0.0;
            for (int k = k0; k < R; ++k) t += s_x[k * R + a] * s_x[k * R + b];
            fac[e] = t;
        }
        if (tid == 0) *status = 0;
        return;
    }
    __syncthreads();
    // ---- fallback: symmetric pseudo-inverse by cyclic Jacobi
    double* A = work;          // R x R, V (destroyed)
    double* Q = work + R * R;  // R x R eigenvectors (columns)
    double* wv = work + 2 * R * R;
    for (int e = tid; e < R * R; e += WG) Q[e] = (e / R == e % R) ? 1.0 : 0.0;
    // s_done (sweep converged) and s_skip (one pair's rotation skipped) are
    // separate flags, and every thread has read s_done before thread 0 can
    // write either again, so no wave acts on a flag of another step.
    __shared__ int s_done, s_skip;
    __shared__ double s_c, s_s;
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        if (tid == 0) {
            double off = 0.0, diag = 0.0;
            for (int i = 0; i < R; ++i)
                for (int j = 0; j < R; ++j) {
                    double x = A[i * R + j];
                    if (i != j) off += x * x;
                    else diag += x * x;
                }
            s_done = (off <= 1e-30 * diag || off == 0.0);
        }
        __syncthreads();
        const bool converged = s_done != 0;
        __syncthreads();
        if (converged) break;
        for (int p = 0; p < R - 1; ++p)
            for (int qq = p + 1; qq < R; ++qq) {
                if (tid == 0) {
                    double apq = A[p * R + qq];
                    if (apq == 0.0) {
                        s_c = 1.0; s_s = 0.0; s_skip = 1;
                    } else {
                        double app = A[p * R + p], aqq = A[qq * R + qq];
                        double tau = (aqq - app) / (2.0 * apq);
                        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        double c = 1.0 / sqrt(1.0 + t * t);
                        s_c = c; s_s = t * c; s_skip = 0;
                    }
                }
                __syncthreads();
                if (s_skip) { __syncthreads(); continue; }
                const double c = s_c, s = s_s;
                for (int k = tid; k < R; k += WG) {
                    double akp = A[k * R + p], akq = A[k * R + qq];
                    A[k * R + p] = c * akp - s * akq;
                    A[k * R + qq] = s * akp + c * akq;
                }
                __threadfence_block();
                __syncthreads();
                for (int k = tid; k < R; k += WG) {
                    double apk = A[p * R + k], aqk = A[qq * R + k];
                    A[p * R + k] = c * apk - s * aqk;
                    A[qq * R + k] = s * apk + c * aqk;
                    double qkp = Q[k * R + p], qkq = Q[k * R + qq];
                    Q[k * R + p] = c * qkp - s * qkq;
                    Q[k * R + qq] = s * qkp + c * qkq;
                }
                __threadfence_block();
                __syncthreads();
            }
    }
    if (tid == 0) {
        double emax = 0.0;
        for (int i = 0; i < R; ++i) emax = fmax(emax, fabs(A[i * R + i]));
        // R * ulp(emax) in the Grams' precision; with eps = 2^-52 this is
        // R * (nextafter(emax, +inf) - emax) of src/cpd_gpu.cu:487-489
        double tol = emax > 0.0 ? R * ldexp(eps, ilogb(emax)) : 0.0;
        for (int i = 0; i < R; ++i) {
            double x = A[i * R + i];
            wv[i] = (fabs(x) > tol) ? 1.0 / x : 0.0;
        }
    }
    __threadfence_block();
    __syncthreads();
    for (int e = tid; e < R * R; e += WG) {
        int i = e / R, j = e % R;
        double s = 0.0;
        for (int k = 0; k < R; ++k) s += Q[i * R + k] * wv[k] * Q[j * R + k];
        fac[e] = s;
    }
    if (tid == 0) *status = 1;
}

// --------------------------------------------------------- pseudo-inverse

// In-place Moore-Penrose pseudo-inverse of a general n x n row-major matrix
// (pseudoinverse_gpu, src/cpd_gpu.cu:452-532, which takes an SVD by gesvdj):
// one-sided Jacobi SVD (Hestenes) in LDS.  W = A is rotated column pair by
// column pair, the rotations accumulated in V, until every pair of columns
// of W is orthogonal to working precision; then sigma_j = ||w_j||,
// A = U diag(sigma) V^T with u_j = w_j / sigma_j, and
// A^+ = V diag(1 / sigma^2) W^T over the sigma_j above the reference's
// cutoff n * (nextafter(sigma_max, +inf) - sigma_max).  A sweep visits all
// pairs as n' - 1 rounds of n' / 2 disjoint pairs (round-robin tournament,
// n' = n rounded up to even), one thread per pair.  *info = 0 when a sweep
// made no rotation, else the sweep count (gesvdj's info > 0: not converged).
__global__ __launch_bounds__(WG) void k_pinv_general(double* __restrict__ A, int n,
                                                     int* __restrict__ info)
{
    extern __shared__ double s_w[];  // W (n x n), then V (n x n), then sigma (n)
    double* s_v = s_w + n * n;
    double* s_sig = s_v + n * n;
    __shared__ int s_rot;
    const int tid = threadIdx.x;
    for (int e = tid; e < n * n; e += WG) {
        s_w[e] = A[e];
        s_v[e] = (e / n == e % n) ? 1.0 : 0.0;
    }
    const int np = (n + 1) & ~1, m = np - 1;
    int sweeps = 0;
    bool converged = false;
    for (; sweeps < 60 && !converged; ++sweeps) {
        if (tid == 0) s_rot = 0;
        __syncthreads();
        for (int r = 0; r < m; ++r) {
            if (tid < np / 2) {
                int p = tid == 0 ? r : (r + tid) % m;
                int q = tid == 0 ? m : (r - tid + m) % m;
                if (p > q) { const int t = p; p = q; q = t; }
                if (q < n) {
                    double a = 0.0, b = 0.0, g = 0.0;
                    for (int i = 0; i < n; ++i) {
                        const double x = s_w[i * n + p], y = s_w[i * n + q];
                        a += x * x;
                        b += y * y;
                        g += x * y;
                    }
                    if (g != 0.0 && fabs(g) > 1e-15 * sqrt(a * b)) {
                        const double z = (b - a) / (2.0 * g);
                        const double t = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
                        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                        for (int i = 0; i < n; ++i) {
                            const double x = s_w[i * n + p], y = s_w[i * n + q];
                            s_w[i * n + p] = c * x - s * y;
                            s_w[i * n + q] = s * x + c * y;
                            const double vx = s_v[i * n + p], vy = s_v[i * n + q];
                            s_v[i * n + p] = c * vx - s * vy;
                            s_v[i * n + q] = s * vx + c * vy;
                        }
                        s_rot = 1;
                    }
                }
            }
            __syncthreads();
        }
        converged = s_rot == 0;  // every thread reads it after the round's barrier
        __syncthreads();
    }
    for (int j = tid; j < n; j += WG) {
        double a = 0.0;
        for (int i = 0; i < n; ++i) a += s_w[i * n + j] * s_w[i * n + j];
        s_sig[j] = sqrt(a);
    }
    __syncthreads();
    double smax = 0.0;
    for (int j = 0; j < n; ++j) smax = fmax(smax, s_sig[j]);
    const double tol = n * (nextafter(smax, INFINITY) - smax);
    for (int e = tid; e < n * n; e += WG) {
        const int i = e / n, k = e % n;
        double t = 0.0;
        for (int j = 0; j < n; ++j) {
            const double sg = s_sig[j];
            if (sg > tol) t += s_v[i * n + j] * s_w[k * n + j] / (sg * sg);
        }
        A[e] = t;
    }
    if (tid == 0 && info) *info = converged ? 0 : sweeps;
}

// ------------------------------------------------------------ solve rows

// U = M F with F = V^{-1} (or V^+), row by row, written as the factor dtype,
// and per-workgroup column statistics (which 0: sum of squares, 1: max
// seeded with 1) into stats[blockIdx][R].  A group of G lanes (G = R rounded
// up to a power of two) takes one row: lane j holds row j of F in registers
// and forms u_j = sum_k F[j][k] m_k, with the row of M staged in LDS (a
// coalesced load).  The workgroup walks rows with a grid stride.
template <typename MT, typename UT, int G>
__global__ __launch_bounds__(WG) void k_solve_rows(const MT* __restrict__ M, UT* __restrict__ U,
                                                   uint64_t rows, int R,
                                                   const double* __restrict__ fac, int which,
                                                   double* __restrict__ stats)
{
    constexpr int SLOTS = WG / G;
    constexpr int RPS = 4;  // rows per slot per step: four loads in flight per lane
    __shared__ double s_m[SLOTS * RPS][G + 1];
    __shared__ double s_st[SLOTS][G];
    const int j = threadIdx.x % G, slot = threadIdx.x / G;
    const bool col = j < R;
    double f[G];
    #pragma unroll
    for (int k = 0; k < G; ++k) f[k] = (col && k < R) ? fac[j * R + k] : 0.0;
    double st = which == 0 ? 0.0 : 1.0;
    const uint64_t step = (uint64_t)gridDim.x * SLOTS * RPS;
    // every lane runs the same number of iterations: the barriers below are uniform
    const uint64_t iters = (rows + step - 1) / step;
    for (uint64_t it = 0; it < iters; ++it) {
        // rows base + q * SLOTS + slot: each q a run of SLOTS consecutive rows
        const uint64_t base = it * step + (uint64_t)blockIdx.x * SLOTS * RPS;
        #pragma unroll
        for (int q = 0; q < RPS; ++q) {
            const uint64_t row = base + (uint64_t)q * SLOTS + slot;
            s_m[q * SLOTS + slot][j] = (row < rows && col) ? (double)M[row * (uint64_t)R + j] : 0.0;
        }
        __syncthreads();
        #pragma unroll
        for (int q = 0; q < RPS; ++q) {
            const uint64_t row = base + (uint64_t)q * SLOTS + slot;
            double t = 0.0;
            #pragma unroll
            for (int k = 0; k < G; ++k) t += f[k] * s_m[q * SLOTS + slot][k];
            if (row < rows && col) {
                U[row * (uint64_t)R + j] = from_f64<UT>(t);
                st = which == 0 ? st + t * t : fmax(st, t);
            }
        }
        __syncthreads();
    }
    s_st[slot][j] = st;
    __syncthreads();
    if (threadIdx.x < G && col) {
        double x = s_st[0][j];
        for (int q = 1; q < SLOTS; ++q) x = which == 0 ? x + s_st[q][j] : fmax(x, s_st[q][j]);
        stats[(uint64_t)blockIdx.x * R + j] = x;
    }
}

// ------------------------------------------------------- fused solve + Gram

// One pass of a mode's ALS update on MFMA: U = M diag(c) F^T row tile by row
// tile (16 rows), then, from the same registers, the tile's contribution to
// G = U^T U, the column statistics of U and (with ip) sum_i U_ic M_ic.
//   solve  v_mfma_f64_16x16x4f64 with A[i][k] = M[r0 + i][4 ks + k] (lane l
//          loads M[r0 + l%16][4 ks + l/16]) and B[k][j] = c_k F[16 ct + j][k]
//          (held in registers for the whole launch); the result D[i][j] sits
//          in register v of lane l as U[r0 + l/16 + 4 v][16 ct + l%16].
//   Gram   that register v is, for the row step r0 + 4 v, exactly the A and B
//          operand of the Gram's MFMA (lane l: U[r0 + 4v + l/16][16 t + l%16]),
//          so the stored (rounded) values feed the Gram tiles directly: f64
//          MFMA for f64 factors, f32 for f32 / bf16 ones, as in k_gram.
// Rows past the end load as 0 and give 0 rows of U (not stored).  Each
// workgroup's waves take tiles with a fixed stride; its partial Gram (and
// ip) and column statistics are combined in LDS in wave order and written to
// part[blockIdx] (R*R + R) and stats[blockIdx] (R): deterministic.
template <typename MT, typename UT, int T>
__global__ __launch_bounds__(WG) void k_solve_gram(const MT* __restrict__ M, UT* __restrict__ U, uint64_t rows,
                                                   int R, const double* __restrict__ fac,
                                                   const double* __restrict__ colscale, int which, bool want_ip,
                                                   double* __restrict__ stats, double* __restrict__ part)
{
    constexpr bool F64 = std::is_same<UT, double>::value;
    using AT = typename std::conditional<F64, double, float>::type;
    using VT = typename std::conditional<F64, d4, f4>::type;
    constexpr int NT = T * (T + 1) / 2;
    constexpr int KS = 4 * T;  // k steps of 4 over the (padded) rank 16 T
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i16 = lane & 15, k4 = lane >> 4;
    // B operands: c_k F[16 ct + j][k] for k = 4 ks + k4, j = i16
    double b[T][KS];
    #pragma unroll
    for (int ct = 0; ct < T; ++ct)
        #pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + k4, j = 16 * ct + i16;
            b[ct][ks] = (k < R && j < R) ? fac[j * R + k] * (colscale ? colscale[k] : 1.0) : 0.0;
        }
    VT g[NT];
    #pragma unroll
    for (int q = 0; q < NT; ++q) g[q] = VT{0, 0, 0, 0};
    double st[T], ipa[T];
    #pragma unroll
    for (int ct = 0; ct < T; ++ct) {
        st[ct] = which == 0 ? 0.0 : 1.0;
        ipa[ct] = 0.0;
    }
    const uint64_t tiles = (rows + 15) / 16;
    const uint64_t stride = (uint64_t)gridDim.x * WAVES_PER_WG;
    for (uint64_t tile = (uint64_t)blockIdx.x * WAVES_PER_WG + w; tile < tiles; tile += stride) {
        const uint64_t r0 = tile * 16;
        d4 d[T];
        #pragma unroll
        for (int ct = 0; ct < T; ++ct) d[ct] = d4{0, 0, 0, 0};
        const uint64_t ra = r0 + i16;
        #pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + k4;
            const double a = (ra < rows && k < R) ? (double)M[ra * (uint64_t)R + k] : 0.0;
            #pragma unroll
            for (int ct = 0; ct < T; ++ct) d[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[ct][ks], d[ct], 0, 0, 0);
        }
        #pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint64_t row = r0 + k4 + 4 * v;
            AT x[T];
            #pragma unroll
            for (int ct = 0; ct < T; ++ct) {
                const int col = 16 * ct + i16;
                const double t = d[ct][v];
                const bool ok = row < rows && col < R;
                const UT u = from_f64<UT>(t);
                if (ok) {
                    U[row * (uint64_t)R + col] = u;
                    st[ct] = which == 0 ? st[ct] + t * t : fmax(st[ct], t);
                    if (want_ip) {
                        const double mv = (double)M[row * (uint64_t)R + col];
                        ipa[ct] += (double)(AT)as_f64(u) * mv;
                    }
                }
                x[ct] = ok ? (AT)as_f64(u) : (AT)0;
            }
            int q = 0;
            #pragma unroll
            for (int ti = 0; ti < T; ++ti)
                #pragma unroll
                for (int tj = ti; tj < T; ++tj) {
                    if constexpr (F64)
                        g[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[ti], x[tj], g[q], 0, 0, 0);
                    else
                        g[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[ti], x[tj], g[q], 0, 0, 0);
                    ++q;
                }
        }
    }
    // combine the waves: Gram tiles as in k_gram; statistics and ip over the
    // four row groups (k4) of a wave, then over the waves, in order
    __shared__ double s_g[WAVES_PER_WG][T * 16][T * 16 + 1];
    __shared__ double s_v[WAVES_PER_WG][2][T * 16];
    int q = 0;
    #pragma unroll
    for (int ti = 0; ti < T; ++ti)
        #pragma unroll
        for (int tj = ti; tj < T; ++tj) {
            #pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = ti * 16 + (F64 ? k4 + 4 * v : 4 * k4 + v), j = tj * 16 + i16;
                const double x = (double)g[q][v];
                s_g[w][i][j] = x;
                s_g[w][j][i] = x;
            }
            ++q;
        }
    #pragma unroll
    for (int ct = 0; ct < T; ++ct) {
        double a = st[ct], c = ipa[ct];
        #pragma unroll
        for (int o = 16; o < 64; o <<= 1) {
            const double y = __shfl_xor(a, o);
            a = which == 0 ? a + y : fmax(a, y);
            c += __shfl_xor(c, o);
        }
        if (k4 == 0) {
            s_v[w][0][ct * 16 + i16] = a;
            s_v[w][1][ct * 16 + i16] = c;
        }
    }
    __syncthreads();
    double* out = part + (uint64_t)blockIdx.x * (uint64_t)(R * R + R);
    for (int e = threadIdx.x; e < R * R; e += WG) {
        const int i = e / R, j = e % R;
        out[e] = ((s_g[0][i][j] + s_g[1][i][j]) + s_g[2][i][j]) + s_g[3][i][j];
    }
    for (int e = threadIdx.x; e < R; e += WG) {
        out[R * R + e] = ((s_v[0][1][e] + s_v[1][1][e]) + s_v[2][1][e]) + s_v[3][1][e];
        double a = s_v[0][0][e];
        for (int x = 1; x < WAVES_PER_WG; ++x) a = which == 0 ? a + s_v[x][0][e] : fmax(a, s_v[x][0][e]);
        stats[(uint64_t)blockIdx.x * R + e] = a;
    }
}

// lambda and 1/lambda (0 for a zero column) from the P2 folded rows of the
// column statistics; dinv (may be null) receives 1/lambda too.
__global__ __launch_bounds__(WAVE) void k_lambda_finish(const double* __restrict__ stats, int p2, int R,
                                                        int which, double* __restrict__ lambda,
                                                        double* __restrict__ inv,
                                                        double* __restrict__ dinv)
{
    const int c = blockIdx.x * WAVE + threadIdx.x;
    if (c >= R) return;
    double y = stats[c];
    for (int q = 1; q < p2; ++q) y = fold_op(which, y, stats[(uint64_t)q * R + c]);
    if (which == 0) y = sqrt(y);
    const double iv = (y == 0.0) ? 0.0 : 1.0 / y;
    lambda[c] = y;
    inv[c] = iv;
    if (dinv) dinv[c] = iv;
}

// ------------------------------------------------------------------- fit

__global__ __launch_bounds__(WG) void k_fit(GramSet gs, int nmodes, int R,
                                            const double* __restrict__ lambda,
                                            const double* __restrict__ ip, double normX2,
                                            double* __restrict__ out)
{
    __shared__ double s_off[WG], s_diag[WG];
    double off = 0.0, diag = 0.0;
    for (int e = threadIdx.x; e < R * R; e += WG) {
        int i = e / R, j = e % R;
        if (j > i) continue;
        double t = lambda[i] * lambda[j];
        for (int m = 0; m < nmodes; ++m) t *= gs.g[m][e];
        if (j < i) off += 2.0 * t;
        else diag += t;
    }
    s_off[threadIdx.x] = off;
    s_diag[threadIdx.x] = diag;
    __syncthreads();
    if (threadIdx.x == 0) {
        double so = 0.0, sd = 0.0, inner = 0.0;
        for (int t = 0; t < WG; ++t) { so += s_off[t]; sd += s_diag[t]; }
        for (int r = 0; r < R; ++r) inner += ip[r] * lambda[r];
        double normU = fabs(so + sd);
        double res = normX2 + normU - 2.0 * inner;
        if (res > 0.0) res = sqrt(res);
        out[0] = 1.0 - res / sqrt(normX2);
        out[1] = inner;
        out[2] = normU;
    }
}

// Sum of squares of n values (f64 or f32 storage) per workgroup.
template <typename T>
__global__ __launch_bounds__(WG) void k_sumsq(const T* __restrict__ x, uint64_t n,
                                              double* __restrict__ part)
{
    __shared__ double s[WAVES_PER_WG];
    double acc = 0.0;
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < n; e += (uint64_t)gridDim.x * WG) {
        double v = (double)x[e];
        acc += v * v;
    }
    #pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

__global__ void k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        out[0] = s;
    }
}

__global__ __launch_bounds__(WG) void k_hadamard(double* __restrict__ x, const double* __restrict__ y,
                                                 uint64_t n)
{
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < n; e += (uint64_t)gridDim.x * WG)
        x[e] *= y[e];
}

// Per-workgroup column statistics of an existing factor (which 0: sum of
// squares, 1: max seeded with 1), for the stand-alone column norm.
template <typename UT>
__global__ __launch_bounds__(WG) void k_colstats(const UT* __restrict__ U, uint64_t rows, int R,
                                                 uint64_t rows_per_wg, int which,
                                                 double* __restrict__ stats)
{
    const uint64_t rb = (uint64_t)blockIdx.x * rows_per_wg;
    const uint64_t re = min(rb + rows_per_wg, rows);
    for (int c = threadIdx.x; c < R; c += WG) {
        double x = (which == 0) ? 0.0 : 1.0;
        for (uint64_t i = rb; i < re; ++i) {
            double v = as_f64(U[i * (uint64_t)R + c]);
            x = (which == 0) ? x + v * v : fmax(x, v);
        }
        stats[(uint64_t)blockIdx.x * R + c] = x;
    }
}

template <typename UT>
__global__ __launch_bounds__(WG) void k_scale(UT* __restrict__ U, uint64_t cells, int R,
                                              const double* __restrict__ inv)
{
    for (uint64_t e = (uint64_t)blockIdx.x * WG + threadIdx.x; e < cells; e += (uint64_t)gridDim.x * WG)
        U[e] = from_f64<UT>(as_f64(U[e]) * inv[e % R]);
}

__global__ void k_vecsum(double* __restrict__ x, uint64_t n)
{
    __shared__ double s[WG];
    double acc = 0.0;
    for (uint64_t e = threadIdx.x; e < n; e += WG) acc += x[e];
    s[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < WG; ++i) t += s[i];
        x[0] = t;
    }
}

// ------------------------------------------------------------ launchers

template <typename UT>
static int gram_t(const UT* U, uint64_t rows, int R, const void* M, double* part,
                  int nparts, uint64_t rpw, hipStream_t st)
{
    int T = (R + 15) / 16;
    switch (T) {
        case 1: hipLaunchKernelGGL((k_gram<UT, 1>), dim3(nparts), dim3(WG), 0, st, U, rows, R, rpw, M, part); break;
        case 2: hipLaunchKernelGGL((k_gram<UT, 2>), dim3(nparts), dim3(WG), 0, st, U, rows, R, rpw, M, part); break;
        case 3: hipLaunchKernelGGL((k_gram<UT, 3>), dim3(nparts), dim3(WG), 0, st, U, rows, R, rpw, M, part); break;
        case 4: hipLaunchKernelGGL((k_gram<UT, 4>), dim3(nparts), dim3(WG), 0, st, U, rows, R, rpw, M, part); break;
        default: return fail(BLCO_EUNSUPPORTED, "rank %d > 64 not supported by the Gram kernel", R);
    }
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int gram_parts(uint64_t rows, int* nparts, uint64_t* rows_per_wg)
{
    uint64_t steps = (rows + 3) / 4;
    uint64_t want = (steps + 63) / 64;  // ~64 row-steps per workgroup (16 per wave)
    if (want < 1) want = 1;
    if (want > GRAM_MAX_PARTS) want = GRAM_MAX_PARTS;  // callers size partial buffers for this
    uint64_t rpw = (rows + want - 1) / want;
    rpw = (rpw + 3) / 4 * 4;
    if (rpw == 0) rpw = 4;
    *rows_per_wg = rpw;
    *nparts = (int)((rows + rpw - 1) / rpw);
    if (*nparts < 1) *nparts = 1;
    return BLCO_OK;
}

int launch_gram(const void* U, int dtype, uint64_t rows, int R, const void* M,
                double* part, double* G, double* ip, hipStream_t st)
{
    if (R < 1 || R > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64] for the Gram kernel", R);
    int nparts;
    uint64_t rpw;
    gram_parts(rows, &nparts, &rpw);
    int rc;
    if (dtype == BLCO_F64) rc = gram_t<double>((const double*)U, rows, R, M, part, nparts, rpw, st);
    else if (dtype == BLCO_F32) rc = gram_t<float>((const float*)U, rows, R, M, part, nparts, rpw, st);
    else rc = gram_t<bf16>((const bf16*)U, rows, R, M, part, nparts, rpw, st);
    if (rc) return rc;
    const int W = R * R + R;
    int p2 = 0;
    rc = fold_parts(part, nparts, W, 0, &p2, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gram_finish, dim3((W + WAVE - 1) / WAVE), dim3(WAVE), 0, st, part, p2, R, nullptr, nullptr,
                       G, ip);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

double gram_eps(int factor_dtype)
{
    return factor_dtype == BLCO_F64 ? 2.220446049250313e-16 : 1.1920928955078125e-07;
}

int launch_factorize(const GramSet& gs, int nmodes, int mode, int R, double eps, double* fac,
                     double* work, int* status, hipStream_t st, const double* dinv, double* colscale)
{
    if (R < 1 || R > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64] in the factorisation", R);
    size_t lds = sizeof(double) * 2 * (size_t)R * R;
    if (lds > 64 * 1024) HIP_TRY(hipFuncSetAttribute((const void*)k_factorize,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_factorize, dim3(1), dim3(WG), lds, st, gs, nmodes, mode, R, eps, fac, work, status,
                       dinv, colscale);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

template <typename MT, typename UT>
static int solve_rows_t(const MT* M, UT* U, uint64_t rows, int R, const double* fac,
                        int which, double* stats, hipStream_t st)
{
    if (rows == 0) return BLCO_OK;
    const uint32_t grid = solve_parts(rows);
    if (R > 64) return fail(BLCO_EUNSUPPORTED, "rank %d > 64 in the row solve", R);
    const int g = R <= 8 ? 8 : R <= 16 ? 16 : R <= 32 ? 32 : 64;
    switch (g) {
        case 8: hipLaunchKernelGGL((k_solve_rows<MT, UT, 8>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, which, stats); break;
        case 16: hipLaunchKernelGGL((k_solve_rows<MT, UT, 16>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, which, stats); break;
        case 32: hipLaunchKernelGGL((k_solve_rows<MT, UT, 32>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, which, stats); break;
        default: hipLaunchKernelGGL((k_solve_rows<MT, UT, 64>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, which, stats); break;
    }
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_solve_rows(const void* M, void* U, int dtype, uint64_t rows, int R, const double* fac,
                      int which, double* stats, hipStream_t st)
{
    if (dtype == BLCO_F64) return solve_rows_t<double, double>((const double*)M, (double*)U, rows, R, fac, which, stats, st);
    if (dtype == BLCO_F32) return solve_rows_t<float, float>((const float*)M, (float*)U, rows, R, fac, which, stats, st);
    return solve_rows_t<float, bf16>((const float*)M, (bf16*)U, rows, R, fac, which, stats, st);
}

// Workgroups of the row solve (each walks rows with a grid stride).
uint32_t solve_parts(uint64_t rows)
{
    const uint64_t p = (rows + 127) / 128;
    return (uint32_t)(p < 1 ? 1 : p > 2048 ? 2048 : p);
}

uint32_t solve_gram_parts(uint64_t rows)
{
    const uint64_t p = (rows + 255) / 256;  // at least 4 row tiles per wave
    return (uint32_t)(p < 1 ? 1 : p > GRAM_MAX_PARTS ? GRAM_MAX_PARTS : p);
}

template <typename MT, typename UT>
static int solve_gram_t(const MT* M, UT* U, uint64_t rows, int R, const double* fac, const double* cs, int which,
                        bool ip, double* stats, double* part, uint32_t grid, hipStream_t st)
{
    switch ((R + 15) / 16) {
        case 1: hipLaunchKernelGGL((k_solve_gram<MT, UT, 1>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, cs, which, ip, stats, part); break;
        case 2: hipLaunchKernelGGL((k_solve_gram<MT, UT, 2>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, cs, which, ip, stats, part); break;
        case 3: hipLaunchKernelGGL((k_solve_gram<MT, UT, 3>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, cs, which, ip, stats, part); break;
        case 4: hipLaunchKernelGGL((k_solve_gram<MT, UT, 4>), dim3(grid), dim3(WG), 0, st, M, U, rows, R, fac, cs, which, ip, stats, part); break;
        default: return fail(BLCO_EUNSUPPORTED, "rank %d > 64 in the fused solve", R);
    }
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_solve_gram(const void* M, void* U, int dtype, uint64_t rows, int R, const double* fac,
                      const double* colscale, int which, bool want_ip, double* stats, double* part, double* lambda,
                      double* inv, double* dinv, double* G, double* ip, hipStream_t st)
{
    if (R < 1 || R > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64] in the fused solve", R);
    if (rows == 0) return fail(BLCO_EINVAL, "fused solve of 0 rows");
    const uint32_t grid = solve_gram_parts(rows);
    int rc;
    if (dtype == BLCO_F64)
        rc = solve_gram_t<double, double>((const double*)M, (double*)U, rows, R, fac, colscale, which, want_ip, stats,
                                          part, grid, st);
    else if (dtype == BLCO_F32)
        rc = solve_gram_t<float, float>((const float*)M, (float*)U, rows, R, fac, colscale, which, want_ip, stats, part,
                                        grid, st);
    else
        rc = solve_gram_t<float, bf16>((const float*)M, (bf16*)U, rows, R, fac, colscale, which, want_ip, stats, part,
                                       grid, st);
    if (!rc) rc = launch_lambda(stats, (int)grid, R, which, lambda, inv, st, dinv);
    if (rc) return rc;
    const int W = R * R + R;
    int p2 = 0;
    rc = fold_parts(part, (int)grid, W, 0, &p2, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gram_finish, dim3((W + WAVE - 1) / WAVE), dim3(WAVE), 0, st, part, p2, R, inv, colscale, G,
                       want_ip ? ip : nullptr);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_pinv_general(double* A, int n, int* info, hipStream_t st)
{
    if (n < 1 || n > 64) return fail(BLCO_EUNSUPPORTED, "pseudo-inverse of order %d outside [1, 64]", n);
    const size_t lds = sizeof(double) * ((size_t)2 * n * n + n);
    hipLaunchKernelGGL(k_pinv_general, dim3(1), dim3(WG), lds, st, A, n, info);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_lambda(double* stats, int nparts, int R, int which, double* lambda, double* inv, hipStream_t st,
                  double* dinv)
{
    int p2 = 0;
    int rc = fold_parts(stats, nparts, R, which, &p2, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lambda_finish, dim3((R + WAVE - 1) / WAVE), dim3(WAVE), 0, st, stats, p2, R, which, lambda,
                       inv, dinv);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_fit(const GramSet& gs, int nmodes, int R, const double* lambda, const double* ip,
               double normX2, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_fit, dim3(1), dim3(WG), 0, st, gs, nmodes, R, lambda, ip, normX2, out);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_sumsq(const void* x, int dtype, uint64_t n, double* part, int nparts, double* out,
                 hipStream_t st)
{
    if (dtype == BLCO_F64)
        hipLaunchKernelGGL(k_sumsq<double>, dim3(nparts), dim3(WG), 0, st, (const double*)x, n, part);
    else
        hipLaunchKernelGGL(k_sumsq<float>, dim3(nparts), dim3(WG), 0, st, (const float*)x, n, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, part, nparts, out);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_hadamard(double* x, const double* y, uint64_t n, hipStream_t st)
{
    uint32_t grid = (uint32_t)std::min<uint64_t>((n + WG - 1) / WG, 4096);
    if (grid == 0) return BLCO_OK;
    hipLaunchKernelGGL(k_hadamard, dim3(grid), dim3(WG), 0, st, x, y, n);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

}  // namespace blco

namespace blco {

template <typename UT>
static void colstats_t(const UT* U, uint64_t rows, int R, uint64_t rpw, int nparts, int which,
                       double* stats, hipStream_t st)
{
    hipLaunchKernelGGL(k_colstats<UT>, dim3(nparts), dim3(WG), 0, st, U, rows, R, rpw, which, stats);
}

int launch_colstats(const void* U, int dtype, uint64_t rows, int R, int which, double* stats,
                    int nparts, hipStream_t st)
{
    uint64_t rpw = (rows + nparts - 1) / nparts;
    if (rpw == 0) rpw = 1;
    if (dtype == BLCO_F64) colstats_t((const double*)U, rows, R, rpw, nparts, which, stats, st);
    else if (dtype == BLCO_F32) colstats_t((const float*)U, rows, R, rpw, nparts, which, stats, st);
    else colstats_t((const bf16*)U, rows, R, rpw, nparts, which, stats, st);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_scale(void* U, int dtype, uint64_t rows, int R, const double* inv, hipStream_t st)
{
    uint64_t cells = rows * (uint64_t)R;
    uint32_t grid = (uint32_t)std::min<uint64_t>((cells + WG - 1) / WG, 8192);
    if (grid == 0) return BLCO_OK;
    if (dtype == BLCO_F64) hipLaunchKernelGGL(k_scale<double>, dim3(grid), dim3(WG), 0, st, (double*)U, cells, R, inv);
    else if (dtype == BLCO_F32) hipLaunchKernelGGL(k_scale<float>, dim3(grid), dim3(WG), 0, st, (float*)U, cells, R, inv);
    else hipLaunchKernelGGL(k_scale<bf16>, dim3(grid), dim3(WG), 0, st, (bf16*)U, cells, R, inv);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

int launch_vecsum(double* x, uint64_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_vecsum, dim3(1), dim3(WG), 0, st, x, n);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

}  // namespace blco
