// cpals.hip -- device-resident CP-ALS driver (cpals_blco_dev,
// src/cpd_gpu.cu:21-242) with the CPU reference's iteration semantics
// (cpd_alto, include/cpd.hpp:44-187):
//
//   for it: for each mode n (0..N-1, or longest first):
//       M   = MTTKRP(n)                         (kernel per mode)
//       F   = (prod_{m != n} G_m)^{-1}, or the pseudo-inverse
//       U_n = M F, column statistics
//       lambda = L2 norms at it 0, max(1, max) after; U_n /= lambda
//       G_n = U_n^T U_n (MFMA); for the last mode also sum_i U_n M
//   fit = 1 - ||X - [[lambda; U]]|| / ||X||, stop when it > 0 and
//   |fit - fit_old| < tolerance.
//
// The division by lambda is not written back: each factor is stored as the
// solve leaves it, U_n diag(lambda_n), with 1 / lambda_n kept aside (dinv).
// Its Gram is formed from the stored factor and scaled by 1 / lambda on
// both sides; the MTTKRP of a later mode then carries the product c of the
// other factors' lambdas in its columns, and that mode's solve multiplies
// M by c^{-1} (folded into F).  Mathematically the same step; each mode's
// update is one MFMA kernel that reads M once and writes U_n once, forming
// the Gram from the solve's registers (launch_solve_gram), instead of a
// solve, an in-place rescale and a Gram pass.  blco_cpals_get applies
// 1 / lambda.
//
// Everything stays on one stream with no host round trip inside a sweep;
// the host reads back the fit (8 bytes) once per iteration.  With use_graph
// one sweep (after iteration 0, whose norm differs) is captured into a
// hipGraph and replayed; the output buffer is cleared by a fill kernel, never
// by a memset node.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "als.hpp"
#include "cpals.hpp"
#include "device.hpp"
#include "devcommon.hpp"
#include "mttkrp.hpp"

using namespace blco;

namespace blco {

void to_dtype(const double* src, void* dst, uint64_t n, int dt)
{
    if (dt == BLCO_F64) std::memcpy(dst, src, n * 8);
    else if (dt == BLCO_F32) {
        float* f = (float*)dst;
        for (uint64_t i = 0; i < n; ++i) f[i] = (float)src[i];
    } else {
        uint16_t* b = (uint16_t*)dst;
        for (uint64_t i = 0; i < n; ++i) {
            float x = (float)src[i];
            uint32_t u;
            std::memcpy(&u, &x, 4);
            if ((u & 0x7f800000u) == 0x7f800000u) b[i] = (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
            else b[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        }
    }
}

void from_dtype(const void* src, double* dst, uint64_t n, int dt)
{
    if (dt == BLCO_F64) std::memcpy(dst, src, n * 8);
    else if (dt == BLCO_F32) {
        const float* f = (const float*)src;
        for (uint64_t i = 0; i < n; ++i) dst[i] = f[i];
    } else {
        const uint16_t* b = (const uint16_t*)src;
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t u = (uint32_t)b[i] << 16;
            float x;
            std::memcpy(&x, &u, 4);
            dst[i] = x;
        }
    }
}

int enqueue_factorize(blco_cpals_state* s, int n, hipStream_t st)
{
    GramSet gs;
    for (int m = 0; m < s->N; ++m) gs.g[m] = s->G[m];
    // the pivot floor of f64 unless the caller asked for the factors' precision
    const double eps = s->o.solve_floor ? gram_eps(s->ft) : gram_eps(BLCO_F64);
    return launch_factorize(gs, s->N, n, s->R, eps, s->fac, s->work, s->status, st, s->dinv, s->colscale);
}

int enqueue_step(blco_cpals_state* s, int n, int which_norm, hipStream_t st)
{
    int rc = enqueue_factorize(s, n, st);
    return rc ? rc : enqueue_update(s, n, which_norm, st);
}

int enqueue_update(blco_cpals_state* s, int n, int which_norm, hipStream_t st)
{
    const int N = s->N, R = s->R;
    const uint64_t I = s->d->dims[n];
    const int last = s->order[N - 1];
    return launch_solve_gram(s->Mbuf, s->U[n], s->ft, I, R, s->fac, s->colscale, which_norm, n == last, s->stats,
                             s->gram_part, s->lambda, s->inv, s->dinv + (size_t)n * R, s->G[n], s->ip, st);
}

int enqueue_fit(blco_cpals_state* s, hipStream_t st)
{
    GramSet gs;
    for (int m = 0; m < s->N; ++m) gs.g[m] = s->G[m];
    return launch_fit(gs, s->N, s->R, s->lambda, s->ip, s->normX2, s->fitbuf, st);
}

// Enqueues one ALS sweep (all modes) and the fit kernel.
static int enqueue_sweep(blco_cpals_state* s, int which_norm, bool timed)
{
    blco_dev* d = s->d;
    const int N = s->N, R = s->R;
    for (int k = 0; k < N; ++k) {
        const int n = s->order[k];
        // V_n's factorisation on the side stream, beside the MTTKRP
        HIP_TRY(hipEventRecord(s->fork[n], s->st));
        HIP_TRY(hipStreamWaitEvent(s->side, s->fork[n], 0));
        int rc = enqueue_factorize(s, n, s->side);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(s->join[n], s->side));
        rc = launch_fill(s->Mbuf, d->dims[n] * (uint64_t)R, s->mt, 0.0, s->st);
        if (rc) return rc;
        if (timed) HIP_TRY(hipEventRecord(s->ev0[n], s->st));
        rc = mttkrp(d, n, R, s->ft, s->U, s->Mbuf, s->o.kernel, s->o.nprtn, s->o.thread_cf, s->st,
                    &s->kern[n]);
        if (rc) return rc;
        if (timed) HIP_TRY(hipEventRecord(s->ev1[n], s->st));
        HIP_TRY(hipStreamWaitEvent(s->st, s->join[n], 0));
        rc = enqueue_update(s, n, which_norm, s->st);
        if (rc) return rc;
    }
    return enqueue_fit(s, s->st);
}

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace blco

extern "C" {

void blco_cpals_free(blco_cpals_state* s)
{
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->side) (void)hipStreamSynchronize(s->side);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (int m = 0; m < BLCO_MAX_MODES; ++m) {
        if (s->U[m]) (void)hipFree(s->U[m]);
        if (s->G[m]) (void)hipFree(s->G[m]);
        if (s->ev0[m]) (void)hipEventDestroy(s->ev0[m]);
        if (s->ev1[m]) (void)hipEventDestroy(s->ev1[m]);
        if (s->fork[m]) (void)hipEventDestroy(s->fork[m]);
        if (s->join[m]) (void)hipEventDestroy(s->join[m]);
    }
    if (s->side) (void)hipStreamDestroy(s->side);
    void* bufs[] = {s->Mbuf, s->lambda, s->inv, s->fac, s->work, s->status, s->gram_part,
                    s->stats, s->ip, s->fitbuf, s->norm_part, s->dinv, s->colscale};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (s->h_fit) (void)hipHostFree(s->h_fit);
    if (s->it0) (void)hipEventDestroy(s->it0);
    if (s->it1) (void)hipEventDestroy(s->it1);
    delete s;
}

int blco_cpals_create(blco_dev* d, const blco_cpals_opts* o, const double* const* U,
                      blco_cpals_state** out)
{
    if (!d || !o || !U || !out) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    const int N = d->nmodes, R = o->rank;
    if (R < 1 || R > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64]", R);
    if (N < 2) return fail(BLCO_EUNSUPPORTED, "CP-ALS needs at least 2 modes");
    if (o->precision != BLCO_F64 && o->precision != BLCO_F32 && o->precision != BLCO_BF16)
        return fail(BLCO_EINVAL, "precision %d", o->precision);
    const int need_v = o->precision == BLCO_F64 ? BLCO_F64 : BLCO_F32;
    if (d->vdtype != need_v)
        return fail(BLCO_EINVAL, "precision %d needs %s values on the device", o->precision,
                    need_v == BLCO_F64 ? "f64" : "f32");
    if (o->max_iters < 0) return fail(BLCO_EINVAL, "max_iters %d", o->max_iters);
    for (int m = 0; m < N; ++m)
        if (!U[m]) return fail(BLCO_EINVAL, "initial factor %d is null", m);
    blco_cpals_state* s = new (std::nothrow) blco_cpals_state();
    if (!s) return fail(BLCO_ENOMEM, "out of host memory");
    auto bad = [&](int rc) { blco_cpals_free(s); return rc; };
    s->d = d;
    s->o = *o;
    if (s->o.kernel == 0) s->o.kernel = BLCO_K_AUTO;
    s->N = N;
    s->R = R;
    s->ft = o->precision;
    s->mt = o->precision == BLCO_F64 ? BLCO_F64 : BLCO_F32;
    s->st = d->stream;
    // mode order: natural (CPU reference) or longest first (GPU reference)
    for (int m = 0; m < N; ++m) s->order[m] = m;
    if (o->mode_order == 1)
        for (int i = 1; i < N; ++i)
            for (int j = i; j > 0 && d->dims[s->order[j]] > d->dims[s->order[j - 1]]; --j)
                std::swap(s->order[j], s->order[j - 1]);
    const int fsz = dtype_size(s->ft), msz = s->mt == BLCO_F64 ? 8 : 4;
    uint64_t maxI = 0;
    for (int m = 0; m < N; ++m) maxI = std::max(maxI, d->dims[m]);
    std::vector<char> stage;
    for (int m = 0; m < N; ++m) {
        uint64_t n = d->dims[m] * (uint64_t)R;
        if (hipMalloc(&s->U[m], n * fsz) != hipSuccess || hipMalloc(&s->G[m], sizeof(double) * R * R) != hipSuccess)
            return bad(fail(BLCO_ENOMEM, "factor allocation failed"));
        stage.resize(n * fsz);
        to_dtype(U[m], stage.data(), n, s->ft);
        if (hipMemcpy(s->U[m], stage.data(), n * fsz, hipMemcpyHostToDevice) != hipSuccess)
            return bad(fail(BLCO_EHIP, "factor upload failed"));
        if (hipEventCreate(&s->ev0[m]) != hipSuccess || hipEventCreate(&s->ev1[m]) != hipSuccess ||
            hipEventCreateWithFlags(&s->fork[m], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->join[m], hipEventDisableTiming) != hipSuccess)
            return bad(fail(BLCO_EHIP, "event creation failed"));
    }
    if (hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(BLCO_EHIP, "stream creation failed"));
    const int gparts = GRAM_MAX_PARTS;  // upper bound of gram_parts() for any row count
    uint32_t sparts = solve_parts(maxI);
    const int nnorm = 1024;
    if (hipMalloc(&s->Mbuf, maxI * R * msz) != hipSuccess ||
        hipMalloc((void**)&s->lambda, sizeof(double) * R) != hipSuccess ||
        hipMalloc((void**)&s->inv, sizeof(double) * R) != hipSuccess ||
        hipMalloc((void**)&s->fac, sizeof(double) * R * R) != hipSuccess ||
        hipMalloc((void**)&s->work, sizeof(double) * (2 * R * R + R)) != hipSuccess ||
        hipMalloc((void**)&s->status, sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&s->gram_part, sizeof(double) * (size_t)gparts * (R * R + R)) != hipSuccess ||
        hipMalloc((void**)&s->stats, sizeof(double) * (size_t)(sparts ? sparts : 1) * R) != hipSuccess ||
        hipMalloc((void**)&s->ip, sizeof(double) * R) != hipSuccess ||
        hipMalloc((void**)&s->fitbuf, sizeof(double) * 4) != hipSuccess ||
        hipMalloc((void**)&s->norm_part, sizeof(double) * (nnorm + 1)) != hipSuccess ||
        hipMalloc((void**)&s->dinv, sizeof(double) * (size_t)N * R) != hipSuccess ||
        hipMalloc((void**)&s->colscale, sizeof(double) * R) != hipSuccess)
        return bad(fail(BLCO_ENOMEM, "solver work allocation failed"));
    if (hipHostMalloc((void**)&s->h_fit, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess)
        return bad(fail(BLCO_ENOMEM, "pinned allocation failed"));
    if (hipEventCreate(&s->it0) != hipSuccess || hipEventCreate(&s->it1) != hipSuccess)
        return bad(fail(BLCO_EHIP, "event creation failed"));
    // lambda = 1 (KruskalModelRandomInit), ||X||^2, initial Grams
    int rc = launch_fill(s->lambda, R, BLCO_F64, 1.0, s->st);
    if (!rc) rc = launch_fill(s->dinv, (uint64_t)N * R, BLCO_F64, 1.0, s->st);
    if (rc) return bad(rc);
    if (d->nnz && !d->stream_data) {
        rc = launch_sumsq(d->d_vals, d->vdtype, d->nnz, s->norm_part, nnorm, s->norm_part + nnorm, s->st);
        if (rc) return bad(rc);
        if (hipMemcpyAsync(&s->normX2, s->norm_part + nnorm, sizeof(double), hipMemcpyDeviceToHost, s->st) != hipSuccess)
            return bad(fail(BLCO_EHIP, "norm read-back failed"));
    } else if (d->nnz) {
        // streamed tensors: the values are in pinned host memory
        double acc = 0.0;
        if (d->vdtype == BLCO_F64) {
            const double* v = (const double*)d->h_vals;
            for (uint64_t i = 0; i < d->nnz; ++i) acc += v[i] * v[i];
        } else {
            const float* v = (const float*)d->h_vals;
            for (uint64_t i = 0; i < d->nnz; ++i) acc += (double)v[i] * (double)v[i];
        }
        s->normX2 = acc;
    }
    for (int m = 0; m < N && !rc; ++m)
        rc = launch_gram(s->U[m], s->ft, d->dims[m], R, nullptr, s->gram_part, s->G[m], nullptr, s->st);
    if (rc) return bad(rc);
    if (hipStreamSynchronize(s->st) != hipSuccess) return bad(fail(BLCO_EHIP, "setup failed"));
    // build the chunk table outside any capture
    if (!chunk_table(d, s->o.thread_cf > 0 ? s->o.thread_cf : 4)) return bad(BLCO_ENOMEM);
    *out = s;
    return BLCO_OK;
}

int blco_cpals_iterate(blco_cpals_state* s, int n_iters, double* fits, int* done)
{
    if (!s) return fail(BLCO_EINVAL, "null solver");
    if (done) *done = 0;
    int ran = 0;
    for (int i = 0; i < n_iters && !s->stopped; ++i) {
        const int it = s->iters_done;
        const int which = it == 0 ? 0 : 1;  // L2 at iteration 0, max norm after
        const bool use_graph = s->o.use_graph && it > 0;
        const double t0 = now_ms();
        if (it == 0) {
            // kernel 14's row-ordered copies of every mode, before any MTTKRP
            // of the sweep runs (their construction would share the GPU with it)
            for (int k = 0; k < s->N; ++k) {
                int rc = mttkrp_prepare(s->d, s->order[k], s->R, s->ft, s->o.kernel);
                if (rc) return rc;
            }
        }
        HIP_TRY(hipEventRecord(s->it0, s->st));
        if (use_graph) {
            if (!s->gexec) {
                HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_sweep(s, 1, false);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(s->st, &g);
                if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
                HIP_TRY(e);
                s->graph = g;
                HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
            }
            HIP_TRY(hipGraphLaunch(s->gexec, s->st));
        } else {
            int rc = enqueue_sweep(s, which, true);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(s->h_fit, s->fitbuf, sizeof(double) * 3, hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipEventRecord(s->it1, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->it0, s->it1));
        s->iter_ms += ms;
        if (!use_graph)
            for (int m = 0; m < s->N; ++m) {
                float mm = 0.f;
                HIP_TRY(hipEventElapsedTime(&mm, s->ev0[m], s->ev1[m]));
                s->mttkrp_ms[m] += mm;
            }
        s->iters_timed++;
        const double fit = s->h_fit[0];
        if (fits) fits[i] = fit;
        if (s->o.verbose)
            printf("--> Iter %3d: f = %0.6f f-delta = %0.6f time = %0.6f, (s)\n", it, fit,
                   fit - s->fit_old, (now_ms() - t0) / 1000.0);
        s->iters_done++;
        ++ran;
        // stop rule of cpd_alto (include/cpd.hpp:158-163); NaN also stops
        // (src/cpd_gpu.cu:195)
        if ((it > 0 && std::fabs(s->fit_old - fit) < s->o.tolerance) || std::isnan(fit)) s->stopped = true;
        s->fit_old = fit;
    }
    if (done) *done = ran;
    return BLCO_OK;
}

int blco_cpals_get(blco_cpals_state* s, double* const* U, double* lambda)
{
    if (!s) return fail(BLCO_EINVAL, "null solver");
    HIP_TRY(hipStreamSynchronize(s->st));
    const int fsz = dtype_size(s->ft), R = s->R;
    std::vector<char> stage;
    std::vector<double> dinv((size_t)s->N * R);
    HIP_TRY(hipMemcpy(dinv.data(), s->dinv, sizeof(double) * dinv.size(), hipMemcpyDeviceToHost));
    for (int m = 0; m < s->N; ++m) {
        if (!U || !U[m]) continue;
        uint64_t n = s->d->dims[m] * (uint64_t)R;
        stage.resize(n * fsz);
        HIP_TRY(hipMemcpy(stage.data(), s->U[m], n * fsz, hipMemcpyDeviceToHost));
        from_dtype(stage.data(), U[m], n, s->ft);
        // the stored factor carries lambda_m in its columns
        const double* iv = dinv.data() + (size_t)m * R;
        for (uint64_t e = 0; e < n; ++e) U[m][e] *= iv[e % R];
    }
    if (lambda) HIP_TRY(hipMemcpy(lambda, s->lambda, sizeof(double) * s->R, hipMemcpyDeviceToHost));
    return BLCO_OK;
}

int blco_cpals_times(const blco_cpals_state* s, double* mttkrp_ms, double* iter_ms, int* iters,
                     int* kernels)
{
    if (!s) return fail(BLCO_EINVAL, "null solver");
    for (int m = 0; m < s->N; ++m) {
        if (mttkrp_ms) mttkrp_ms[m] = s->mttkrp_ms[m];
        if (kernels) kernels[m] = s->kern[m];
    }
    if (iter_ms) *iter_ms = s->iter_ms;
    if (iters) *iters = s->iters_timed;
    return BLCO_OK;
}

void blco_cpals_reset_times(blco_cpals_state* s)
{
    if (!s) return;
    for (int m = 0; m < BLCO_MAX_MODES; ++m) s->mttkrp_ms[m] = 0.0;
    s->iter_ms = 0.0;
    s->iters_timed = 0;
}

int blco_cpals(blco_dev* d, const blco_cpals_opts* o, double* const* U, double* lambda,
               double* fit_trace, int* iters, double* final_fit)
{
    blco_cpals_state* s = nullptr;
    int rc = blco_cpals_create(d, o, (const double* const*)U, &s);
    if (rc) return rc;
    if (o->verbose)
        printf("Begin CP-ALS decomposition, rank=%d maxiter=%d tolerance=%f\n", o->rank, o->max_iters,
               o->tolerance);
    int done = 0;
    rc = blco_cpals_iterate(s, o->max_iters, fit_trace, &done);
    if (!rc) rc = blco_cpals_get(s, U, lambda);
    if (!rc) {
        if (iters) *iters = done;
        if (final_fit) *final_fit = s->fit_old;
        if (o->verbose) printf("Final fit: %0.9f\n", s->fit_old);
    }
    blco_cpals_free(s);
    return rc;
}

}  // extern "C"
