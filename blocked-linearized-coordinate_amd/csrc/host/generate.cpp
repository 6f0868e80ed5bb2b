// generate.cpp -- synthetic Poisson sparse tensors and CP initial factors.
//
// Index draws: uniform floor(u * I) or Zipf(s) by inversion of the exact CDF
// over ranks 1..I (rank k -> row k - 1).  Draw d of mode n uses counter d of
// stream STREAM_INDEX + n.  Duplicate coordinates are merged keeping the
// first draw; with topup more draws are made until target_nnz coordinates
// are unique and the first target_nnz unique draws (in draw order) are kept.
// Rates: rho = sum_r prod_n A_n(i_n, r), A_n entries Exp(1) = Gamma(1,1)
// (counter row * gen_rank + r of stream STREAM_GEN_FACTOR + n).  Values:
// zero-truncated Poisson(rho) from counters 64 * d + j of STREAM_VALUE.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "internal.hpp"
#include "radix.hpp"
#include "rng.hpp"

typedef unsigned __int128 u128;

namespace blco {

struct IndexSampler {
    uint64_t I = 1;
    double s = 0.0;
    std::vector<double> cdf;  // Zipf only

    void init(uint64_t len, double zs)
    {
        I = len;
        s = zs;
        if (s > 0.0) {
            cdf.resize(I);
            double acc = 0.0;
            for (uint64_t k = 0; k < I; ++k) {
                acc += std::pow((double)(k + 1), -s);
                cdf[k] = acc;
            }
            for (uint64_t k = 0; k < I; ++k) cdf[k] /= acc;
            cdf[I - 1] = 1.0;
        }
    }
    uint64_t draw(double u) const
    {
        if (s > 0.0) {
            uint64_t k = (uint64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            return k < I ? k : I - 1;
        }
        uint64_t k = (uint64_t)(u * (double)I);
        return k < I ? k : I - 1;
    }
};

}  // namespace blco

using namespace blco;

extern "C" int blco_generate(const blco_gen_spec* sp, blco_coo** out)
{
    if (!sp || !out) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    const int N = sp->nmodes;
    Layout L;
    int rc = make_layout(N, sp->dims, &L);
    if (rc) return rc;
    if (sp->gen_rank < 1) return fail(BLCO_EINVAL, "gen_rank must be >= 1");
    if (sp->target_nnz == 0) return fail(BLCO_EINVAL, "target_nnz must be >= 1");
    long double cells = 1.0L;
    for (int n = 0; n < N; ++n) cells *= (long double)sp->dims[n];
    if ((long double)sp->target_nnz > cells)
        return fail(BLCO_EINVAL, "target_nnz exceeds the number of cells");

    IndexSampler smp[BLCO_MAX_MODES];
    for (int n = 0; n < N; ++n) smp[n].init(sp->dims[n], sp->zipf_s[n]);

    // Draw, merge duplicates (stable sort by linear key keeps the first
    // draw), and top up until enough unique coordinates exist.
    std::vector<uint64_t> draws;  // draw-major: draws[d * N + n]
    std::vector<uint8_t> keep;
    uint64_t ndraw = 0, unique = 0;
    uint64_t want = sp->target_nnz;
    int total_bits = L.total_bits;
    // Each round's size follows the unique yield of the draws so far, so a
    // skewed (Zipf) tensor with many duplicates converges in a few rounds.
    // The result does not depend on the round sizes: draws are a prefix of
    // one counter sequence and the first unique draws in draw order are kept.
    double yield = 1.0;
    for (int round = 0; round < 64; ++round) {
        uint64_t add = want;
        if (round > 0) {
            yield = std::max(0.01, (double)unique / (double)ndraw);
            add = (uint64_t)((double)(want - unique) / yield * 1.15) + 1024;
        }
        uint64_t old = ndraw;
        ndraw += add;
        try {
            draws.resize(ndraw * N);
        } catch (const std::bad_alloc&) {
            return fail(BLCO_ENOMEM, "generator allocation failed");
        }
        #pragma omp parallel for schedule(static)
        for (uint64_t d = old; d < ndraw; ++d)
            for (int n = 0; n < N; ++n)
                draws[d * N + n] = smp[n].draw(rng_u01(sp->seed, STREAM_INDEX + n, d));
        // linear key with per-mode bit fields (bits_n), which is injective
        std::vector<u128> key(ndraw);
        std::vector<uint64_t> id(ndraw);
        #pragma omp parallel for schedule(static)
        for (uint64_t d = 0; d < ndraw; ++d) {
            u128 k = 0;
            int sh = 0;
            for (int n = 0; n < N; ++n) {
                k |= (u128)draws[d * N + n] << sh;
                sh += L.bits[n];
            }
            key[d] = k;
            id[d] = d;
        }
        progress("generate: round %d, %llu draws, sorting", round, (unsigned long long)ndraw);
        radix_sort(ndraw, key.data(), id.data(), total_bits);
        keep.assign(ndraw, 0);
        unique = 0;
        for (uint64_t i = 0; i < ndraw; ++i)
            if (i == 0 || key[i] != key[i - 1]) { keep[id[i]] = 1; ++unique; }
        if (!sp->topup || unique >= want) break;
    }
    uint64_t nnz = sp->topup ? std::min(unique, want) : unique;
    progress("generate: %llu unique coordinates, drawing values", (unsigned long long)nnz);

    blco_coo* t = new (std::nothrow) blco_coo();
    if (!t) return fail(BLCO_ENOMEM, "out of host memory");
    t->nmodes = N;
    for (int n = 0; n < N; ++n) t->dims[n] = sp->dims[n];
    t->nnz = nnz;
    std::vector<uint64_t> src(nnz);  // draw index of each kept nonzero
    {
        uint64_t k = 0;
        for (uint64_t d = 0; d < ndraw && k < nnz; ++d)
            if (keep[d]) src[k++] = d;
    }
    try {
        t->cidx.resize((uint64_t)N * nnz);
        t->vals.resize(nnz);
    } catch (const std::bad_alloc&) {
        delete t;
        return fail(BLCO_ENOMEM, "generator allocation failed");
    }
    // generating factors A_n (dims[n] x gen_rank), Exp(1) entries
    const int Rg = sp->gen_rank;
    std::vector<std::vector<double>> A(N);
    for (int n = 0; n < N; ++n) {
        A[n].resize(sp->dims[n] * (uint64_t)Rg);
        double* a = A[n].data();
        uint64_t cnt = sp->dims[n] * (uint64_t)Rg;
        #pragma omp parallel for schedule(static)
        for (uint64_t e = 0; e < cnt; ++e)
            a[e] = -std::log1p(-rng_u01(sp->seed, STREAM_GEN_FACTOR + n, e));
    }
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t d = src[i];
        double rho = 0.0;
        for (int r = 0; r < Rg; ++r) {
            double p = 1.0;
            for (int n = 0; n < N; ++n) p *= A[n][draws[d * N + n] * Rg + r];
            rho += p;
        }
        for (int n = 0; n < N; ++n) t->cidx[(uint64_t)n * nnz + i] = draws[d * N + n];
        uint64_t c = ztpoisson(rho, sp->seed, d * 64);
        t->vals[i] = sp->value_kind == 1 ? std::log1p((double)c) : (double)c;
    }
    *out = t;
    return BLCO_OK;
}

extern "C" int blco_init_factor(uint64_t seed, int mode, uint64_t rows, int rank, double* U)
{
    if (!U || rank < 1 || mode < 0) return fail(BLCO_EINVAL, "bad argument");
    uint64_t cnt = rows * (uint64_t)rank;
    #pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < cnt; ++e) U[e] = rng_u01(seed, STREAM_INIT + (uint64_t)mode, e);
    return BLCO_OK;
}
