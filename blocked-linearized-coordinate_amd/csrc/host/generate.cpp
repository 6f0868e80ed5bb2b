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
#include <limits>
#include <new>

#include "internal.hpp"
#include "radix.hpp"
#include "rng.hpp"

typedef unsigned __int128 u128;

namespace blco {

void sampler_tables(uint64_t I, double s, std::vector<double>* cdf, std::vector<uint64_t>* guide)
{
    cdf->clear();
    guide->clear();
    if (!(s > 0.0)) return;
    cdf->resize(I);
    double* c = cdf->data();
    double acc = 0.0;
    for (uint64_t k = 0; k < I; ++k) {
        acc += std::pow((double)(k + 1), -s);
        c[k] = acc;
    }
    for (uint64_t k = 0; k < I; ++k) c[k] /= acc;
    c[I - 1] = 1.0;
    guide->resize(SAMPLER_GUIDE + 1);
    for (uint32_t g = 0; g <= SAMPLER_GUIDE; ++g)
        (*guide)[g] = (uint64_t)(std::upper_bound(c, c + I, (double)g / SAMPLER_GUIDE) - c);
}

void generating_factor(uint64_t seed, int mode, uint64_t rows, int Rg, double* a)
{
    const uint64_t cnt = rows * (uint64_t)Rg;
    #pragma omp parallel for schedule(static)
    for (uint64_t e = 0; e < cnt; ++e) a[e] = -std::log1p(-rng_u01(seed, STREAM_GEN_FACTOR + (uint64_t)mode, e));
}

double log1p_count(uint64_t c) { return std::log1p((double)c); }

struct IndexSampler {
    static constexpr uint32_t GUIDE = SAMPLER_GUIDE;
    uint64_t I = 1;
    double s = 0.0;
    std::vector<double> cdf;      // Zipf only
    std::vector<uint64_t> guide;  // guide[g] = first k with cdf[k] > g / GUIDE

    void init(uint64_t len, double zs)
    {
        I = len;
        s = zs;
        sampler_tables(I, s, &cdf, &guide);
    }
    // Inverse CDF: the first k with cdf[k] > u (an upper bound search over
    // the guide interval [guide[g], guide[g + 1]] that holds it).
    uint64_t draw(double u) const
    {
        if (s > 0.0) {
            const uint32_t g = (uint32_t)(u * GUIDE);
            const uint64_t lo = guide[g], hi = std::min<uint64_t>(guide[g + 1] + 1, I);
            uint64_t k = (uint64_t)(std::upper_bound(cdf.begin() + lo, cdf.begin() + hi, u) - cdf.begin());
            return k < I ? k : I - 1;
        }
        uint64_t k = (uint64_t)(u * (double)I);
        return k < I ? k : I - 1;
    }
};

// Linear key of draw d: per-mode bit fields of bits_n bits (injective).
template <typename K>
inline K draw_key(const blco_gen_spec* sp, const IndexSampler* smp, const Layout& L, uint64_t d)
{
    K k = 0;
    int sh = 0;
    for (int n = 0; n < sp->nmodes; ++n) {
        k |= (K)smp[n].draw(rng_u01(sp->seed, STREAM_INDEX + n, d)) << sh;
        sh += L.bits[n];
    }
    return k;
}

// Draws until `want` unique coordinates exist (or once, without top-up) and
// returns the draw indices of the first unique draws in draw order.  Each
// round's size follows the unique yield of the draws so far, so a skewed
// (Zipf) tensor with many duplicates converges in a few rounds; the result
// does not depend on the round sizes (the draws are a prefix of one counter
// sequence).  32-bit draw ids while they fit.
template <typename K, typename P>
int unique_draws_ids(const blco_gen_spec* sp, const IndexSampler* smp, const Layout& L, uint64_t want,
                     std::vector<uint64_t>* src, uint64_t* ndraw_out)
{
    uint64_t ndraw = 0, unique = 0;
    std::vector<uint64_t> first;  // bitmap: draw is the first of its coordinate
    for (int round = 0; round < 64; ++round) {
        uint64_t add = want;
        if (round > 0) {
            const double yield = std::max(0.01, (double)unique / (double)ndraw);
            add = (uint64_t)((double)(want - unique) / yield * 1.15) + 1024;
        }
        if ((double)ndraw + (double)add > (double)std::numeric_limits<P>::max()) break;
        ndraw += add;
        std::vector<K> key;
        std::vector<P> id;
        try {
            key.resize(ndraw);
            id.resize(ndraw);
            first.assign((ndraw + 63) / 64, 0);
        } catch (const std::bad_alloc&) {
            return fail(BLCO_ENOMEM, "generator allocation failed");
        }
        #pragma omp parallel for schedule(static)
        for (uint64_t d = 0; d < ndraw; ++d) {
            key[d] = draw_key<K>(sp, smp, L, d);
            id[d] = (P)d;
        }
        progress("generate: round %d, %llu draws, sorting", round, (unsigned long long)ndraw);
        radix_sort(ndraw, key.data(), id.data(), L.total_bits);
        progress("generate: round %d sorted", round);
        uint64_t u = 0;
        uint64_t* bits = first.data();
        #pragma omp parallel for schedule(static) reduction(+ : u)
        for (uint64_t i = 0; i < ndraw; ++i)
            if (i == 0 || key[i] != key[i - 1]) {
                __atomic_fetch_or(&bits[id[i] >> 6], 1ull << (id[i] & 63), __ATOMIC_RELAXED);
                ++u;
            }
        unique = u;
        if (!sp->topup || unique >= want) break;
    }
    const uint64_t nnz = sp->topup ? std::min(unique, want) : unique;
    try {
        src->resize(nnz);
    } catch (const std::bad_alloc&) {
        return fail(BLCO_ENOMEM, "generator allocation failed");
    }
    uint64_t k = 0;
    for (uint64_t d = 0; d < ndraw && k < nnz; ++d)
        if (first[d >> 6] >> (d & 63) & 1) (*src)[k++] = d;
    *ndraw_out = ndraw;
    return BLCO_OK;
}

template <typename K>
int unique_draws(const blco_gen_spec* sp, const IndexSampler* smp, const Layout& L, uint64_t want,
                 std::vector<uint64_t>* src, uint64_t* ndraw)
{
    // top-up rounds draw at most a few times `want`
    if ((double)want * 8.0 < (double)std::numeric_limits<uint32_t>::max())
        return unique_draws_ids<K, uint32_t>(sp, smp, L, want, src, ndraw);
    return unique_draws_ids<K, uint64_t>(sp, smp, L, want, src, ndraw);
}

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
    // draw), and top up until enough unique coordinates exist.  Draws are
    // not stored: draw d's indices are a function of (seed, d) and are
    // recomputed where needed.
    const uint64_t want = sp->target_nnz;
    std::vector<uint64_t> src;  // draw index of each kept nonzero, in draw order
    uint64_t ndraw = 0;
    if (L.total_bits <= 64)
        rc = unique_draws<uint64_t>(sp, smp, L, want, &src, &ndraw);
    else
        rc = unique_draws<u128>(sp, smp, L, want, &src, &ndraw);
    if (rc) return rc;
    const uint64_t nnz = src.size();
    progress("generate: %llu unique coordinates from %llu draws, drawing values", (unsigned long long)nnz,
             (unsigned long long)ndraw);

    blco_coo* t = new (std::nothrow) blco_coo();
    if (!t) return fail(BLCO_ENOMEM, "out of host memory");
    t->nmodes = N;
    for (int n = 0; n < N; ++n) t->dims[n] = sp->dims[n];
    t->nnz = nnz;
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
        generating_factor(sp->seed, n, sp->dims[n], Rg, A[n].data());
    }
    progress("generate: generating factors drawn");
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        const uint64_t d = src[i];
        uint64_t ix[BLCO_MAX_MODES];
        for (int n = 0; n < N; ++n) ix[n] = smp[n].draw(rng_u01(sp->seed, STREAM_INDEX + n, d));
        double rho = 0.0;
        for (int r = 0; r < Rg; ++r) {
            double p = 1.0;
            for (int n = 0; n < N; ++n) p *= A[n][ix[n] * Rg + r];
            rho += p;
        }
        for (int n = 0; n < N; ++n) t->cidx[(uint64_t)n * nnz + i] = ix[n];
        uint64_t c = ztpoisson(rho, sp->seed, d * 64);
        t->vals[i] = sp->value_kind == 1 ? std::log1p((double)c) : (double)c;
    }
    progress("generate: values drawn");
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
