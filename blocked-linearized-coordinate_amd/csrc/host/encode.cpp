// encode.cpp -- ALTO linearisation and BLCO re-encoding on the host.
//
// Follows gen_alto / setup_packed_alto / sort_alto (include/alto.hpp:107-186,
// 597-695, 697-754) and gen_blcotensor_host (include/blco.hpp:196-331):
//   1. bits_n = 64 - clz(I_n - 1); modes are dealt one bit at a time from the
//      LSB, shortest mode first with ties in mode order, until every mode has
//      its bits (ALTO masks).
//   2. key_i = OR_n pdep(i_n, mask_n), nonzeros sorted by key.
//   3. BLCO keeps, per mode, the index bits that fell in the low 64 key bits,
//      packed mode after mode (masks 2^trunc_n - 1 at prefix positions); the
//      upper key half is the block id, and block_coords_n = pext(block id,
//      upper mask_n) << trunc_n.  Runs of one block id are cut into pieces of
//      at most max_block_size nonzeros.
// The sort is a stable parallel LSD radix sort (equal keys keep input order);
// the dense block histogram of the reference (1 << (bits - 64) counters) is
// replaced by walking runs of the sorted upper halves, which visits the same
// blocks in the same order without the 2^(bits-64) array.
#include <immintrin.h>
#include <omp.h>

#include <cstdarg>
#include <cstring>
#include <new>

#include "internal.hpp"
#include "radix.hpp"

typedef unsigned __int128 u128;

namespace blco {

void progress(const char* fmt, ...)
{
    static const bool on = getenv("BLCO_PROGRESS") != nullptr;
    if (!on) return;
    static const double t0 = omp_get_wtime();
    va_list ap;
    va_start(ap, fmt);
    fprintf(stderr, "[blco %8.2f s] ", omp_get_wtime() - t0);
    vfprintf(stderr, fmt, ap);
    fputc('\n', stderr);
    fflush(stderr);
    va_end(ap);
}

static thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

void clear_error() { g_err.clear(); }

int make_layout(int nmodes, const uint64_t* dims, Layout* L)
{
    if (nmodes < 1 || nmodes > BLCO_MAX_MODES)
        return fail(BLCO_EINVAL, "nmodes %d outside [1, %d]", nmodes, BLCO_MAX_MODES);
    *L = Layout();
    L->nmodes = nmodes;
    int order[BLCO_MAX_MODES], maxb = 0;
    for (int n = 0; n < nmodes; ++n) {
        if (dims[n] == 0) return fail(BLCO_EINVAL, "mode %d has length 0", n);
        L->dims[n] = dims[n];
        L->bits[n] = mode_bits(dims[n]);
        L->total_bits += L->bits[n];
        if (L->bits[n] > maxb) maxb = L->bits[n];
        order[n] = n;
    }
    if (L->total_bits > 128)
        return fail(BLCO_EUNSUPPORTED, "index needs %d bits (> 128)", L->total_bits);
    // shortest first, stable
    for (int i = 1; i < nmodes; ++i) {
        int v = order[i], j = i - 1;
        while (j >= 0 && L->bits[order[j]] > L->bits[v]) { order[j + 1] = order[j]; --j; }
        order[j + 1] = v;
    }
    u128 m[BLCO_MAX_MODES] = {};
    int shift = 0;
    for (int level = 0; level < maxb; ++level)
        for (int k = 0; k < nmodes; ++k)
            if (level < L->bits[order[k]]) m[order[k]] |= (u128)1 << shift++;
    for (int n = 0; n < nmodes; ++n) {
        L->alto_lo[n] = (uint64_t)m[n];
        L->alto_hi[n] = (uint64_t)(m[n] >> 64);
        L->trunc[n] = __builtin_popcountll(L->alto_lo[n]);
        L->masks[n] = L->trunc[n] >= 64 ? ~0ull : ((1ull << L->trunc[n]) - 1);
        L->pos[n] = n == 0 ? 0 : L->pos[n - 1] + L->trunc[n - 1];
    }
    return BLCO_OK;
}

static inline uint64_t pdep64(uint64_t x, uint64_t m) { return _pdep_u64(x, m); }
static inline uint64_t pext64(uint64_t x, uint64_t m) { return _pext_u64(x, m); }

// Blocks of a BLCO tensor (include/blco.hpp:250-321): runs of equal upper
// key half (hi[i] for sorted nonzero i; none: one run), each split into
// pieces of at most mbs nonzeros; block coordinates are the upper half's
// bits of every mode, shifted above its truncated lower bits.
void build_blocks_runs(const Layout& L, uint64_t nnz, uint64_t nruns, const uint64_t* start, const uint64_t* hival,
                       uint64_t mbs, blco_host* h)
{
    const int N = L.nmodes;
    for (uint64_t r = 0; r < nruns; ++r) {
        const uint64_t i = start[r], j = r + 1 < nruns ? start[r + 1] : nnz;
        const uint64_t bid = hival ? hival[r] : 0;
        for (uint64_t s = i; s < j; s += mbs) {
            const uint64_t cnt = (j - s) < mbs ? (j - s) : mbs;
            h->block_start.push_back(s);
            h->block_nnz.push_back(cnt);
            if (cnt > h->max_block_nnz) h->max_block_nnz = cnt;
            for (int n = 0; n < N; ++n)
                h->block_coords.push_back(L.alto_hi[n] ? pext64(bid, L.alto_hi[n]) << L.trunc[n] : 0);
        }
    }
}

void build_blocks(const Layout& L, uint64_t nnz, const uint64_t* hi, uint64_t mbs, blco_host* h)
{
    // runs of equal upper key half (one run without upper halves)
    std::vector<uint64_t> start, val;
    for (uint64_t i = 0; i < nnz; ++i)
        if (i == 0 || (hi && hi[i] != hi[i - 1])) {
            start.push_back(i);
            val.push_back(hi ? hi[i] : 0);
        }
    build_blocks_runs(L, nnz, start.size(), start.data(), val.data(), mbs, h);
}

template <typename K, typename P>
static int encode_impl(const Layout& L, uint64_t nnz, const uint64_t* cidx,
                       const double* vals, uint64_t mbs, blco_host* h)
{
    const int N = L.nmodes;
    std::vector<K> keys;
    std::vector<P> perm;
    try {
        keys.resize(nnz);
        perm.resize(nnz);
        h->coords.resize(nnz);
        h->vals.resize(nnz);
    } catch (const std::bad_alloc&) {
        return fail(BLCO_ENOMEM, "host allocation for %llu nonzeros failed", (unsigned long long)nnz);
    }
    int bad = 0;
    #pragma omp parallel for schedule(static) reduction(| : bad)
    for (uint64_t i = 0; i < nnz; ++i) {
        K k = 0;
        for (int n = 0; n < N; ++n) {
            uint64_t x = cidx[(uint64_t)n * nnz + i];
            bad |= (x >= L.dims[n]);
            uint64_t lo = pdep64(x, L.alto_lo[n]);
            if (sizeof(K) > 8) {
                int s = __builtin_popcountll(L.alto_lo[n]);
                uint64_t hi = s >= 64 ? 0 : pdep64(x >> s, L.alto_hi[n]);
                k |= (K)(((u128)hi << 64) | lo);
            } else {
                k |= (K)lo;
            }
        }
        keys[i] = k;
        perm[i] = (P)i;
    }
    if (bad) return fail(BLCO_EINVAL, "a coordinate is outside its mode length");
    progress("encode: %llu ALTO keys formed, sorting", (unsigned long long)nnz);
    radix_sort(nnz, keys.data(), perm.data(), L.total_bits);
    progress("encode: keys sorted, re-encoding to BLCO");
    #pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < nnz; ++i) {
        uint64_t klo = (uint64_t)keys[i];
        uint64_t c = 0;
        for (int n = 0; n < N; ++n)
            if (L.trunc[n]) c |= pext64(klo, L.alto_lo[n]) << L.pos[n];
        h->coords[i] = c;
        h->vals[i] = vals[perm[i]];
    }
    std::vector<uint64_t> hi;
    if (sizeof(K) > 8) {
        hi.resize(nnz);
        #pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < nnz; ++i) hi[i] = (uint64_t)((u128)keys[i] >> 64);
    }
    build_blocks(L, nnz, sizeof(K) > 8 ? hi.data() : nullptr, mbs, h);
    return BLCO_OK;
}

}  // namespace blco

using namespace blco;

extern "C" {

const char* blco_last_error(void) { return g_err.c_str(); }

int blco_host_encode(int nmodes, const uint64_t* dims, uint64_t nnz,
                     const uint64_t* cidx, const double* vals,
                     uint64_t max_block_size, blco_host** out)
{
    if (!out || !dims || (nnz && (!cidx || !vals))) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    blco_host* h = new (std::nothrow) blco_host();
    if (!h) return fail(BLCO_ENOMEM, "out of host memory");
    int rc = make_layout(nmodes, dims, &h->L);
    if (rc) { delete h; return rc; }
    h->nnz = nnz;
    h->max_block_size = max_block_size;
    uint64_t mbs = max_block_size ? max_block_size : (nnz ? nnz : 1);  // blco.hpp:264
    bool wide = h->L.total_bits > 64;
    bool idx64 = nnz >= (1ull << 32);
    if (wide && idx64) rc = encode_impl<u128, uint64_t>(h->L, nnz, cidx, vals, mbs, h);
    else if (wide) rc = encode_impl<u128, uint32_t>(h->L, nnz, cidx, vals, mbs, h);
    else if (idx64) rc = encode_impl<uint64_t, uint64_t>(h->L, nnz, cidx, vals, mbs, h);
    else rc = encode_impl<uint64_t, uint32_t>(h->L, nnz, cidx, vals, mbs, h);
    if (rc) { delete h; return rc; }
    *out = h;
    return BLCO_OK;
}

int blco_host_encode_coo(const blco_coo* t, uint64_t max_block_size, blco_host** out)
{
    if (!t) return fail(BLCO_EINVAL, "null tensor");
    return blco_host_encode(t->nmodes, t->dims, t->nnz, t->cidx.data(), t->vals.data(),
                            max_block_size, out);
}

int blco_host_info(const blco_host* h, int* nmodes, uint64_t* dims, uint64_t* nnz,
                   uint64_t* nblocks, uint64_t* max_block_nnz)
{
    if (!h) return fail(BLCO_EINVAL, "null tensor");
    if (nmodes) *nmodes = h->L.nmodes;
    if (dims) std::memcpy(dims, h->L.dims, sizeof(uint64_t) * h->L.nmodes);
    if (nnz) *nnz = h->nnz;
    if (nblocks) *nblocks = h->block_nnz.size();
    if (max_block_nnz) *max_block_nnz = h->max_block_nnz;
    return BLCO_OK;
}

int blco_host_layout(const blco_host* h, uint64_t* masks, int* pos, uint64_t* alto_lo,
                     uint64_t* alto_hi)
{
    if (!h) return fail(BLCO_EINVAL, "null tensor");
    for (int n = 0; n < h->L.nmodes; ++n) {
        if (masks) masks[n] = h->L.masks[n];
        if (pos) pos[n] = h->L.pos[n];
        if (alto_lo) alto_lo[n] = h->L.alto_lo[n];
        if (alto_hi) alto_hi[n] = h->L.alto_hi[n];
    }
    return BLCO_OK;
}

int blco_host_blocks(const blco_host* h, uint64_t* block_start, uint64_t* block_nnz,
                     uint64_t* block_coords)
{
    if (!h) return fail(BLCO_EINVAL, "null tensor");
    size_t nb = h->block_nnz.size();
    if (block_start) std::memcpy(block_start, h->block_start.data(), nb * sizeof(uint64_t));
    if (block_nnz) std::memcpy(block_nnz, h->block_nnz.data(), nb * sizeof(uint64_t));
    if (block_coords)
        std::memcpy(block_coords, h->block_coords.data(), h->block_coords.size() * sizeof(uint64_t));
    return BLCO_OK;
}

const uint64_t* blco_host_coords(const blco_host* h) { return h ? h->coords.data() : nullptr; }
const double* blco_host_values(const blco_host* h) { return h ? h->vals.data() : nullptr; }
void blco_host_free(blco_host* h) { delete h; }

}  // extern "C"
