// device.hip -- BLCO tensors in HBM (or streamed from pinned host memory),
// MTTKRP dispatch and the stand-alone ALS-step entry points of the C ABI.
//
// Layout in HBM: coords u64[nnz] and values (f64 or f32)[nnz] in ALTO order,
// one contiguous allocation each, blocks back to back (block b occupies
// [block_start[b], block_start[b] + block_nnz[b])); block coordinates
// u64[nblocks][nmodes]; a chunk table per tiles-per-wave setting (one entry
// per wave of work: begin, length, block).  Streamed tensors keep coords and
// values in pinned host memory and pass them through two device slots of
// slot_nnz nonzeros, copying on a side stream while the previous slot is
// being consumed (src/alto_dev.cu:1869-1874, 1987 stream one block per
// launch through NUM_STREAMS slots).
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "als.hpp"
#include "device.hpp"
#include "devcommon.hpp"
#include "mttkrp.hpp"

using namespace blco;

namespace blco {

// LDS for a privatised output: with the <= 12 KiB of per-wave staging the
// workgroup stays within the 64 KiB a launch gets without an attribute.
constexpr uint64_t LDS_OUT_MAX = 48 * 1024;
constexpr int LVL3_MAX_MODE_LENGTH = 100;     // src/utils.hpp:13
constexpr int DEFAULT_TPW = 4;                // tiles of 64 per wave
constexpr int SORTED_TPW = 16;                // tiles of 64 per wave, row-ordered kernel (most)
constexpr uint64_t SORTED_TARGET_CHUNKS = 8192;
constexpr int DEFAULT_NPRTN = 1;
constexpr uint32_t LVL3_CHUNKS_PER_WG = 64;
constexpr uint32_t HOT_EXCLUDE = 64;  // hot rows kept out of the row windows

static int value_size(int vdt) { return vdt == BLCO_F64 ? 8 : 4; }

// Default tiles of 64 nonzeros per chunk of the row-ordered kernel.  A
// tensor of up to 2^22 nonzeros takes shorter chunks, so that the grid keeps
// 4096-8192 waves (2 tiles up to 2^20 nonzeros, 4 up to 2^21, 8 up to 2^22):
// with 16 tiles such a tensor fills a fraction of the 256 CUs, and 2 tiles
// ran tensors of 1e5-3e5 nonzeros in a third of the time (the total over
// chunk targets from 2000 to 64000 differs by 2%).  From 2^22 nonzeros 16,
// and 32 for FP64 rows of 256 bytes or more from 2^26 nonzeros: fewer
// chunk ends, each of which adds its open rows to HBM (3-7% faster on the
// 77M-nonzero modes, 0.5% on the 1.7B-nonzero ones; FP32 rows of that size
// were faster at 16).
static int sorted_tpw(const blco_dev* d, int rank, int ftype)
{
    if (ftype == BLCO_F64 && rank >= 32 && d->nnz >= (1ull << 26)) return 32;
    int t = 2;
    while (t < SORTED_TPW && d->nnz > (uint64_t)WAVE * t * SORTED_TARGET_CHUNKS) t *= 2;
    return t;
}

const ChunkTab* chunk_table(blco_dev* d, int tpw)
{
    for (auto& t : d->tabs)
        if (t.tpw == tpw) return &t;
    ChunkTab t;
    t.tpw = tpw;
    const uint64_t CH = (uint64_t)WAVE * tpw;
    std::vector<Chunk> h;
    t.first.push_back(0);
    for (uint64_t b = 0; b < d->nblocks; ++b) {
        for (uint64_t s = 0; s < d->bnnz[b]; s += CH) {
            Chunk c;
            c.begin = d->bstart[b] + s;
            c.len = (uint32_t)std::min<uint64_t>(CH, d->bnnz[b] - s);
            c.block = (uint32_t)b;
            h.push_back(c);
        }
        t.first.push_back((uint32_t)h.size());
    }
    t.n = (uint32_t)h.size();
    // streaming units: chunk ranges of one block holding at most slot_nnz
    if (d->stream_data) {
        for (uint64_t b = 0; b < d->nblocks; ++b) {
            uint32_t c0 = t.first[b], c1 = t.first[b + 1];
            uint32_t u0 = c0;
            uint64_t acc = 0;
            for (uint32_t c = c0; c < c1; ++c) {
                if (acc + h[c].len > d->slot_nnz && c > u0) {
                    t.units.push_back({u0, c, h[u0].begin, acc});
                    u0 = c;
                    acc = 0;
                }
                acc += h[c].len;
            }
            if (c1 > u0) t.units.push_back({u0, c1, h[u0].begin, acc});
        }
    }
    if (t.n) {
        if (hipMalloc(&t.d, sizeof(Chunk) * t.n) != hipSuccess) {
            fail(BLCO_ENOMEM, "device allocation of the chunk table failed");
            return nullptr;
        }
        if (hipMemcpy(t.d, h.data(), sizeof(Chunk) * t.n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(t.d);
            fail(BLCO_EHIP, "upload of the chunk table failed");
            return nullptr;
        }
        d->bytes += sizeof(Chunk) * t.n;
    }
    d->tabs.push_back(t);
    return &d->tabs.back();
}

const WindowTab* window_table(blco_dev* d, int k)
{
    for (auto& t : d->wtabs)
        if (t.k == k) return &t;
    WindowTab t;
    t.k = k;
    const uint64_t nfine = d->fine_begin.size();
    const int N = d->nmodes;
    const uint32_t G = 1u << k;
    t.first.push_back(0);
    std::vector<uint64_t> lo, hi;  // per entry per mode
    uint64_t f = 0;
    for (uint64_t b = 0; b < d->nblocks; ++b) {
        while (f < nfine && d->fine_block[f] == b) {
            uint64_t g1 = f;
            SuperChunk sc;
            sc.begin = d->fine_begin[f];
            sc.len = 0;
            sc.block = (uint32_t)b;
            uint64_t l[BLCO_MAX_MODES], h[BLCO_MAX_MODES];
            for (int m = 0; m < N; ++m) { l[m] = ~0ull; h[m] = 0; }
            while (g1 < nfine && d->fine_block[g1] == b && g1 - f < G) {
                sc.len += d->fine_len[g1];
                for (int m = 0; m < N; ++m) {
                    l[m] = std::min(l[m], d->fine_lo[m * nfine + g1]);
                    h[m] = std::max(h[m], d->fine_hi[m * nfine + g1]);
                }
                ++g1;
            }
            t.h_sc.push_back(sc);
            for (int m = 0; m < N; ++m) { lo.push_back(l[m]); hi.push_back(h[m]); }
            f = g1;
        }
        t.first.push_back((uint32_t)t.h_sc.size());
    }
    t.n = (uint32_t)t.h_sc.size();
    t.h_range.resize((uint64_t)N * t.n * 2);
    for (uint32_t e = 0; e < t.n; ++e)
        for (int m = 0; m < N; ++m) {
            t.h_range[((uint64_t)m * t.n + e) * 2] = lo[(uint64_t)e * N + m];
            t.h_range[((uint64_t)m * t.n + e) * 2 + 1] = hi[(uint64_t)e * N + m];
        }
    if (t.n) {
        if (hipMalloc(&t.d_sc, sizeof(SuperChunk) * t.n) != hipSuccess ||
            hipMalloc(&t.d_range, sizeof(uint64_t) * t.h_range.size()) != hipSuccess) {
            if (t.d_sc) (void)hipFree(t.d_sc);
            fail(BLCO_ENOMEM, "device allocation of the window table failed");
            return nullptr;
        }
        if (hipMemcpy(t.d_sc, t.h_sc.data(), sizeof(SuperChunk) * t.n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(t.d_range, t.h_range.data(), sizeof(uint64_t) * t.h_range.size(),
                      hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(t.d_sc);
            (void)hipFree(t.d_range);
            fail(BLCO_EHIP, "upload of the window table failed");
            return nullptr;
        }
        d->bytes += sizeof(SuperChunk) * t.n + sizeof(uint64_t) * t.h_range.size();
    }
    d->wtabs.push_back(t);
    return &d->wtabs.back();
}

// Partial rows of one window plan stay under this many bytes; past it the
// cost model moves to larger super-chunks.
constexpr uint64_t PART_MAX_BYTES = 4ull << 30;
// An HBM atomic add of one row costs about this many plain row stores.
constexpr double ATOMIC_ROW_COST = 32.0;
// Partial rows summed by one reduce lane group before it adds to the output.
constexpr uint64_t REDUCE_SEG = 16;

// Per-entry use of the LDS window: it pays when storing the window's rows
// (and reading them back in the reduce) costs less than the HBM atomics of
// the entry's nonzeros that it absorbs.
static inline bool use_window(uint64_t rows, uint64_t cap, double len)
{
    const uint64_t w = std::min(rows, cap);
    const double absorbed = rows ? len * (double)w / (double)rows : 0.0;
    return w > 0 && 2.0 * (double)w < ATOMIC_ROW_COST * absorbed;
}

// Window rows of each super-chunk of 2^k fine chunks for a mode.
template <typename F>
static void walk_entries(const blco_dev* d, int mode, int k, F&& visit)
{
    const uint64_t nfine = d->fine_begin.size();
    const uint64_t G = 1ull << k;
    uint64_t f = 0;
    while (f < nfine) {
        const uint32_t b = d->fine_block[f];
        uint64_t g1 = f, lo = ~0ull, hi = 0, len = 0;
        while (g1 < nfine && d->fine_block[g1] == b && g1 - f < G) {
            lo = std::min(lo, d->fine_lo[mode * nfine + g1]);
            hi = std::max(hi, d->fine_hi[mode * nfine + g1]);
            len += d->fine_len[g1];
            ++g1;
        }
        visit(hi >= lo ? hi - lo + 1 : 0, len);
        f = g1;
    }
}

// Super-chunk size (2^k fine chunks) and hot-area size for a mode: minimise
// an estimate of the HBM work -- each used window row is stored once and
// read once by the reduce, each hot area likewise; nonzeros outside a window
// (or in hot rows when there is no hot area) are added into the output one
// by one -- keeping >= 1024 workgroups and the partial rows under
// PART_MAX_BYTES.
static void choose_window(const blco_dev* d, int mode, uint32_t total, uint64_t row_bytes, int* k_out,
                          uint32_t* hot_out)
{
    const uint32_t H = (uint32_t)std::min<size_t>(d->hot_list_h[mode].size(), total / 2);
    uint64_t hot_nnz = 0;
    for (uint32_t h = 0; h < H; ++h) hot_nnz += d->hot_cnt_h[mode][h];
    const double hf = d->nnz ? (double)hot_nnz / (double)d->nnz : 0.0;
    int best_k = 0, last_k = 0;
    uint32_t best_h = 0;
    double best_cost = -1.0;
    for (int k = 0; k <= 8; ++k) {
        for (uint32_t hot : {H, 0u}) {
            const uint64_t cap = total - hot;
            double cost = 0.0;
            uint64_t entries = 0, prows = 0;
            walk_entries(d, mode, k, [&](uint64_t rows, uint64_t len) {
                const double l = (double)len * (1.0 - hf);  // nonzeros outside hot rows
                if (use_window(rows, cap, l)) {
                    const uint64_t w = std::min(rows, cap);
                    prows += w;
                    cost += 2.0 * (double)w + ATOMIC_ROW_COST * l * (double)(rows - w) / (double)rows;
                } else {
                    cost += ATOMIC_ROW_COST * l;
                }
                ++entries;
            });
            if (k > 0 && entries < 1024) goto done;  // keep >= 4 workgroups per CU
            prows += (uint64_t)hot * entries;
            cost += hot ? 2.0 * hot * (double)entries : 4.0 * ATOMIC_ROW_COST * (double)hot_nnz;
            cost += 4.0 * (double)entries;  // fixed per-workgroup work
            last_k = k;
            if (prows * row_bytes > PART_MAX_BYTES) continue;
            if (best_cost < 0 || cost < best_cost) {
                best_cost = cost;
                best_k = k;
                best_h = hot;
            }
            if (H == 0) break;
        }
    }
done:
    if (best_cost < 0) {  // nothing under the cap: largest super-chunks, no hot area
        best_k = last_k;
        best_h = 0;
    }
    *k_out = best_k;
    *hot_out = best_h;
}

const WindowPlan* window_plan(blco_dev* d, int mode, int rank, int acc)
{
    for (auto& p : d->wplans)
        if (p.mode == mode && p.rank == rank && p.acc == acc) return &p;
    WindowPlan p;
    p.mode = mode;
    p.rank = rank;
    p.acc = acc;
    const uint64_t row_bytes = (uint64_t)rank * acc;
    const uint32_t total = (uint32_t)(WINDOW_LDS_BYTES / row_bytes);
    int k = 0;
    choose_window(d, mode, total, row_bytes, &k, &p.hot_rows);
    p.lds_rows = total - p.hot_rows;
    p.wt = window_table(d, k);
    if (!p.wt) return nullptr;
    const WindowTab& t = *p.wt;
    const uint64_t* rg = t.h_range.data() + (uint64_t)mode * t.n * 2;
    uint64_t hot_nnz = 0;
    for (uint32_t h = 0; h < p.hot_rows; ++h) hot_nnz += d->hot_cnt_h[mode][h];
    const double hf = d->nnz ? (double)hot_nnz / (double)d->nnz : 0.0;
    // partial rows: each entry's window (when it pays), then each hot area;
    // poff[e + 1] - poff[e] is entry e's window height
    std::vector<uint64_t> poff(t.n + 1);
    uint64_t P = 0;
    for (uint32_t e = 0; e < t.n; ++e) {
        poff[e] = P;
        const uint64_t lo = rg[2 * e], hi = rg[2 * e + 1];
        const uint64_t rows = hi >= lo ? hi - lo + 1 : 0;
        if (use_window(rows, p.lds_rows, (double)t.h_sc[e].len * (1.0 - hf)))
            P += std::min<uint64_t>(rows, p.lds_rows);
    }
    poff[t.n] = P;
    p.hot_poff = P;
    P += (uint64_t)t.n * p.hot_rows;
    p.prows = P;
    if (P > 0xFFFFFFFFull) {
        fail(BLCO_EUNSUPPORTED, "%llu window partial rows exceed the 32-bit reduce index", (unsigned long long)P);
        return nullptr;
    }
    // output row of every partial row, then the partial rows grouped by it
    std::vector<uint64_t> prow(P);
    for (uint32_t e = 0; e < t.n; ++e) {
        const uint64_t lo = rg[2 * e];
        for (uint64_t i = 0; i < poff[e + 1] - poff[e]; ++i) prow[poff[e] + i] = lo + i;
        for (uint32_t h = 0; h < p.hot_rows; ++h)
            prow[p.hot_poff + (uint64_t)e * p.hot_rows + h] = d->hot_list_h[mode][h];
    }
    std::vector<uint32_t> order(P);
    if (d->dims[mode] <= (1ull << 27)) {  // counting sort by output row
        std::vector<uint32_t> at(d->dims[mode] + 1, 0);
        for (uint64_t i = 0; i < P; ++i) ++at[prow[i] + 1];
        for (uint64_t r = 0; r < d->dims[mode]; ++r) at[r + 1] += at[r];
        for (uint64_t i = 0; i < P; ++i) order[at[prow[i]]++] = (uint32_t)i;
    } else {
        for (uint64_t i = 0; i < P; ++i) order[i] = (uint32_t)i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return prow[x] < prow[y]; });
    }
    // segments of at most REDUCE_SEG partial rows of one output row; a row
    // split over several segments is marked (top bit) for an atomic add
    std::vector<uint64_t> rows, ptr;
    for (uint64_t i = 0; i < P;) {
        uint64_t j = i + 1;
        while (j < P && prow[order[j]] == prow[order[i]]) ++j;
        const uint64_t flag = j - i > REDUCE_SEG ? REDUCE_SHARED : 0;
        for (uint64_t s0 = i; s0 < j; s0 += REDUCE_SEG) {
            rows.push_back(prow[order[i]] | flag);
            ptr.push_back(s0);
        }
        i = j;
    }
    ptr.push_back(P);
    p.nred = rows.size();
    auto up = [&](auto** dst, const auto& v) {
        const size_t bytes = sizeof(v[0]) * v.size();
        if (!bytes) return true;
        if (hipMalloc(dst, bytes) != hipSuccess) return false;
        d->bytes += bytes;
        return hipMemcpy(*dst, v.data(), bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(&p.d_poff, poff) || !up(&p.d_rows, rows) || !up(&p.d_ptr, ptr) || !up(&p.d_idx, order)) {
        if (p.d_poff) (void)hipFree(p.d_poff);
        if (p.d_rows) (void)hipFree(p.d_rows);
        if (p.d_ptr) (void)hipFree(p.d_ptr);
        if (p.d_idx) (void)hipFree(p.d_idx);
        fail(BLCO_ENOMEM, "device allocation or upload of the window reduce plan failed");
        return nullptr;
    }
    d->wplans.push_back(p);
    return &d->wplans.back();
}

bool lvl3_output_fits(const blco_dev* d, int mode, int rank, int ftype)
{
    uint64_t esz = ftype == BLCO_F64 ? 8 : 4;
    return d->dims[mode] * (uint64_t)rank * esz <= LDS_OUT_MAX;
}

int auto_kernel(const blco_dev* d, int mode, int rank, int ftype)
{
    // src/alto_dev.cu:1975,1996-2001: short modes -> lvl3, else lvl1; here
    // the long modes run the row-ordered kernel when the tensor is resident
    // (the row window when it is streamed or the copy does not fit).
    if (d->dims[mode] <= (uint64_t)LVL3_MAX_MODE_LENGTH && lvl3_output_fits(d, mode, rank, ftype))
        return BLCO_K_LVL3;
    return d->stream_data || d->sorted_failed[mode] ? BLCO_K_BATCHED2 : BLCO_K_SORTED;
}

static int ensure_partials(blco_dev* d, size_t bytes)
{
    if (d->partials_bytes >= bytes) return BLCO_OK;
    if (d->partials) (void)hipFree(d->partials);
    d->partials = nullptr;
    d->partials_bytes = 0;
    HIP_TRY(hipMalloc(&d->partials, bytes));
    d->partials_bytes = bytes;
    return BLCO_OK;
}

// Tiles of the split mode's index for a row-ordered copy: NUM_XCD tiles cut
// at nonzero quantiles (tile_of[i] for every index i), or 1 when the copy is
// left untiled:
//  - small tensors (fewer than XCD_TILE_MIN_NNZ nonzeros);
//  - a split factor that one XCD's L2 holds whole anyway (below
//    XCD_TILE_MIN_FACTOR bytes at the rank and factor type of the first
//    call; tiling a smaller split mode, such as a short outer mode of a
//    fiber-ordered copy, only multiplies the row pieces);
//  - output rows with few nonzeros: tiling cuts every row into up to NUM_XCD
//    pieces, each flushed to HBM on its own, so it is declined when the added
//    pieces exceed one per XCD_TILE_RUN_RATIO nonzeros;
//  - split modes longer than 2^32, or a split index so hot that the largest
//    tile would hold more than XCD_TILE_SLACK times its share.
// BLCO_XCD_TILES=0 turns tiling off and =1 forces it (tests).
constexpr uint64_t XCD_TILE_MIN_NNZ = 1ull << 22;
constexpr uint64_t XCD_TILE_MIN_FACTOR = 2ull << 20;
constexpr uint64_t XCD_TILE_RUN_RATIO = 256;
constexpr double XCD_TILE_SLACK = 1.15;
// The fiber kernel runs on a fiber-ordered copy whose fibers average at
// least this many nonzeros (below, the plain row-ordered kernel runs on it).
// BLCO_FIBERS=0 keeps the ALTO order within rows (no fibers), =1 runs the
// fiber kernel whatever the fiber length (tests).
constexpr double FIBER_MIN_LEN = 1.25;
// Fiber order is chosen only when the inner factor holds at least this many
// bytes at the first call's rank and type (smaller ones stay in L2, where
// the gathers the fibers save are cheap).  The FP64 fiber kernel holds 108
// VGPRs (4 waves per SIMD against 5 for the plain kernel), so for FP64
// factors it pays only once most inner gathers miss the caches.
constexpr uint64_t FIBER_MIN_INNER = 2ull << 20;
constexpr uint64_t FIBER_MIN_INNER_F64 = 64ull << 20;

// Counts behind the decision come from the host copy (coords, bc) or, when
// coords is empty, from the device tensor.
static int xcd_tiles(const blco_dev* d, int mode, int split, uint64_t row_bytes,
                     const std::vector<uint64_t>& coords, const std::vector<uint64_t>& bc,
                     std::vector<uint8_t>* tile_of)
{
    const bool on_device = coords.empty();
    const char* env = getenv("BLCO_XCD_TILES");
    const bool force = env && env[0] == '1', off = env && env[0] == '0';
    if (off || split < 0) return 1;
    const uint64_t n = d->nnz, I = d->dims[split], rows = d->dims[mode];
    if (I > (1ull << 32) || rows > (1ull << 32)) return 1;
    if (!force && (n < XCD_TILE_MIN_NNZ || I * row_bytes < XCD_TILE_MIN_FACTOR)) return 1;
    const int N = d->nmodes;
    auto index = [&](int m, uint64_t b, uint64_t i) {
        return ((coords[i] >> d->pos[m]) & d->masks[m]) | bc[b * N + m];
    };
    std::vector<uint64_t> cnt(I, 0);
    if (on_device) {
        if (dev_split_counts(d, split, &cnt)) return 1;  // no tiling when the counts fail
    } else {
        for (uint64_t b = 0; b < d->nblocks; ++b)
            for (uint64_t i = d->bstart[b]; i < d->bstart[b] + d->bnnz[b]; ++i) ++cnt[index(split, b, i)];
    }
    tile_of->assign(I, 0);
    uint64_t cum = 0, tile_nnz[NUM_XCD] = {};
    for (uint64_t i = 0; i < I; ++i) {
        const uint64_t t = std::min<uint64_t>(NUM_XCD - 1, (cum + cnt[i] / 2) * NUM_XCD / n);
        (*tile_of)[i] = (uint8_t)t;
        tile_nnz[t] += cnt[i];
        cum += cnt[i];
    }
    if (force) return NUM_XCD;
    const uint64_t most = *std::max_element(tile_nnz, tile_nnz + NUM_XCD);
    if ((double)most > XCD_TILE_SLACK * (double)n / NUM_XCD) return 1;
    // the tiles each output row meets
    std::vector<uint8_t> met(rows, 0);
    const uint8_t* to = tile_of->data();
    if (on_device) {
        if (dev_row_tiles(d, mode, split, *tile_of, &met)) return 1;
    } else {
        for (uint64_t b = 0; b < d->nblocks; ++b)
            for (uint64_t i = d->bstart[b]; i < d->bstart[b] + d->bnnz[b]; ++i)
                met[index(mode, b, i)] |= (uint8_t)(1u << to[index(split, b, i)]);
    }
    uint64_t added = 0;
    for (uint64_t r = 0; r < rows; ++r)
        if (met[r]) added += (uint64_t)__builtin_popcount(met[r]) - 1;
    return added * XCD_TILE_RUN_RATIO <= n ? NUM_XCD : 1;
}

// Device bytes that hold nonzeros: the resident ones, the two streaming
// slots and the row-ordered copies (what a device budget caps; the chunk,
// window and hot-row tables are outside it).
static uint64_t nonzero_bytes(const blco_dev* d)
{
    const uint64_t per_nz = 8 + (uint64_t)value_size(d->vdtype);
    uint64_t b = d->res_nnz * per_nz;
    if (d->stream_data) b += 2 * std::min<uint64_t>(d->slot_nnz, d->nnz ? d->nnz : 1) * per_nz;
    for (int m = 0; m < d->nmodes; ++m)
        if (d->d_scoords[m]) b += d->nnz * per_nz;
    return b;
}

// Builds the row-ordered copy of a mode (kernel 14) from the resident
// tensor, once per mode: on the GPU (roworder_dev.hip), or with
// BLCO_ROWORDER=host by a download, a stable sort on the host and an upload
// (the same copy; the host build is the reference of the tests).  The copy
// is grouped by (tile of the split mode, block) and sorted by the mode index
// within a group, where the split mode is the longest other mode: the
// chunks of tile t all run on XCD t, whose L2 then holds only 1/NUM_XCD of
// the split mode's factor rows.
static int ensure_row_ordered(blco_dev* d, int mode, int rank, int ftype)
{
    if (d->d_scoords[mode] || !d->nnz) return BLCO_OK;
    const uint64_t n = d->nnz, nb = d->nblocks;
    const int N = d->nmodes;
    const size_t vsz = value_size(d->vdtype);
    if (d->budget && nonzero_bytes(d) + n * (8 + vsz) > d->budget)
        return fail(BLCO_ENOMEM, "the row-ordered copy of mode %d (%llu bytes) exceeds the device budget "
                    "(%llu of %llu bytes held by nonzeros)", mode, (unsigned long long)(n * (8 + vsz)),
                    (unsigned long long)nonzero_bytes(d), (unsigned long long)d->budget);
    const char* renv = getenv("BLCO_ROWORDER");
    bool on_host = renv && std::strcmp(renv, "host") == 0;
    progress("row-ordered copy of mode %d: %llu nonzeros (%s)", mode, (unsigned long long)n,
             on_host ? "host" : "device");
    // fiber order: the inner mode is the longest input mode, the others are
    // outer; the split mode for the XCD tiles is then the longest outer mode
    // (a split of the inner mode would cut the fibers), else (ALTO order
    // within rows) the longest input mode.
    RowOrderSpec o;
    o.mode = mode;
    o.nmodes = N;
    o.masks = d->masks;
    o.pos = d->pos;
    o.nblocks = nb;
    o.block_start = d->bstart.data();
    o.block_nnz = d->bnnz.data();
    int inner = -1;
    const char* fenv = getenv("BLCO_FIBERS");
    if (N >= 3 && !(fenv && fenv[0] == '0')) {
        for (int m = 0; m < N; ++m)
            if (m != mode && (inner < 0 || d->dims[m] > d->dims[inner])) inner = m;
        int bits = 0;
        for (int m = 0; m < N; ++m)
            if (m != mode && m != inner) {
                o.outer[o.nouter++] = m;
                bits += d->masks[m] ? 64 - __builtin_clzll(d->masks[m]) : 0;
            }
        // fibers save inner-factor gathers, which pay when that factor is
        // too large for one XCD's L2
        const bool big = d->dims[inner] * (uint64_t)rank * dtype_size(ftype) >=
                         (ftype == BLCO_F64 ? FIBER_MIN_INNER_F64 : FIBER_MIN_INNER);
        if (bits > 64 || !(big || (fenv && fenv[0] == '1'))) {
            inner = -1;
            o.nouter = 0;
        }
    }
    int split = -1;
    for (int m = 0; m < N; ++m)
        if (m != mode && m != inner && (split < 0 || d->dims[m] > d->dims[split])) split = m;
    std::vector<uint64_t> coords, bc;
    std::vector<char> vals;
    auto download = [&]() -> int {
        try {
            coords.resize(n);
            vals.resize(n * vsz);
            bc.resize(nb * N);
        } catch (const std::bad_alloc&) {
            return fail(BLCO_ENOMEM, "host copy for the row-ordered copy of mode %d", mode);
        }
        HIP_TRY(hipMemcpy(coords.data(), d->d_coords, n * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(vals.data(), d->d_vals, n * vsz, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(bc.data(), d->d_bcoords, nb * N * 8, hipMemcpyDeviceToHost));
        return BLCO_OK;
    };
    if (on_host) {
        int rc = download();
        if (rc) return rc;
    }
    o.bcoords = on_host ? bc.data() : nullptr;
    std::vector<uint8_t> tile_of;
    const int ntiles = xcd_tiles(d, mode, split, (uint64_t)rank * dtype_size(ftype), coords, bc, &tile_of);
    o.split = split;
    o.ntiles = ntiles;
    o.tile_of = tile_of.data();
    std::vector<SGroup> groups;
    std::vector<uint64_t> runs;
    uint64_t fibers = 0;
    uint64_t* dc = nullptr;
    void* dv = nullptr;
    if (!on_host) {
        const int rc = dev_row_copy(d, o, &groups, &runs, inner >= 0 ? &fibers : nullptr, &dc, &dv);
        if (rc == BLCO_EUNSUPPORTED) {  // key wider than 64 bits: the host build
            on_host = true;
            int r2 = download();
            if (r2) return r2;
            o.bcoords = bc.data();
        } else if (rc) {
            return rc;
        }
    }
    if (on_host) {
        std::vector<uint64_t> perm(n), sc(n);
        std::vector<char> sv(n * vsz);
        std::vector<uint64_t> gnnz((uint64_t)ntiles * nb);
        row_order(o, n, coords.data(), perm.data(), gnnz.data());
        gather_bytes(coords.data(), 8, perm.data(), n, sc.data());
        gather_bytes(vals.data(), vsz, perm.data(), n, sv.data());
        // groups in storage order, and run starts of equal (group, row)
        const uint64_t mk = d->masks[mode];
        const int ps = d->pos[mode];
        uint64_t start = 0;
        for (uint64_t g = 0; g < gnnz.size(); ++g) {
            if (!gnnz[g]) continue;
            groups.push_back(SGroup{start, gnnz[g], (uint32_t)(g % nb), (uint32_t)(g / nb)});
            for (uint64_t i = start; i < start + gnnz[g]; ++i)
                if (i == start || ((sc[i] >> ps) & mk) != ((sc[i - 1] >> ps) & mk)) runs.push_back(i);
            start += gnnz[g];
        }
        runs.push_back(n);
        // fibers: positions where the group, the row or an outer index changes
        if (inner >= 0) {
            uint64_t fm[BLCO_MAX_MODES];
            int fp[BLCO_MAX_MODES], nf = 0;
            fm[nf] = mk;
            fp[nf++] = ps;
            for (int j = 0; j < o.nouter; ++j) {
                fm[nf] = d->masks[o.outer[j]];
                fp[nf++] = d->pos[o.outer[j]];
            }
            for (const SGroup& g : groups) {
                const uint64_t lo = g.start, hi = lo + g.nnz;
                uint64_t f = 1;
                for (uint64_t i = lo + 1; i < hi; ++i) {
                    bool first = false;
                    for (int k = 0; k < nf && !first; ++k)
                        first = ((sc[i] >> fp[k]) & fm[k]) != ((sc[i - 1] >> fp[k]) & fm[k]);
                    f += first;
                }
                fibers += f;
            }
        }
        auto release = [&]() {
            if (dc) (void)hipFree(dc);
            if (dv) (void)hipFree(dv);
        };
        if (hipMalloc(&dc, n * 8) != hipSuccess || hipMalloc(&dv, n * vsz) != hipSuccess) {
            release();
            return fail(BLCO_ENOMEM, "no device memory for the row-ordered copy of mode %d (%llu bytes)", mode,
                        (unsigned long long)(n * (8 + vsz)));
        }
        if (hipMemcpy(dc, sc.data(), n * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dv, sv.data(), n * vsz, hipMemcpyHostToDevice) != hipSuccess) {
            release();
            return fail(BLCO_EHIP, "upload of the row-ordered copy failed");
        }
    }
    progress("row-ordered copy of mode %d: sorted (%d tiles of mode %d, inner mode %d)", mode, ntiles, split,
             inner);
    const bool use_fibers = inner >= 0 && ((fenv && fenv[0] == '1') || (double)n >= FIBER_MIN_LEN * (double)fibers);
    if (inner >= 0)
        progress("row-ordered copy of mode %d: %.2f nonzeros per fiber%s", mode, (double)n / (double)fibers,
                 use_fibers ? "" : " (fiber kernel off)");
    d->d_scoords[mode] = dc;
    d->d_svals[mode] = dv;
    d->bytes += n * (8 + vsz);
    d->stiles[mode] = ntiles;
    d->sinner[mode] = inner;
    d->sfiber[mode] = use_fibers;
    d->sgrp_h[mode].swap(groups);
    d->srun_h[mode].swap(runs);
    return BLCO_OK;
}

// Chunks of the row-ordered copy of a mode: within a group, a chunk closes at
// the first row boundary once it holds 64 * tpw nonzeros, and never holds more
// than twice that (a long row is cut), so most rows end inside a chunk while
// every chunk stays short enough to spread a hot row over many waves.  A tiled
// copy pads every tile's chunk list with empty chunks to one common length, a
// multiple of the waves per workgroup: the grid then has NUM_XCD equal ranges
// of workgroups and the XCD-contiguous workgroup numbering puts tile t on XCD t.
static const ChunkTab* sorted_chunk_table(blco_dev* d, int mode, int tpw)
{
    for (auto& t : d->stabs[mode])
        if (t.tpw == tpw) return &t;
    const uint64_t base = (uint64_t)WAVE * tpw, cap = 2 * base;
    const int ntiles = std::max(1, d->stiles[mode]);
    std::vector<std::vector<Chunk>> tiles(ntiles);
    const std::vector<uint64_t>& runs = d->srun_h[mode];
    size_t r = 0;
    for (const SGroup& g : d->sgrp_h[mode]) {
        std::vector<Chunk>& h = tiles[g.tile];
        const uint64_t end = g.start + g.nnz;
        uint64_t c0 = g.start;
        while (r < runs.size() && runs[r] < g.start) ++r;
        auto emit = [&](uint64_t lo, uint64_t hi) { h.push_back(Chunk{lo, (uint32_t)(hi - lo), g.block}); };
        // walk the runs of this group: [runs[r], runs[r+1]) clipped to the group
        while (c0 < end) {
            const uint64_t run_end = (r + 1 < runs.size()) ? std::min(runs[r + 1], end) : end;
            if (run_end - c0 > cap) {  // inside a long row
                emit(c0, c0 + cap);
                c0 += cap;
                continue;
            }
            if (run_end - c0 >= base || run_end == end) {
                emit(c0, run_end);
                c0 = run_end;
            }
            ++r;
            if (r >= runs.size() || runs[r] >= end) {
                if (c0 < end) emit(c0, end);
                c0 = end;
            }
        }
    }
    std::vector<Chunk> h;
    if (ntiles == 1) {
        h.swap(tiles[0]);
    } else {
        size_t most = 0;
        for (auto& tl : tiles) most = std::max(most, tl.size());
        most = (most + WAVES_PER_WG - 1) / WAVES_PER_WG * WAVES_PER_WG;
        h.reserve(most * ntiles);
        for (auto& tl : tiles) {
            h.insert(h.end(), tl.begin(), tl.end());
            h.resize(h.size() + (most - tl.size()), Chunk{0, 0, 0});
        }
    }
    ChunkTab t;
    t.tpw = tpw;
    t.n = (uint32_t)h.size();
    if (t.n) {
        if (hipMalloc(&t.d, sizeof(Chunk) * t.n) != hipSuccess) {
            fail(BLCO_ENOMEM, "device allocation of the row-ordered chunk table failed");
            return nullptr;
        }
        if (hipMemcpy(t.d, h.data(), sizeof(Chunk) * t.n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(t.d);
            fail(BLCO_EHIP, "upload of the row-ordered chunk table failed");
            return nullptr;
        }
        d->bytes += sizeof(Chunk) * t.n;
    }
    d->stabs[mode].push_back(t);
    return &d->stabs[mode].back();
}

// One kernel launch over chunks [c0, c1) whose data lives at coords/vals
// indexed from coord_base.
static int launch_range(blco_dev* d, int which, MttkrpArgs a, const ChunkTab* t, uint32_t c0,
                        uint32_t c1, int ftype, hipStream_t st)
{
    if (c1 <= c0) return BLCO_OK;
    a.chunks = t->d;
    a.chunk0 = c0;
    a.nchunks = c1 - c0;
    uint32_t grid;
    size_t lds = 0;
    if (which == W_LVL3) {
        a.chunks_per_wg = LVL3_CHUNKS_PER_WG;
        grid = (a.nchunks + LVL3_CHUNKS_PER_WG - 1) / LVL3_CHUNKS_PER_WG;
        lds = d->dims[a.mode] * (uint64_t)a.rank * (ftype == BLCO_F64 ? 8 : 4);
    } else {
        grid = (a.nchunks + WAVES_PER_WG - 1) / WAVES_PER_WG;
    }
    return launch_mttkrp(which, a, ftype, grid, lds, st);
}

static size_t plan_rows(blco_dev* d, int mode, int rank, int ftype, bool fib, MttkrpArgs* a);

// The best window of h consecutive rows of mode m (most nonzeros); its
// start, and its nonzero count in *covered.
static uint64_t best_window(blco_dev* d, int m, uint64_t h, uint64_t* covered)
{
    std::vector<uint64_t>& pre = d->idx_prefix[m];
    *covered = 0;
    if (pre.empty()) {
        std::vector<uint64_t> cnt;
        if (dev_split_counts(d, m, &cnt)) return 0;  // no window when the counts fail
        pre.assign(cnt.size() + 1, 0);
        for (size_t i = 0; i < cnt.size(); ++i) pre[i + 1] = pre[i] + cnt[i];
    }
    const uint64_t n = pre.size() - 1;
    if (h > n) h = n;
    uint64_t best = 0, lo = 0;
    for (uint64_t i = 0; i + h <= n; ++i) {
        const uint64_t c = pre[i + h] - pre[i];
        if (c > best) { best = c; lo = i; }
    }
    *covered = best;
    return lo;
}

// Chunks per wave and the LDS windows of hot rows for the row-ordered
// kernels (k_mttkrp_rows), into a; returns the dynamic LDS bytes.  Cached
// per (mode, rank, factor type, form, chunk count).
size_t rows_kernel_plan(blco_dev* d, int mode, int rank, int ftype, bool fib, MttkrpArgs* a)
{
    for (const RowsPlan& p : d->rows_plans[mode])
        if (p.rank == rank && p.ftype == ftype && p.fib == fib && p.nchunks == a->nchunks) {
            a->cpw = p.cpw;
            a->cpw_interleave = p.interleave;
            for (int t = 0; t < 2; ++t) {
                a->hot_lo[t] = p.hot_lo[t];
                a->hot_n[t] = p.hot_n[t];
            }
            return p.lds;
        }
    const size_t lds = plan_rows(d, mode, rank, ftype, fib, a);
    d->rows_plans[mode].push_back(RowsPlan{rank, ftype, fib, a->nchunks, a->cpw, a->cpw_interleave,
                                           {a->hot_lo[0], a->hot_lo[1]}, {a->hot_n[0], a->hot_n[1]}, lds});
    return lds;
}

// The windows cover, for the first two input modes (the inner and first
// outer mode with fibers, 2/3 and 1/3 of HOT_TABLE_BYTES' rows; else
// halves), the mode's consecutive rows with the most nonzeros; a window that
// would hold fewer than 2% of the nonzeros is left out.  A workgroup fills
// its windows once, so its waves take several chunks (as many as leave 1024
// workgroups, at least half of the cap): up to 4 consecutive ones each in
// the plain form, for rows of at most 128 bytes (FP64 rows of 256 bytes
// were slower with windows); in the fiber form up to 8, with the grid walking the
// chunks in step (chunk j of a wave is j grid widths on), and only when the
// inner window holds a fifth of the nonzeros and the copy is not split into
// XCD tiles -- the forms and counts that were faster on the 77M-nonzero
// and 1.7B-nonzero workloads (a wave per chunk, no windows, otherwise).
static size_t plan_rows(blco_dev* d, int mode, int rank, int ftype, bool fib, MttkrpArgs* a)
{
    const uint32_t n = a->nchunks;
    a->cpw = 1;
    a->cpw_interleave = 0;
    a->hot_lo[0] = a->hot_lo[1] = 0;
    a->hot_n[0] = a->hot_n[1] = 0;
    const uint64_t row_bytes = (uint64_t)rank * dtype_size(ftype);
    // up to 8 (fibers) or 4 chunks per wave while 1024 workgroups remain
    const uint32_t cap = fib ? 8 : 4, fit = n / (1024u * WAVES_PER_WG);
    const uint32_t cpw = fit < cap ? fit : cap;
    // windows need whole 16-byte pieces and at least half the chunks per wave
    if (row_bytes % 16 != 0 || row_bytes > HOT_TABLE_BYTES / 8 || cpw < cap / 2) return 0;
    if (fib && d->stiles[mode] > 1) return 0;
    if (!fib && row_bytes > 128) return 0;
    int in[2] = {-1, -1}, k = 0;
    if (fib) in[k++] = a->inner;
    for (int m = 0; m < d->nmodes && k < 2; ++m)
        if (m != mode && !(fib && m == a->inner)) in[k++] = m;
    const uint64_t total = HOT_TABLE_BYTES / row_bytes;
    const uint64_t h0 = k == 1 ? total : fib ? total * 2 / 3 : total / 2;
    const uint64_t want[2] = {h0, total - h0};
    uint64_t lo[2] = {0, 0}, h[2] = {0, 0}, covered[2] = {0, 0};
    for (int t = 0; t < k; ++t) {
        h[t] = std::min<uint64_t>(want[t], d->dims[in[t]]);
        lo[t] = best_window(d, in[t], h[t], &covered[t]);
    }
    if (fib && covered[0] * 5 < d->nnz) return 0;
    a->cpw = cpw;
    a->cpw_interleave = fib ? 1 : 0;
    size_t lds = 0;
    for (int t = 0; t < k; ++t) {
        if (h[t] == 0 || covered[t] * 50 < d->nnz) continue;
        a->hot_lo[t] = (uint32_t)lo[t];
        a->hot_n[t] = (uint32_t)h[t];
        lds += h[t] * row_bytes;
        progress("row-ordered kernel of mode %d: rows [%llu, %llu) of mode %d in LDS, %.1f%% of the nonzeros",
                 mode, (unsigned long long)lo[t], (unsigned long long)(lo[t] + h[t]), in[t],
                 100.0 * covered[t] / d->nnz);
    }
    return lds;
}

int mttkrp_prepare(blco_dev* d, int mode, int rank, int ftype, int kernel)
{
    if (!d || mode < 0 || mode >= d->nmodes || !d->nnz || d->stream_data) return BLCO_OK;
    int np = 0;
    int stpw = 0;
    if (kernel == BLCO_K_SELECT && !select_kernel(d, mode, rank, ftype, &kernel, &np, &stpw)) kernel = BLCO_K_AUTO;
    const bool chosen = kernel == BLCO_K_AUTO;
    if (chosen) kernel = auto_kernel(d, mode, rank, ftype);
    if (kernel != BLCO_K_SORTED) return BLCO_OK;
    const int rc = ensure_row_ordered(d, mode, rank, ftype);
    if (rc == BLCO_ENOMEM && chosen) {  // no room for the copy: mttkrp() runs the row window
        d->sorted_failed[mode] = true;
        return BLCO_OK;
    }
    return rc;
}

int mttkrp(blco_dev* d, int mode, int rank, int ftype, void* const* factors, void* out, int kernel,
           int nprtn, int tpw, hipStream_t st, int* used_kernel)
{
    if (!d) return fail(BLCO_EINVAL, "null tensor");
    if (mode < 0 || mode >= d->nmodes) return fail(BLCO_EINVAL, "mode %d outside [0, %d)", mode, d->nmodes);
    if (rank < 1 || rank > 4096) return fail(BLCO_EINVAL, "rank %d outside [1, 4096]", rank);
    if (ftype != BLCO_F64 && ftype != BLCO_F32 && ftype != BLCO_BF16)
        return fail(BLCO_EINVAL, "factor dtype %d", ftype);
    const int need_v = ftype == BLCO_F64 ? BLCO_F64 : BLCO_F32;
    if (d->vdtype != need_v)
        return fail(BLCO_EUNSUPPORTED, "%s factors need %s values on the device",
                    ftype == BLCO_F64 ? "f64" : "f32/bf16", need_v == BLCO_F64 ? "f64" : "f32");
    if (!factors || !out) return fail(BLCO_EINVAL, "null factors or output");
    for (int m = 0; m < d->nmodes; ++m)
        if (m != mode && !factors[m]) return fail(BLCO_EINVAL, "factor %d is null", m);
    if (d->nmodes < 2) return fail(BLCO_EUNSUPPORTED, "MTTKRP needs at least 2 modes");
    if (tpw > 64) return fail(BLCO_EINVAL, "thread_cf %d > 64", tpw);
    if (nprtn <= 0) nprtn = DEFAULT_NPRTN;
    if (nprtn > 1024) return fail(BLCO_EINVAL, "nprtn %d > 1024", nprtn);
    if (!st) st = d->stream;
    if (kernel == BLCO_K_SELECT) {
        int k = 0, np = 0, stpw = 0;
        if (select_kernel(d, mode, rank, ftype, &k, &np, &stpw)) {
            kernel = k;
            if (k == BLCO_K_LVL3) nprtn = np;
            if (tpw <= 0) tpw = stpw;  // the caller's thread_cf wins
        } else {
            kernel = BLCO_K_AUTO;
        }
    }
    const bool chosen = kernel == BLCO_K_AUTO;
    if (chosen) kernel = auto_kernel(d, mode, rank, ftype);
    int which;
    switch (kernel) {
        case BLCO_K_LVL1:
        case BLCO_K_BATCHED: which = W_LVL1; break;
        case BLCO_K_LVL2:
        case BLCO_K_BATCHED2: which = W_WINDOW; break;
        case BLCO_K_LVL3:
            which = W_LVL3;
            if (!lvl3_output_fits(d, mode, rank, ftype)) {
                // the reference warns and proceeds (src/alto_dev.cu:1866); a
                // privatised copy must fit in LDS here, so lvl1 runs instead
                which = W_WINDOW;
                kernel = BLCO_K_BATCHED2;
            }
            break;
        case BLCO_K_ATOMIC: which = W_ATOMIC; break;
        case BLCO_K_SORTED:
            which = W_SORTED;
            if (d->stream_data) {  // the copies are device-resident only
                which = W_WINDOW;
                kernel = BLCO_K_BATCHED2;
            }
            break;
        default: return fail(BLCO_EINVAL, "unknown kernel id %d", kernel);
    }
    if (tpw <= 0) tpw = which == W_SORTED ? sorted_tpw(d, rank, ftype) : DEFAULT_TPW;
    if (used_kernel) *used_kernel = kernel;
    if (d->nnz == 0) return BLCO_OK;
    if (which == W_SORTED) {
        int rc = ensure_row_ordered(d, mode, rank, ftype);
        if (rc == BLCO_ENOMEM && chosen) {  // no room for the copy: the row window
            d->sorted_failed[mode] = true;
            which = W_WINDOW;
            kernel = BLCO_K_BATCHED2;
            if (used_kernel) *used_kernel = kernel;
            tpw = DEFAULT_TPW;
        } else if (rc) {
            return rc;
        }
    }
    const ChunkTab* t = chunk_table(d, tpw);
    if (!t) return BLCO_ENOMEM;

    MttkrpArgs a;
    std::memset(&a, 0, sizeof a);
    a.bcoords = d->d_bcoords;
    for (int m = 0; m < d->nmodes; ++m) {
        a.mask[m] = d->masks[m];
        a.pos[m] = d->pos[m];
        a.f[m] = factors[m];
    }
    a.nmodes = d->nmodes;
    a.mode = mode;
    a.rank = rank;
    a.out = out;
    a.out_rows = d->dims[mode];
    a.nprtn = which == W_LVL3 ? nprtn : 1;
    const int otype = ftype == BLCO_F64 ? BLCO_F64 : BLCO_F32;
    const uint64_t cells = d->dims[mode] * (uint64_t)rank;
    if (which == W_LVL3 && nprtn > 1) {
        size_t bytes = (size_t)nprtn * cells * (otype == BLCO_F64 ? 8 : 4);
        int rc = ensure_partials(d, bytes);
        if (rc) return rc;
        a.partials = d->partials;
        rc = launch_fill(d->partials, (uint64_t)nprtn * cells, otype, 0.0, st);
        if (rc) return rc;
    }

    int rc = BLCO_OK;
    if (which == W_WINDOW) {
        const int acc = ftype == BLCO_F64 ? 8 : 4;
        const WindowPlan* wp = window_plan(d, mode, rank, acc);
        if (!wp) return BLCO_ENOMEM;
        const WindowTab* wt = wp->wt;
        const size_t need = wp->prows * (size_t)rank * acc;
        if (d->part_bytes < need) {
            if (d->part) (void)hipFree(d->part);
            d->part = nullptr;
            d->part_bytes = 0;
            HIP_TRY(hipMalloc(&d->part, need));
            d->part_bytes = need;
        }
        a.sc = wt->d_sc;
        a.sc_range = wt->d_range + (uint64_t)mode * wt->n * 2;
        a.hot_rank = wp->hot_rows ? d->d_hot_rank[mode] : nullptr;
        a.hot_rows = wp->hot_rows;
        a.lds_rows = wp->lds_rows;
        a.sc_poff = wp->d_poff;
        a.hot_poff = wp->hot_poff;
        a.part = d->part;
        const size_t lds = (size_t)(wp->hot_rows + wp->lds_rows) * rank * acc;
        bool wide = false;  // any mode index >= 2^32 needs 64-bit indices
        for (int m = 0; m < d->nmodes; ++m) wide |= d->dims[m] > 0xFFFFFFFFull;
        const int wk = wide ? W_WINDOW64 : W_WINDOW;
        const bool per_block = kernel == BLCO_K_LVL2;
        const int vsz = value_size(d->vdtype);
        if (!d->stream_data) {
            a.coords = d->d_coords;
            a.vals = d->d_vals;
            a.coord_base = 0;
            if (per_block) {
                for (uint64_t b = 0; b < d->nblocks && !rc; ++b) {
                    a.chunk0 = wt->first[b];
                    a.nchunks = wt->first[b + 1] - wt->first[b];
                    if (a.nchunks) rc = launch_mttkrp(wk, a, ftype, a.nchunks, lds, st);
                }
            } else {
                a.chunk0 = 0;
                a.nchunks = wt->n;
                if (wt->n) rc = launch_mttkrp(wk, a, ftype, wt->n, lds, st);
            }
        } else {
            // stream entry ranges of at most slot_nnz nonzeros within a block
            size_t u = 0;
            for (uint64_t b = 0; b < d->nblocks && !rc; ++b) {
                uint32_t e = wt->first[b];
                while (e < wt->first[b + 1] && !rc) {
                    uint32_t e1 = e;
                    uint64_t cnt = 0;
                    while (e1 < wt->first[b + 1] && (e1 == e || cnt + wt->h_sc[e1].len <= d->slot_nnz)) {
                        cnt += wt->h_sc[e1].len;
                        ++e1;
                    }
                    if (cnt > d->slot_nnz)
                        return fail(BLCO_EUNSUPPORTED, "a window of %llu nonzeros exceeds the %llu-nonzero slot",
                                    (unsigned long long)cnt, (unsigned long long)d->slot_nnz);
                    const uint64_t begin = wt->h_sc[e].begin;
                    a.chunk0 = e;
                    a.nchunks = e1 - e;
                    if (begin + cnt <= d->res_nnz) {  // within the resident prefix
                        a.coords = d->d_coords;
                        a.vals = d->d_vals;
                        a.coord_base = 0;
                        rc = launch_mttkrp(wk, a, ftype, a.nchunks, lds, st);
                        e = e1;
                        continue;
                    }
                    const int s = (int)(u & 1);
                    HIP_TRY(hipStreamWaitEvent(d->copy_stream, d->ev_done[s], 0));
                    HIP_TRY(hipMemcpyAsync(d->slot_coords[s], d->h_coords + begin, cnt * 8,
                                           hipMemcpyHostToDevice, d->copy_stream));
                    HIP_TRY(hipMemcpyAsync(d->slot_vals[s], (const char*)d->h_vals + begin * vsz,
                                           cnt * vsz, hipMemcpyHostToDevice, d->copy_stream));
                    HIP_TRY(hipEventRecord(d->ev_copied[s], d->copy_stream));
                    HIP_TRY(hipStreamWaitEvent(st, d->ev_copied[s], 0));
                    a.coords = d->slot_coords[s];
                    a.vals = d->slot_vals[s];
                    a.coord_base = begin;
                    rc = launch_mttkrp(wk, a, ftype, a.nchunks, lds, st);
                    HIP_TRY(hipEventRecord(d->ev_done[s], st));
                    ++u;
                    e = e1;
                }
            }
        }
        // every partial row was written by its workgroup: fold them in
        if (!rc)
            rc = launch_window_reduce(out, d->part, wp->d_rows, wp->d_ptr, wp->d_idx, wp->nred, rank,
                                      otype, st);
    } else if (which == W_SORTED) {
        a.coords = d->d_scoords[mode];
        a.vals = d->d_svals[mode];
        a.coord_base = 0;
        const ChunkTab* stab = sorted_chunk_table(d, mode, tpw);
        if (!stab) return BLCO_ENOMEM;
        // 32-bit row offsets (and byte offsets into the factors' buffer
        // resources) when every factor holds fewer than FACTOR_BYTES_MAX bytes
        bool wide = false;
        for (int m = 0; m < d->nmodes; ++m)
            wide |= d->dims[m] * (uint64_t)rank * dtype_size(ftype) >= FACTOR_BYTES_MAX;
        int kind = wide ? W_SORTED64 : W_SORTED;
        if (d->sfiber[mode]) {
            kind = wide ? W_FIBER64 : W_FIBER;
            a.inner = d->sinner[mode];
        }
        if (kind == W_SORTED || kind == W_FIBER) {
            a.chunks = stab->d;
            a.chunk0 = 0;
            a.nchunks = stab->n;
            const size_t lds = rows_kernel_plan(d, mode, rank, ftype, kind == W_FIBER, &a);
            // the form with LDS windows only where the plan holds some
            if (lds) kind = kind == W_FIBER ? W_FIBER_LDS : W_SORTED_LDS;
            const uint32_t per_wg = WAVES_PER_WG * a.cpw;
            if (stab->n) rc = launch_mttkrp(kind, a, ftype, (stab->n + per_wg - 1) / per_wg, lds, st);
        } else {
            rc = launch_range(d, kind, a, stab, 0, stab->n, ftype, st);
        }
    } else if (!d->stream_data) {
        a.coords = d->d_coords;
        a.vals = d->d_vals;
        a.coord_base = 0;
        if (kernel == BLCO_K_LVL1) {
            for (uint64_t b = 0; b < d->nblocks && !rc; ++b)
                rc = launch_range(d, which, a, t, t->first[b], t->first[b + 1], ftype, st);
        } else {
            rc = launch_range(d, which, a, t, 0, t->n, ftype, st);
        }
    } else {
        // double-buffered streaming: the k-th unit beyond the resident
        // prefix goes through slot k % 2
        const int vsz = value_size(d->vdtype);
        size_t k = 0;
        for (size_t u = 0; u < t->units.size() && !rc; ++u) {
            const StreamUnit& su = t->units[u];
            if (su.begin + su.nnz <= d->res_nnz) {
                a.coords = d->d_coords;
                a.vals = d->d_vals;
                a.coord_base = 0;
                rc = launch_range(d, which, a, t, su.c0, su.c1, ftype, st);
                continue;
            }
            const int s = (int)(k++ & 1);
            HIP_TRY(hipStreamWaitEvent(d->copy_stream, d->ev_done[s], 0));
            HIP_TRY(hipMemcpyAsync(d->slot_coords[s], d->h_coords + su.begin, su.nnz * 8,
                                   hipMemcpyHostToDevice, d->copy_stream));
            HIP_TRY(hipMemcpyAsync(d->slot_vals[s], (const char*)d->h_vals + su.begin * vsz,
                                   su.nnz * vsz, hipMemcpyHostToDevice, d->copy_stream));
            HIP_TRY(hipEventRecord(d->ev_copied[s], d->copy_stream));
            HIP_TRY(hipStreamWaitEvent(st, d->ev_copied[s], 0));
            a.coords = d->slot_coords[s];
            a.vals = d->slot_vals[s];
            a.coord_base = su.begin;
            rc = launch_range(d, which, a, t, su.c0, su.c1, ftype, st);
            HIP_TRY(hipEventRecord(d->ev_done[s], st));
        }
    }
    if (rc) return rc;
    if (which == W_LVL3 && nprtn > 1)
        rc = launch_reduce_partials(out, d->partials, cells, nprtn, otype, st);
    return rc;
}

static bool have_kfd() { return access("/dev/kfd", F_OK) == 0; }

}  // namespace blco

// ------------------------------------------------------------------ C ABI

extern "C" {

int blco_malloc(void** ptr, uint64_t bytes)
{
    if (!ptr) return fail(BLCO_EINVAL, "null output");
    *ptr = nullptr;
    if (!have_kfd()) return fail(BLCO_ENODEV, "no GPU (/dev/kfd absent)");
    if (bytes == 0) return BLCO_OK;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        *ptr = nullptr;
        return fail(BLCO_ENOMEM, "device allocation of %llu bytes failed", (unsigned long long)bytes);
    }
    return BLCO_OK;
}

int blco_free(void* ptr)
{
    if (ptr) HIP_TRY(hipFree(ptr));
    return BLCO_OK;
}

int blco_memcpy_htod(void* dst, const void* src, uint64_t bytes)
{
    if (bytes == 0) return BLCO_OK;
    if (!dst || !src) return fail(BLCO_EINVAL, "null pointer");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BLCO_OK;
}

int blco_memcpy_dtoh(void* dst, const void* src, uint64_t bytes)
{
    if (bytes == 0) return BLCO_OK;
    if (!dst || !src) return fail(BLCO_EINVAL, "null pointer");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BLCO_OK;
}

int blco_memcpy_dtod(void* dst, const void* src, uint64_t bytes)
{
    if (bytes == 0) return BLCO_OK;
    if (!dst || !src) return fail(BLCO_EINVAL, "null pointer");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(hipDeviceSynchronize());
    return BLCO_OK;
}

int blco_memset_zero(void* ptr, uint64_t bytes)
{
    if (bytes == 0) return BLCO_OK;
    if (!ptr) return fail(BLCO_EINVAL, "null pointer");
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(ptr, 0, bytes));
    HIP_TRY(hipDeviceSynchronize());
    return BLCO_OK;
}

int blco_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return BLCO_OK;
}

int blco_device_count(void)
{
    if (!have_kfd()) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

namespace blco {

// Device tensor from host arrays (uploaded), or from device arrays that it
// takes over (adopt_coords / adopt_vals: resident only; `coords` is then a
// host copy of the device words for the host-side tables and `vals` is
// unused).  Adopted arrays belong to the tensor (freed by blco_dev_free),
// and are freed here on failure.
int create_device_tensor(int nmodes, const uint64_t* dims, const uint64_t* masks, const int* pos,
                            uint64_t nnz, const uint64_t* coords, const double* vals,
                            uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz,
                            const uint64_t* block_coords, int value_dtype, int stream_data,
                            uint64_t slot_nnz, uint64_t device_budget, uint64_t* adopt_coords, void* adopt_vals,
                            blco_dev** out)
{
    const bool adopt = adopt_coords != nullptr;
    auto drop_adopted = [&](int rc) {
        if (adopt_coords) (void)hipFree(adopt_coords);
        if (adopt_vals) (void)hipFree(adopt_vals);
        return rc;
    };
    if (!out) return drop_adopted(fail(BLCO_EINVAL, "null output"));
    *out = nullptr;
    if (nmodes < 1 || nmodes > BLCO_MAX_MODES) return drop_adopted(fail(BLCO_EINVAL, "nmodes %d", nmodes));
    if (!dims || !masks || !pos) return drop_adopted(fail(BLCO_EINVAL, "null layout"));
    if (adopt && (stream_data || !adopt_vals))
        return drop_adopted(fail(BLCO_EINVAL, "adopted device arrays must be resident"));
    if (nnz && (!coords || (!vals && !adopt) || !block_start || !block_nnz || !block_coords || nblocks == 0))
        return drop_adopted(fail(BLCO_EINVAL, "null tensor arrays"));
    if (value_dtype != BLCO_F64 && value_dtype != BLCO_F32)
        return drop_adopted(fail(BLCO_EINVAL, "value dtype %d", value_dtype));
    uint64_t sum = 0;
    for (uint64_t b = 0; b < nblocks; ++b) {
        if (block_start[b] != sum) return drop_adopted(fail(BLCO_EINVAL, "blocks are not contiguous at %llu", (unsigned long long)b));
        sum += block_nnz[b];
    }
    if (sum != nnz) return drop_adopted(fail(BLCO_EINVAL, "block sizes sum to %llu, nnz is %llu", (unsigned long long)sum, (unsigned long long)nnz));
    for (int m = 0; m < nmodes; ++m) {
        if (dims[m] == 0) return drop_adopted(fail(BLCO_EINVAL, "mode %d has length 0", m));
        if (masks[m] && (pos[m] < 0 || pos[m] > 63)) return drop_adopted(fail(BLCO_EINVAL, "mode %d position %d", m, pos[m]));
    }
    // every decoded index must lie inside its mode (kernels gather factor
    // rows and scatter output rows with these indices unchecked); the same
    // passes find the hot rows and each fine chunk's per-mode row range
    progress("device tensor: checking %llu nonzeros in %llu blocks", (unsigned long long)nnz,
             (unsigned long long)nblocks);
    TensorTables tt;
    {
        int rc = analyze_tensor(nmodes, dims, masks, pos, nnz, coords, nblocks, block_start, block_nnz,
                                block_coords, FINE_CHUNK, HOT_EXCLUDE, &tt);
        if (rc) return drop_adopted(rc);
    }
    std::vector<uint64_t>& fbeg = tt.fbeg;
    std::vector<uint32_t>& flen = tt.flen;
    std::vector<uint32_t>& fblk = tt.fblk;
    std::vector<uint64_t>& flo = tt.flo;
    std::vector<uint64_t>& fhi = tt.fhi;
    std::vector<std::vector<uint32_t>>& hot = tt.hot;
    std::vector<std::vector<uint64_t>>& hot_cnt = tt.hot_cnt;
    if (!have_kfd()) return drop_adopted(fail(BLCO_ENODEV, "no GPU (/dev/kfd absent)"));
    const int vsz = value_size(value_dtype);
    const uint64_t per_nz = 8 + (uint64_t)vsz;
    // a device budget keeps a tensor that fits resident and streams the
    // rest of one that does not; the nonzeros it cannot hold beside the two
    // slots pass through them
    uint64_t res_nnz = 0;
    if (device_budget && !adopt) stream_data = nnz * per_nz > device_budget;
    if (stream_data && slot_nnz == 0) slot_nnz = 1ull << 24;  // --max-block-size default
    // a slot holds at least one chunk at the largest tiles-per-wave (64 x 64)
    if (stream_data && slot_nnz < 4096) slot_nnz = 4096;
    if (stream_data && device_budget) {
        const uint64_t slots = 2 * std::min<uint64_t>(slot_nnz, nnz ? nnz : 1) * per_nz;
        if (device_budget < slots)
            return drop_adopted(fail(BLCO_EINVAL, "device budget of %llu bytes is below the two streaming slots "
                                     "(%llu bytes at %llu nonzeros each)", (unsigned long long)device_budget,
                                     (unsigned long long)slots, (unsigned long long)slot_nnz));
        res_nnz = std::min<uint64_t>(nnz, (device_budget - slots) / per_nz);
    }

    blco_dev* d = new (std::nothrow) blco_dev();
    if (!d) return drop_adopted(fail(BLCO_ENOMEM, "out of host memory"));
    // adopted arrays belong to the tensor from here on (blco_dev_free)
    d->d_coords = adopt_coords;
    d->d_vals = adopt_vals;
    d->nmodes = nmodes;
    for (int m = 0; m < nmodes; ++m) {
        d->dims[m] = dims[m];
        d->masks[m] = masks[m];
        d->pos[m] = masks[m] ? pos[m] : 0;  // a 0-bit mode never shifts
    }
    d->nnz = nnz;
    d->vdtype = value_dtype;
    d->stream_data = stream_data ? 1 : 0;
    d->slot_nnz = slot_nnz;
    d->res_nnz = stream_data ? res_nnz : nnz;
    d->budget = device_budget;
    d->nblocks = nblocks;
    d->bstart.assign(block_start, block_start + nblocks);
    d->bnnz.assign(block_nnz, block_nnz + nblocks);
    progress("device tensor: selector features");
    compute_static_features(nmodes, d->dims, d->masks, d->pos, nnz, coords, nblocks, block_start, block_nnz,
                            block_coords, BLCO_SEL_NFEAT, &d->sfeat[0][0]);
    d->fine_begin.swap(fbeg);
    d->fine_len.swap(flen);
    d->fine_block.swap(fblk);
    d->fine_lo.swap(flo);
    d->fine_hi.swap(fhi);
    for (int m = 0; m < nmodes; ++m) {
        d->hot_list_h[m] = hot[m];
        d->hot_cnt_h[m] = hot_cnt[m];
    }
    auto cleanup = [&](int rc) { blco_dev_free(d); return rc; };
    if (hipGetDevice(&d->device) != hipSuccess) return cleanup(fail(BLCO_EHIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(BLCO_EHIP, "stream creation failed"));
    for (int m = 0; m < nmodes; ++m) {
        const std::vector<uint32_t>& hl = d->hot_list_h[m];
        if (hl.empty()) continue;
        std::vector<int32_t> rk(dims[m], -1);
        for (size_t h = 0; h < hl.size(); ++h) rk[hl[h]] = (int32_t)h;
        if (hipMalloc(&d->d_hot_list[m], sizeof(uint32_t) * hl.size()) != hipSuccess ||
            hipMalloc(&d->d_hot_rank[m], sizeof(int32_t) * dims[m]) != hipSuccess)
            return cleanup(fail(BLCO_ENOMEM, "device allocation failed (hot rows)"));
        if (hipMemcpy(d->d_hot_list[m], hl.data(), sizeof(uint32_t) * hl.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d->d_hot_rank[m], rk.data(), sizeof(int32_t) * dims[m], hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(BLCO_EHIP, "upload of the hot-row tables failed"));
        d->bytes += sizeof(uint32_t) * hl.size() + sizeof(int32_t) * dims[m];
    }
    if (nblocks) {
        if (hipMalloc(&d->d_bcoords, sizeof(uint64_t) * nblocks * nmodes) != hipSuccess)
            return cleanup(fail(BLCO_ENOMEM, "device allocation failed (block coordinates)"));
        if (hipMemcpy(d->d_bcoords, block_coords, sizeof(uint64_t) * nblocks * nmodes,
                      hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(BLCO_EHIP, "upload of block coordinates failed"));
        d->bytes += sizeof(uint64_t) * nblocks * nmodes;
    }
    // values converted to the device dtype through a bounded host staging buffer
    auto convert = [&](uint64_t off, uint64_t cnt, void* dst) {
        if (value_dtype == BLCO_F64) std::memcpy(dst, vals + off, cnt * 8);
        else {
            float* f = (float*)dst;
            for (uint64_t i = 0; i < cnt; ++i) f[i] = (float)vals[off + i];
        }
    };
    if (adopt) {
        d->bytes += nnz * (8 + vsz);
    } else if (!d->stream_data) {
        if (nnz) {
            if (hipMalloc(&d->d_coords, nnz * 8) != hipSuccess || hipMalloc(&d->d_vals, nnz * vsz) != hipSuccess)
                return cleanup(fail(BLCO_ENOMEM, "device allocation of %llu nonzeros failed",
                                    (unsigned long long)nnz));
            d->bytes += nnz * (8 + vsz);
            if (hipMemcpy(d->d_coords, coords, nnz * 8, hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(fail(BLCO_EHIP, "coordinate upload failed"));
            const uint64_t STAGE = 1ull << 24;
            std::vector<char> buf;
            try { buf.resize(std::min(nnz, STAGE) * vsz); } catch (...) { return cleanup(fail(BLCO_ENOMEM, "staging")); }
            for (uint64_t off = 0; off < nnz; off += STAGE) {
                uint64_t cnt = std::min(STAGE, nnz - off);
                convert(off, cnt, buf.data());
                if (hipMemcpy((char*)d->d_vals + off * vsz, buf.data(), cnt * vsz, hipMemcpyHostToDevice) != hipSuccess)
                    return cleanup(fail(BLCO_EHIP, "value upload failed"));
            }
        }
    } else {
        if (nnz) {
            if (hipHostMalloc((void**)&d->h_coords, nnz * 8, hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&d->h_vals, nnz * vsz, hipHostMallocDefault) != hipSuccess)
                return cleanup(fail(BLCO_ENOMEM, "pinned host allocation of %llu nonzeros failed",
                                    (unsigned long long)nnz));
            std::memcpy(d->h_coords, coords, nnz * 8);
            convert(0, nnz, d->h_vals);
        }
        if (res_nnz) {  // the resident prefix, from the pinned copy
            if (hipMalloc(&d->d_coords, res_nnz * 8) != hipSuccess || hipMalloc(&d->d_vals, res_nnz * vsz) != hipSuccess)
                return cleanup(fail(BLCO_ENOMEM, "device allocation of %llu resident nonzeros failed",
                                    (unsigned long long)res_nnz));
            d->bytes += res_nnz * per_nz;
            if (hipMemcpy(d->d_coords, d->h_coords, res_nnz * 8, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d->d_vals, d->h_vals, res_nnz * vsz, hipMemcpyHostToDevice) != hipSuccess)
                return cleanup(fail(BLCO_EHIP, "upload of the resident nonzeros failed"));
        }
        uint64_t cap = std::min<uint64_t>(slot_nnz, nnz ? nnz : 1);
        for (int s = 0; s < 2; ++s) {
            if (hipMalloc(&d->slot_coords[s], cap * 8) != hipSuccess || hipMalloc(&d->slot_vals[s], cap * vsz) != hipSuccess)
                return cleanup(fail(BLCO_ENOMEM, "device slot allocation failed"));
            d->bytes += cap * (8 + vsz);
            if (hipEventCreateWithFlags(&d->ev_copied[s], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&d->ev_done[s], hipEventDisableTiming) != hipSuccess)
                return cleanup(fail(BLCO_EHIP, "event creation failed"));
        }
        if (hipStreamCreateWithFlags(&d->copy_stream, hipStreamNonBlocking) != hipSuccess)
            return cleanup(fail(BLCO_EHIP, "copy stream creation failed"));
    }
    // every block must fit a slot when streaming (chunks never straddle blocks)
    *out = d;
    return BLCO_OK;
}

}  // namespace blco

extern "C" {

int blco_dev_create_arrays(int nmodes, const uint64_t* dims, const uint64_t* masks, const int* pos,
                           uint64_t nnz, const uint64_t* coords, const double* vals,
                           uint64_t nblocks, const uint64_t* block_start, const uint64_t* block_nnz,
                           const uint64_t* block_coords, int value_dtype, int stream_data,
                           uint64_t slot_nnz, blco_dev** out)
{
    return create_device_tensor(nmodes, dims, masks, pos, nnz, coords, vals, nblocks, block_start, block_nnz,
                                block_coords, value_dtype, stream_data, slot_nnz, 0, nullptr, nullptr, out);
}

int blco_dev_create_budget(const blco_host* h, int value_dtype, uint64_t device_budget, uint64_t slot_nnz,
                           blco_dev** out)
{
    if (!h) return fail(BLCO_EINVAL, "null host tensor");
    return create_device_tensor(h->L.nmodes, h->L.dims, h->L.masks, h->L.pos, h->nnz, h->coords.data(),
                                h->vals.data(), h->block_nnz.size(), h->block_start.data(), h->block_nnz.data(),
                                h->block_coords.data(), value_dtype, 0, slot_nnz, device_budget, nullptr, nullptr,
                                out);
}

int blco_dev_rows_plan(const blco_dev* d, int mode, int rank, int factor_dtype, uint32_t* chunks_per_wave,
                       int* interleaved, uint64_t* lds_rows)
{
    if (!d || mode < 0 || mode >= d->nmodes || !chunks_per_wave || !interleaved || !lds_rows)
        return fail(BLCO_EINVAL, "bad plan query");
    for (const RowsPlan& p : d->rows_plans[mode])
        if (p.rank == rank && p.ftype == factor_dtype) {
            *chunks_per_wave = p.cpw;
            *interleaved = (int)p.interleave;
            *lds_rows = (uint64_t)p.hot_n[0] + p.hot_n[1];
            return BLCO_OK;
        }
    return fail(BLCO_EINVAL, "mode %d has run no row-ordered MTTKRP at rank %d", mode, rank);
}

int blco_dev_residency(const blco_dev* d, uint64_t* resident_nnz, uint64_t* streamed_nnz, uint64_t* device_bytes)
{
    if (!d) return fail(BLCO_EINVAL, "null tensor");
    if (resident_nnz) *resident_nnz = d->res_nnz;
    if (streamed_nnz) *streamed_nnz = d->nnz - d->res_nnz;
    if (device_bytes) *device_bytes = d->bytes;
    return BLCO_OK;
}

int blco_dev_create(const blco_host* h, int value_dtype, int stream_data, uint64_t slot_nnz,
                    blco_dev** out)
{
    if (!h) return fail(BLCO_EINVAL, "null host tensor");
    return blco_dev_create_arrays(h->L.nmodes, h->L.dims, h->L.masks, h->L.pos, h->nnz,
                                  h->coords.data(), h->vals.data(), h->block_nnz.size(),
                                  h->block_start.data(), h->block_nnz.data(),
                                  h->block_coords.data(), value_dtype, stream_data, slot_nnz, out);
}

int blco_dev_row_copy(blco_dev* d, int mode, int rank, int factor_dtype, uint64_t* info, uint64_t* coords,
                      void* vals, uint64_t* runs, uint64_t* groups)
{
    if (!d || !info || mode < 0 || mode >= d->nmodes || rank < 1) return fail(BLCO_EINVAL, "bad row-copy arguments");
    if (d->stream_data) return fail(BLCO_EINVAL, "a streamed tensor has no row-ordered copies");
    if (factor_dtype < BLCO_F64 || factor_dtype > BLCO_BF16) return fail(BLCO_EINVAL, "factor dtype %d", factor_dtype);
    int rc = ensure_row_ordered(d, mode, rank, factor_dtype);
    if (rc) return rc;
    const uint64_t n = d->nnz;
    info[0] = (uint64_t)d->stiles[mode];
    info[1] = (uint64_t)(int64_t)d->sinner[mode];
    info[2] = d->sfiber[mode] ? 1 : 0;
    info[3] = d->sgrp_h[mode].size();
    info[4] = d->srun_h[mode].size();
    if (n && coords) HIP_TRY(hipMemcpy(coords, d->d_scoords[mode], n * 8, hipMemcpyDeviceToHost));
    if (n && vals) HIP_TRY(hipMemcpy(vals, d->d_svals[mode], n * value_size(d->vdtype), hipMemcpyDeviceToHost));
    if (runs) std::memcpy(runs, d->srun_h[mode].data(), d->srun_h[mode].size() * 8);
    if (groups)
        for (size_t g = 0; g < d->sgrp_h[mode].size(); ++g) {
            const SGroup& x = d->sgrp_h[mode][g];
            groups[4 * g] = x.start;
            groups[4 * g + 1] = x.nnz;
            groups[4 * g + 2] = x.block;
            groups[4 * g + 3] = x.tile;
        }
    return BLCO_OK;
}

int blco_dev_info(const blco_dev* d, int* nmodes, uint64_t* dims, uint64_t* nnz, uint64_t* nblocks)
{
    if (!d) return fail(BLCO_EINVAL, "null device tensor");
    if (nmodes) *nmodes = d->nmodes;
    if (dims)
        for (int m = 0; m < d->nmodes; ++m) dims[m] = d->dims[m];
    if (nnz) *nnz = d->nnz;
    if (nblocks) *nblocks = d->nblocks;
    return BLCO_OK;
}

int blco_dev_to_host(const blco_dev* d, blco_host** out)
{
    if (!d || !out) return fail(BLCO_EINVAL, "null argument");
    *out = nullptr;
    if (d->stream_data) return fail(BLCO_EINVAL, "a streamed tensor is already in host memory");
    blco_host* h = new (std::nothrow) blco_host();
    if (!h) return fail(BLCO_ENOMEM, "out of host memory");
    int rc = make_layout(d->nmodes, d->dims, &h->L);
    if (rc) {
        delete h;
        return rc;
    }
    const uint64_t n = d->nnz, nb = d->nblocks;
    try {
        h->coords.resize(n);
        h->vals.resize(n);
        h->block_coords.resize(nb * d->nmodes);
    } catch (const std::bad_alloc&) {
        delete h;
        return fail(BLCO_ENOMEM, "host copy of %llu nonzeros", (unsigned long long)n);
    }
    h->nnz = n;
    h->block_start = d->bstart;
    h->block_nnz = d->bnnz;
    for (uint64_t b = 0; b < nb; ++b) h->max_block_nnz = std::max(h->max_block_nnz, d->bnnz[b]);
    bool ok = (n == 0 || hipMemcpy(h->coords.data(), d->d_coords, n * 8, hipMemcpyDeviceToHost) == hipSuccess) &&
              (nb == 0 || hipMemcpy(h->block_coords.data(), d->d_bcoords, nb * d->nmodes * 8,
                                    hipMemcpyDeviceToHost) == hipSuccess);
    if (ok && n) {
        if (d->vdtype == BLCO_F64) {
            ok = hipMemcpy(h->vals.data(), d->d_vals, n * 8, hipMemcpyDeviceToHost) == hipSuccess;
        } else {
            std::vector<float> f(n);
            ok = hipMemcpy(f.data(), d->d_vals, n * 4, hipMemcpyDeviceToHost) == hipSuccess;
            for (uint64_t i = 0; ok && i < n; ++i) h->vals[i] = f[i];
        }
    }
    if (!ok) {
        delete h;
        return fail(BLCO_EHIP, "download of the device tensor failed");
    }
    *out = h;
    return BLCO_OK;
}

__global__ void k_sample_gather(uint64_t cnt, uint64_t step, int vsz, const uint64_t* __restrict__ coords,
                                const void* __restrict__ vals, uint64_t* __restrict__ gc, void* __restrict__ gv)
{
    for (uint64_t k = (uint64_t)blockIdx.x * WG + threadIdx.x; k < cnt; k += (uint64_t)gridDim.x * WG) {
        const uint64_t i = k * step;
        gc[k] = coords[i];
        if (vsz == 8) static_cast<double*>(gv)[k] = static_cast<const double*>(vals)[i];
        else static_cast<float*>(gv)[k] = static_cast<const float*>(vals)[i];
    }
}

int blco_dev_sample(const blco_dev* d, uint64_t step, blco_coo** out)
{
    if (!d || !out || step == 0) return fail(BLCO_EINVAL, "bad sample arguments");
    *out = nullptr;
    if (d->stream_data) return fail(BLCO_EINVAL, "sampling takes a resident tensor");
    const uint64_t n = d->nnz, cnt = (n + step - 1) / step;
    const int N = d->nmodes;
    std::vector<uint64_t> c(cnt), bc(d->nblocks * N);
    std::vector<double> v(cnt);
    const size_t vsz = value_size(d->vdtype);
    std::vector<char> vb(cnt * vsz);
    if (cnt) {
        // gathered on the device into contiguous buffers, then one copy each
        uint64_t* gc = nullptr;
        void* gv = nullptr;
        if (hipMalloc(&gc, cnt * 8) != hipSuccess || hipMalloc(&gv, cnt * vsz) != hipSuccess) {
            if (gc) (void)hipFree(gc);
            return fail(BLCO_ENOMEM, "sample buffers of %llu nonzeros", (unsigned long long)cnt);
        }
        const uint32_t grid = (uint32_t)std::min<uint64_t>((cnt + WG - 1) / WG, 65536);
        hipLaunchKernelGGL(k_sample_gather, dim3(grid), dim3(WG), 0, 0, cnt, step, (int)vsz, d->d_coords,
                           d->d_vals, gc, gv);
        const bool ok = hipGetLastError() == hipSuccess &&
                        hipMemcpy(c.data(), gc, cnt * 8, hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(vb.data(), gv, cnt * vsz, hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(bc.data(), d->d_bcoords, bc.size() * 8, hipMemcpyDeviceToHost) == hipSuccess;
        (void)hipFree(gc);
        (void)hipFree(gv);
        if (!ok) return fail(BLCO_EHIP, "sample download failed");
    }
    blco_coo* t = new (std::nothrow) blco_coo();
    if (!t) return fail(BLCO_ENOMEM, "out of host memory");
    t->nmodes = N;
    for (int m = 0; m < N; ++m) t->dims[m] = d->dims[m];
    t->nnz = cnt;
    t->cidx.resize((uint64_t)N * cnt);
    t->vals.resize(cnt);
    uint64_t b = 0;
    for (uint64_t k = 0; k < cnt; ++k) {
        const uint64_t i = k * step;
        while (b + 1 < d->nblocks && d->bstart[b + 1] <= i) ++b;
        for (int m = 0; m < N; ++m)
            t->cidx[(uint64_t)m * cnt + k] = ((c[k] >> d->pos[m]) & d->masks[m]) | bc[b * N + m];
        t->vals[k] = vsz == 8 ? reinterpret_cast<const double*>(vb.data())[k]
                              : (double)reinterpret_cast<const float*>(vb.data())[k];
    }
    *out = t;
    return BLCO_OK;
}

__global__ void k_vals_to_f32(uint64_t n, const double* __restrict__ src, float* __restrict__ dst)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        dst[i] = (float)src[i];
}

__global__ void k_vals_to_f64(uint64_t n, const float* __restrict__ src, double* __restrict__ dst)
{
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG)
        dst[i] = (double)src[i];
}

int blco_dev_set_value_dtype(blco_dev* d, int value_dtype)
{
    if (!d) return fail(BLCO_EINVAL, "null device tensor");
    if (value_dtype != BLCO_F64 && value_dtype != BLCO_F32) return fail(BLCO_EINVAL, "value dtype %d", value_dtype);
    if (value_dtype == d->vdtype) return BLCO_OK;
    if (d->stream_data) return fail(BLCO_EINVAL, "the value type of a streamed tensor is fixed");
    HIP_TRY(hipStreamSynchronize(d->stream));
    const uint64_t n = d->nnz;
    const size_t old_sz = value_size(d->vdtype), new_sz = value_size(value_dtype);
    if (n) {
        void* nv = nullptr;
        if (hipMalloc(&nv, n * new_sz) != hipSuccess)
            return fail(BLCO_ENOMEM, "values of %llu nonzeros", (unsigned long long)n);
        const uint32_t grid = (uint32_t)std::min<uint64_t>((n + WG - 1) / WG, 65536);
        if (value_dtype == BLCO_F32)
            hipLaunchKernelGGL(k_vals_to_f32, dim3(grid), dim3(WG), 0, d->stream, n, (const double*)d->d_vals, (float*)nv);
        else
            hipLaunchKernelGGL(k_vals_to_f64, dim3(grid), dim3(WG), 0, d->stream, n, (const float*)d->d_vals, (double*)nv);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(d->stream) != hipSuccess) {
            (void)hipFree(nv);
            return fail(BLCO_EHIP, "value conversion failed");
        }
        (void)hipFree(d->d_vals);
        d->d_vals = nv;
        d->bytes = d->bytes - n * old_sz + n * new_sz;
    }
    // the row-ordered copies hold values of the old type: rebuilt on next use
    for (int m = 0; m < BLCO_MAX_MODES; ++m) {
        if (d->d_scoords[m]) {
            (void)hipFree(d->d_scoords[m]);
            d->d_scoords[m] = nullptr;
            d->bytes -= n * (8 + old_sz);
        }
        if (d->d_svals[m]) (void)hipFree(d->d_svals[m]);
        d->d_svals[m] = nullptr;
        for (auto& t : d->stabs[m])
            if (t.d) (void)hipFree(t.d);
        d->stabs[m].clear();
        d->sgrp_h[m].clear();
        d->srun_h[m].clear();
        d->sorted_failed[m] = false;
        d->stiles[m] = 0;
        d->sinner[m] = 0;
        d->sfiber[m] = false;
    }
    d->sel_cache.clear();
    d->vdtype = value_dtype;
    return BLCO_OK;
}

void blco_dev_free(blco_dev* d)
{
    if (!d) return;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    if (d->copy_stream) (void)hipStreamSynchronize(d->copy_stream);
    for (auto& t : d->tabs)
        if (t.d) (void)hipFree(t.d);
    for (auto& t : d->wtabs) {
        if (t.d_sc) (void)hipFree(t.d_sc);
        if (t.d_range) (void)hipFree(t.d_range);
    }
    for (int m = 0; m < BLCO_MAX_MODES; ++m) {
        if (d->d_hot_list[m]) (void)hipFree(d->d_hot_list[m]);
        if (d->d_hot_rank[m]) (void)hipFree(d->d_hot_rank[m]);
    }
    for (auto& p : d->wplans) {
        if (p.d_poff) (void)hipFree(p.d_poff);
        if (p.d_rows) (void)hipFree(p.d_rows);
        if (p.d_ptr) (void)hipFree(p.d_ptr);
        if (p.d_idx) (void)hipFree(p.d_idx);
    }
    if (d->part) (void)hipFree(d->part);
    for (int m = 0; m < BLCO_MAX_MODES; ++m) {
        if (d->d_scoords[m]) (void)hipFree(d->d_scoords[m]);
        if (d->d_svals[m]) (void)hipFree(d->d_svals[m]);
        for (auto& t : d->stabs[m])
            if (t.d) (void)hipFree(t.d);
    }
    if (d->d_bcoords) (void)hipFree(d->d_bcoords);
    if (d->d_coords) (void)hipFree(d->d_coords);
    if (d->d_vals) (void)hipFree(d->d_vals);
    if (d->partials) (void)hipFree(d->partials);
    for (int s = 0; s < 2; ++s) {
        if (d->slot_coords[s]) (void)hipFree(d->slot_coords[s]);
        if (d->slot_vals[s]) (void)hipFree(d->slot_vals[s]);
        if (d->ev_copied[s]) (void)hipEventDestroy(d->ev_copied[s]);
        if (d->ev_done[s]) (void)hipEventDestroy(d->ev_done[s]);
    }
    if (d->h_coords) (void)hipHostFree(d->h_coords);
    if (d->h_vals) (void)hipHostFree(d->h_vals);
    if (d->copy_stream) (void)hipStreamDestroy(d->copy_stream);
    if (d->stream) (void)hipStreamDestroy(d->stream);
    delete d;
}

uint64_t blco_dev_bytes(const blco_dev* d) { return d ? d->bytes : 0; }

int blco_dev_arrays(const blco_dev* d, uint64_t** coords, void** vals, int* value_dtype, uint64_t** block_coords)
{
    if (!d) return fail(BLCO_EINVAL, "null device tensor");
    const bool res = !d->stream_data;
    if (coords) *coords = res ? d->d_coords : nullptr;
    if (vals) *vals = res ? d->d_vals : nullptr;
    if (value_dtype) *value_dtype = d->vdtype;
    if (block_coords) *block_coords = res ? d->d_bcoords : nullptr;
    return BLCO_OK;
}

int blco_mttkrp(blco_dev* d, int mode, int rank, int factor_dtype, void* const* factors, void* out,
                int kernel, int nprtn, int thread_cf, void* stream)
{
    return mttkrp(d, mode, rank, factor_dtype, factors, out, kernel, nprtn, thread_cf,
                  (hipStream_t)stream, nullptr);
}

int blco_mttkrp_time(blco_dev* d, int mode, int rank, int factor_dtype, void* const* factors, void* out,
                     int kernel, int nprtn, int thread_cf, int reps, float* ms, int* used_kernel)
{
    if (!ms || reps < 1) return fail(BLCO_EINVAL, "bad timing arguments");
    // one untimed call builds the tables and the scratch
    int rc = mttkrp(d, mode, rank, factor_dtype, factors, out, kernel, nprtn, thread_cf, nullptr, used_kernel);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(BLCO_EHIP, "event creation failed");
    }
    hipError_t e = hipEventRecord(e0, d->stream);
    for (int r = 0; r < reps && !rc && e == hipSuccess; ++r)
        rc = mttkrp(d, mode, rank, factor_dtype, factors, out, kernel, nprtn, thread_cf, nullptr, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, d->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess && !rc) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    HIP_TRY(e);
    *ms = t / (float)reps;
    return BLCO_OK;
}

int blco_auto_kernel(const blco_dev* d, int mode, int rank)
{
    if (!d || mode < 0 || mode >= d->nmodes) return fail(BLCO_EINVAL, "bad mode");
    return auto_kernel(d, mode, rank, d->vdtype);
}

// ----------------------------------------------------- ALS-step primitives

int blco_gram(const void* U, int dtype, uint64_t rows, int rank, double* G, void* stream)
{
    if (!U || !G || rank < 1 || rank > 64) return fail(BLCO_EINVAL, "bad Gram arguments");
    hipStream_t st = (hipStream_t)stream;
    int nparts;
    uint64_t rpw;
    gram_parts(rows, &nparts, &rpw);
    double* part = nullptr;
    HIP_TRY(hipMalloc(&part, sizeof(double) * (size_t)nparts * (rank * rank + rank)));
    int rc = launch_gram(U, dtype, rows, rank, nullptr, part, G, nullptr, st);
    hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(part);
    if (rc) return rc;
    HIP_TRY(e);
    return BLCO_OK;
}

int blco_hadamard(double* x, const double* y, uint64_t n, void* stream)
{
    if (!x || !y) return fail(BLCO_EINVAL, "null vector");
    return launch_hadamard(x, y, n, (hipStream_t)stream);
}

int blco_value_fill(double* x, uint64_t n, double val, void* stream)
{
    if (!x) return fail(BLCO_EINVAL, "null vector");
    return launch_fill(x, n, BLCO_F64, val, (hipStream_t)stream);
}

int blco_vecsum(double* x, uint64_t n, void* stream)
{
    if (!x) return fail(BLCO_EINVAL, "null vector");
    return launch_vecsum(x, n, (hipStream_t)stream);
}

int blco_solve(double* const* grams, int nmodes, int mode, int rank, const void* M, void* U, int dtype,
               uint64_t rows, int* used_fallback, void* stream)
{
    if (!grams || !M || !U || nmodes < 2 || nmodes > BLCO_MAX_MODES || mode < 0 || mode >= nmodes)
        return fail(BLCO_EINVAL, "bad solve arguments");
    if (rank < 1 || rank > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64] in the solve", rank);
    if (dtype < BLCO_F64 || dtype > BLCO_BF16) return fail(BLCO_EINVAL, "factor dtype %d", dtype);
    hipStream_t st = (hipStream_t)stream;
    GramSet gs;
    for (int m = 0; m < nmodes; ++m) gs.g[m] = grams[m];
    double *fac = nullptr, *work = nullptr, *stats = nullptr;
    int* status = nullptr;
    uint32_t np = solve_parts(rows);
    auto release = [&]() {
        if (fac) (void)hipFree(fac);
        if (work) (void)hipFree(work);
        if (status) (void)hipFree(status);
        if (stats) (void)hipFree(stats);
    };
    if (hipMalloc(&fac, sizeof(double) * rank * rank) != hipSuccess ||
        hipMalloc(&work, sizeof(double) * (2 * rank * rank + rank)) != hipSuccess ||
        hipMalloc(&status, sizeof(int)) != hipSuccess ||
        hipMalloc(&stats, sizeof(double) * (size_t)(np ? np : 1) * rank) != hipSuccess) {
        release();
        return fail(BLCO_ENOMEM, "solve work allocation failed");
    }
    int rc = launch_factorize(gs, nmodes, mode, rank, gram_eps(BLCO_F64), fac, work, status, st);
    if (!rc) rc = launch_solve_rows(M, U, dtype, rows, rank, fac, 0, stats, st);
    int h_status = 0;
    hipError_t e = hipMemcpyAsync(&h_status, status, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    release();
    if (rc) return rc;
    HIP_TRY(e);
    if (used_fallback) *used_fallback = h_status;
    return BLCO_OK;
}

int blco_pinv(double* A, uint64_t n, int* info, void* stream)
{
    if (!A) return fail(BLCO_EINVAL, "null matrix");
    if (n < 1 || n > 64) return fail(BLCO_EUNSUPPORTED, "pseudo-inverse of order %llu outside [1, 64]",
                                     (unsigned long long)n);
    hipStream_t st = (hipStream_t)stream;
    int rc = launch_pinv_general(A, (int)n, info, st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return BLCO_OK;
}

int blco_colnorm(void* U, int dtype, uint64_t rows, int rank, int which, double* lambda, void* stream)
{
    if (!U || !lambda || rank < 1 || (which != 0 && which != 1)) return fail(BLCO_EINVAL, "bad norm arguments");
    hipStream_t st = (hipStream_t)stream;
    int np = (int)std::min<uint64_t>((rows + 1023) / 1024, 1024);
    if (np < 1) np = 1;
    double *stats = nullptr, *inv = nullptr;
    HIP_TRY(hipMalloc(&stats, sizeof(double) * (size_t)np * rank));
    HIP_TRY(hipMalloc(&inv, sizeof(double) * rank));
    int rc = launch_colstats(U, dtype, rows, rank, which, stats, np, st);
    if (!rc) rc = launch_lambda(stats, np, rank, which, lambda, inv, st);
    if (!rc) rc = launch_scale(U, dtype, rows, rank, inv, st);
    hipError_t e = hipStreamSynchronize(st);
    (void)hipFree(stats); (void)hipFree(inv);
    if (rc) return rc;
    HIP_TRY(e);
    return BLCO_OK;
}

int blco_fit(double* const* grams, int nmodes, int rank, const double* lambda, const void* U_last,
             const void* M_last, int dtype, uint64_t rows, double normX2, double* fit, void* stream)
{
    if (!grams || !lambda || !U_last || !M_last || !fit || nmodes < 1 || nmodes > BLCO_MAX_MODES ||
        rank < 1 || rank > 64)
        return fail(BLCO_EINVAL, "bad fit arguments");
    hipStream_t st = (hipStream_t)stream;
    GramSet gs;
    for (int m = 0; m < nmodes; ++m) gs.g[m] = grams[m];
    int nparts;
    uint64_t rpw;
    gram_parts(rows, &nparts, &rpw);
    double *part = nullptr, *G = nullptr, *ip = nullptr, *out = nullptr;
    HIP_TRY(hipMalloc(&part, sizeof(double) * (size_t)nparts * (rank * rank + rank)));
    HIP_TRY(hipMalloc(&G, sizeof(double) * rank * rank));
    HIP_TRY(hipMalloc(&ip, sizeof(double) * rank));
    HIP_TRY(hipMalloc(&out, sizeof(double) * 4));
    int rc = launch_gram(U_last, dtype, rows, rank, M_last, part, G, ip, st);
    if (!rc) rc = launch_fit(gs, nmodes, rank, lambda, ip, normX2, out, st);
    double h[3] = {0, 0, 0};
    hipError_t e = hipMemcpyAsync(h, out, sizeof(double) * 3, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part); (void)hipFree(G); (void)hipFree(ip); (void)hipFree(out);
    if (rc) return rc;
    HIP_TRY(e);
    *fit = h[0];
    return BLCO_OK;
}

}  // extern "C"
