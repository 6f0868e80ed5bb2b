// batch.hip -- many small tensors decomposed together.
//
// One persistent kernel per mode runs the MTTKRP of every tensor of the batch
// that has that mode: the work list holds chunks (64 x tpw nonzeros of one
// BLCO block of one tensor), each workgroup walks a contiguous range of it
// (XCD-aware), each wave one chunk at a time with the segmented tile
// reduction of the lvl1 kernel, reading its tensor's masks, positions,
// factors and output from a per-tensor descriptor.  The chunk size (tpw),
// waves per workgroup and workgroups per CU are picked per (mode, rank,
// factor type) by timing the candidates on the batch itself.
//
// The CP-ALS driver keeps one solver state per tensor (cpals.hpp) and runs a
// sweep as: per mode, clear every output (one stream per tensor), the
// persistent MTTKRP (batch stream), then every tensor's ALS step on its own
// stream; the fit of every tensor at the end.  After iteration 0 the sweep is
// captured once into a hipGraph (the per-tensor streams fork from and join
// the capturing stream through events) and replayed.  The reference
// decomposes one tensor per run (src/main.cpp:322-390); this has no
// counterpart there.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <new>
#include <vector>

#include "als.hpp"
#include "cpals.hpp"
#include "devcommon.hpp"
#include "device.hpp"
#include "mttkrp_kern.hpp"

namespace blco {

struct BatchTensor {
    const uint64_t* coords;
    const void* vals;
    const uint64_t* bcoords;
    uint64_t mask[BLCO_MAX_MODES];
    int pos[BLCO_MAX_MODES];
    const void* f[BLCO_MAX_MODES];
    void* out;
    int nmodes;
};

struct BatchChunk {
    uint64_t begin;
    uint32_t len;
    uint32_t block;
    uint32_t tensor;  // index into the launch's descriptor array
    uint32_t pad;
};

constexpr int BATCH_MAX_WAVES = 8;

namespace {

template <int L>
constexpr size_t wave_lds_bytes(size_t acc)
{
    return (size_t)BLCO_MAX_MODES * (WAVE + WAVE / L) * 8 + (WAVE + WAVE / L) * acc;
}

template <int N, typename FT, typename VT, int C, int L, bool EXACT>
__device__ __forceinline__ void batch_chunk(const MttkrpArgs& a, const BatchChunk& ch, int lane,
                                            uint64_t (*s_idx)[WAVE + WAVE / L],
                                            typename Acc<FT>::type* s_val)
{
    using AT = typename Acc<FT>::type;
    uint64_t bc[N];
    #pragma unroll
    for (int m = 0; m < N; ++m) bc[m] = a.bcoords[(uint64_t)ch.block * N + m];
    const int slot = lane / L, sub = lane % L;
    AT* out = reinterpret_cast<AT*>(a.out);
    const int R = a.rank;
    for (uint32_t t0 = 0; t0 < ch.len; t0 += WAVE) {
        const uint32_t nvalid = min((uint32_t)WAVE, ch.len - t0);
        stage_tile<N, L, VT, AT>(a, bc, ch.begin + t0, nvalid, lane, s_idx, s_val);
        for (int cb = 0; cb < (EXACT ? 1 : R); cb += C * L) {
            const int col = cb + sub * C;
            segment_tile<N, FT, C, L, EXACT>(a, nvalid, slot, col, s_idx, s_val,
                [&](uint64_t row, const AT (&acc)[C]) {
                    AT* o = out + row * (uint64_t)R + col;
                    #pragma unroll
                    for (int cc = 0; cc < C; ++cc)
                        if (EXACT || col + cc < R) atomic_add_global(o + cc, acc[cc]);
                });
        }
        wave_lds_sync();
    }
}

// blockDim.x = 64 * waves; dynamic LDS = waves * wave_lds_bytes<L>.
template <typename FT, typename VT, int C, int L, bool EXACT>
__global__ __launch_bounds__(BATCH_MAX_WAVES * WAVE) void k_mttkrp_batch(
    const BatchTensor* __restrict__ tens, const BatchChunk* __restrict__ chunks, uint32_t nchunks,
    uint32_t per_wg, int mode, int rank)
{
    using AT = typename Acc<FT>::type;
    extern __shared__ __align__(16) unsigned char s_dyn[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, waves = blockDim.x >> 6;
    unsigned char* mine = s_dyn + (size_t)w * wave_lds_bytes<L>(sizeof(AT));
    auto s_idx = reinterpret_cast<uint64_t (*)[WAVE + WAVE / L]>(mine);
    AT* s_val = reinterpret_cast<AT*>(mine + (size_t)BLCO_MAX_MODES * (WAVE + WAVE / L) * 8);
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const uint64_t c0 = (uint64_t)g * per_wg;
    const uint64_t c1 = min(c0 + per_wg, (uint64_t)nchunks);
    for (uint64_t c = c0 + w; c < c1; c += waves) {  // whole waves; no block barriers
        const BatchChunk ch = chunks[c];
        const BatchTensor& td = tens[ch.tensor];
        MttkrpArgs a;  // the fields the tile helpers read
        a.coords = td.coords;
        a.vals = td.vals;
        a.coord_base = 0;
        a.bcoords = td.bcoords;
        #pragma unroll
        for (int m = 0; m < BLCO_MAX_MODES; ++m) {
            a.mask[m] = td.mask[m];
            a.pos[m] = td.pos[m];
            a.f[m] = td.f[m];
        }
        a.nmodes = td.nmodes;
        a.mode = mode;
        a.rank = rank;
        a.out = td.out;
        switch (td.nmodes) {
            case 2: batch_chunk<2, FT, VT, C, L, EXACT>(a, ch, lane, s_idx, s_val); break;
            case 3: batch_chunk<3, FT, VT, C, L, EXACT>(a, ch, lane, s_idx, s_val); break;
            case 4: batch_chunk<4, FT, VT, C, L, EXACT>(a, ch, lane, s_idx, s_val); break;
            default: batch_chunk<5, FT, VT, C, L, EXACT>(a, ch, lane, s_idx, s_val); break;
        }
    }
}

template <typename FT, typename VT>
const void* batch_kernel(const Shape& s, size_t* per_wave)
{
    using AT = typename Acc<FT>::type;
#define BK(C_, L_, EX_)                                         \
    do {                                                        \
        *per_wave = wave_lds_bytes<L_>(sizeof(AT));             \
        return (const void*)&k_mttkrp_batch<FT, VT, C_, L_, EX_>; \
    } while (0)
    if (s.exact) {
        if (s.C == 1 && s.L == 8) BK(1, 8, true);
        if (s.C == 2 && s.L == 8) BK(2, 8, true);
        if constexpr (sizeof(FT) == 8) {
            if (s.C == 4 && s.L == 8) BK(4, 8, true);
            if (s.C == 4 && s.L == 16) BK(4, 16, true);
        } else if constexpr (sizeof(FT) == 4) {
            if (s.C == 4 && s.L == 8) BK(4, 8, true);
            if (s.C == 4 && s.L == 16) BK(4, 16, true);
        } else {
            if (s.C == 4 && s.L == 8) BK(4, 8, true);
            if (s.C == 8 && s.L == 8) BK(8, 8, true);
        }
        return nullptr;
    }
    if (s.L == 8) BK(1, 8, false);
    if (s.L == 16) BK(1, 16, false);
    if (s.L == 32) BK(1, 32, false);
    BK(1, 64, false);
#undef BK
}

struct ChunkList {
    int tpw = 0, mode = 0;
    BatchChunk* d = nullptr;
    uint32_t n = 0;
};

struct Tuned {
    int mode, rank, ftype;
    int tpw, waves, wpc;
    float ms;
};

int num_cus()
{
    static int n = 0;
    if (!n) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            n = v;
        else
            n = 256;
    }
    return n;
}

}  // namespace
}  // namespace blco

using namespace blco;

struct blco_batch {
    std::vector<blco_dev*> devs;
    int maxN = 0;
    int device = 0;
    hipStream_t st = nullptr;
    std::list<ChunkList> lists;
    std::vector<Tuned> tuned;
    BatchTensor* d_tens = nullptr;  // descriptors of blco_batch_mttkrp calls
    size_t tens_cap = 0;
};

namespace blco {
namespace {

// Tensors that have mode `mode`, in batch order (the descriptor order).
std::vector<int> members(const blco_batch* b, int mode)
{
    std::vector<int> t;
    for (size_t i = 0; i < b->devs.size(); ++i)
        if (b->devs[i]->nmodes > mode) t.push_back((int)i);
    return t;
}

const ChunkList* chunk_list(blco_batch* b, int tpw, int mode)
{
    for (auto& l : b->lists)
        if (l.tpw == tpw && l.mode == mode) return &l;
    ChunkList l;
    l.tpw = tpw;
    l.mode = mode;
    std::vector<BatchChunk> h;
    const uint64_t span = (uint64_t)WAVE * tpw;
    const std::vector<int> mem = members(b, mode);
    for (size_t k = 0; k < mem.size(); ++k) {
        const blco_dev* d = b->devs[mem[k]];
        for (uint64_t blk = 0; blk < d->nblocks; ++blk)
            for (uint64_t s0 = 0; s0 < d->bnnz[blk]; s0 += span) {
                BatchChunk c;
                c.begin = d->bstart[blk] + s0;
                c.len = (uint32_t)std::min<uint64_t>(span, d->bnnz[blk] - s0);
                c.block = (uint32_t)blk;
                c.tensor = (uint32_t)k;
                c.pad = 0;
                h.push_back(c);
            }
    }
    l.n = (uint32_t)h.size();
    if (l.n) {
        if (hipMalloc(&l.d, sizeof(BatchChunk) * l.n) != hipSuccess) {
            fail(BLCO_ENOMEM, "device allocation of the batch work list failed");
            return nullptr;
        }
        if (hipMemcpy(l.d, h.data(), sizeof(BatchChunk) * l.n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(l.d);
            fail(BLCO_EHIP, "upload of the batch work list failed");
            return nullptr;
        }
    }
    b->lists.push_back(l);
    return &b->lists.back();
}

// Descriptors of the members of `mode` (host), factors F[t * MAX_MODES + m]
// and outputs O[t] indexed by batch position.
std::vector<BatchTensor> descriptors(const blco_batch* b, int mode, void* const* F, void* const* O)
{
    std::vector<BatchTensor> v;
    for (int t : members(b, mode)) {
        const blco_dev* d = b->devs[t];
        BatchTensor x{};
        x.coords = d->d_coords;
        x.vals = d->d_vals;
        x.bcoords = d->d_bcoords;
        for (int m = 0; m < d->nmodes; ++m) {
            x.mask[m] = d->masks[m];
            x.pos[m] = d->pos[m];
            x.f[m] = F[(size_t)t * BLCO_MAX_MODES + m];
        }
        x.out = O[t];
        x.nmodes = d->nmodes;
        v.push_back(x);
    }
    return v;
}

int launch_batch(blco_batch* b, const BatchTensor* d_tens, int mode, int rank, int ftype, int tpw, int waves,
                 int wpc, hipStream_t st)
{
    const ChunkList* l = chunk_list(b, tpw, mode);
    if (!l) return BLCO_ENOMEM;
    if (!l->n) return BLCO_OK;
    const Shape s = pick_shape(rank, dtype_size(ftype));
    size_t per_wave = 0;
    const void* k = ftype == BLCO_F64   ? batch_kernel<double, double>(s, &per_wave)
                    : ftype == BLCO_F32 ? batch_kernel<float, float>(s, &per_wave)
                                        : batch_kernel<bf16, float>(s, &per_wave);
    if (!k) return fail(BLCO_EUNSUPPORTED, "no batch MTTKRP kernel for rank %d", rank);
    const uint64_t wgs_needed = (l->n + waves - 1) / waves;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(wgs_needed, (uint64_t)wpc * num_cus()));
    uint32_t per_wg = (uint32_t)((l->n + grid - 1) / grid);
    const size_t lds = per_wave * waves;
    uint32_t nchunks = l->n;
    const BatchChunk* chunks = l->d;
    void* params[] = {(void*)&d_tens, (void*)&chunks, &nchunks, &per_wg, &mode, &rank};
    if (lds > 64 * 1024) HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIP_TRY(hipLaunchKernel(k, dim3(grid), dim3(waves * WAVE), params, lds, st));
    return BLCO_OK;
}

const int TPW_SET[] = {1, 2, 4, 8};
const int WAVES_SET[] = {1, 2, 4, 8};
const int WPC_SET[] = {1, 2, 4, 8};

// Times every (tpw, waves, workgroups per CU) on the batch (outputs are
// overwritten) and keeps the fastest.
int tune(blco_batch* b, const BatchTensor* d_tens, int mode, int rank, int ftype, Tuned* out)
{
    for (const Tuned& t : b->tuned)
        if (t.mode == mode && t.rank == rank && t.ftype == ftype) {
            *out = t;
            return BLCO_OK;
        }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    Tuned best{mode, rank, ftype, 4, 4, 4, -1.f};
    int rc = BLCO_OK;
    for (int tpw : TPW_SET)
        for (int waves : WAVES_SET)
            for (int wpc : WPC_SET) {
                if (waves * wpc > 32) continue;  // beyond the waves a CU holds
                rc = launch_batch(b, d_tens, mode, rank, ftype, tpw, waves, wpc, b->st);  // warm
                if (rc) break;
                (void)hipEventRecord(e0, b->st);
                for (int r = 0; r < 2 && !rc; ++r) rc = launch_batch(b, d_tens, mode, rank, ftype, tpw, waves, wpc, b->st);
                (void)hipEventRecord(e1, b->st);
                if (hipEventSynchronize(e1) != hipSuccess) rc = fail(BLCO_EHIP, "tuning run failed");
                if (rc) break;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms /= 2.f;
                if (best.ms < 0 || ms < best.ms) best = Tuned{mode, rank, ftype, tpw, waves, wpc, ms};
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->tuned.push_back(best);
    *out = best;
    return BLCO_OK;
}

int check_call(const blco_batch* b, int mode, int rank, int ftype)
{
    if (!b) return fail(BLCO_EINVAL, "null batch");
    if (mode < 0 || mode >= b->maxN) return fail(BLCO_EINVAL, "mode %d: no tensor of the batch has it", mode);
    if (rank < 1 || rank > 64) return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, 64]", rank);
    if (ftype < BLCO_F64 || ftype > BLCO_BF16) return fail(BLCO_EINVAL, "factor dtype %d", ftype);
    const int need_v = ftype == BLCO_F64 ? BLCO_F64 : BLCO_F32;
    for (const blco_dev* d : b->devs)
        if (d->vdtype != need_v) return fail(BLCO_EINVAL, "factor dtype %d needs %s values in every tensor",
                                             ftype, need_v == BLCO_F64 ? "f64" : "f32");
    return BLCO_OK;
}

int upload_descriptors(blco_batch* b, const std::vector<BatchTensor>& v)
{
    if (b->tens_cap < v.size()) {
        if (b->d_tens) (void)hipFree(b->d_tens);
        b->d_tens = nullptr;
        b->tens_cap = 0;
        HIP_TRY(hipMalloc(&b->d_tens, sizeof(BatchTensor) * v.size()));
        b->tens_cap = v.size();
    }
    if (!v.empty()) HIP_TRY(hipMemcpy(b->d_tens, v.data(), sizeof(BatchTensor) * v.size(), hipMemcpyHostToDevice));
    return BLCO_OK;
}

}  // namespace
}  // namespace blco

extern "C" {

int blco_batch_create(int ntensors, blco_dev* const* devs, blco_batch** out)
{
    if (!out) return fail(BLCO_EINVAL, "null output handle");
    *out = nullptr;
    if (ntensors < 1 || !devs) return fail(BLCO_EINVAL, "empty batch");
    for (int t = 0; t < ntensors; ++t) {
        if (!devs[t]) return fail(BLCO_EINVAL, "tensor %d is null", t);
        if (devs[t]->stream_data) return fail(BLCO_EUNSUPPORTED, "tensor %d is streamed; a batch needs resident tensors", t);
        if (devs[t]->nnz >= (1ull << 40)) return fail(BLCO_EUNSUPPORTED, "tensor %d is too large for a batch", t);
    }
    blco_batch* b = new (std::nothrow) blco_batch();
    if (!b) return fail(BLCO_ENOMEM, "out of host memory");
    b->devs.assign(devs, devs + ntensors);
    for (blco_dev* d : b->devs) b->maxN = std::max(b->maxN, d->nmodes);
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return fail(BLCO_EHIP, "stream creation failed");
    }
    *out = b;
    return BLCO_OK;
}

void blco_batch_free(blco_batch* b)
{
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    for (auto& l : b->lists)
        if (l.d) (void)hipFree(l.d);
    if (b->d_tens) (void)hipFree(b->d_tens);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

int blco_batch_tune(blco_batch* b, int mode, int rank, int factor_dtype, void* const* factors,
                    void* const* outs, int* tpw, int* waves, int* wgs_per_cu, float* ms)
{
    int rc = check_call(b, mode, rank, factor_dtype);
    if (rc) return rc;
    if (!factors || !outs) return fail(BLCO_EINVAL, "null factor or output table");
    rc = upload_descriptors(b, descriptors(b, mode, factors, outs));
    if (rc) return rc;
    Tuned t;
    rc = tune(b, b->d_tens, mode, rank, factor_dtype, &t);
    if (rc) return rc;
    if (tpw) *tpw = t.tpw;
    if (waves) *waves = t.waves;
    if (wgs_per_cu) *wgs_per_cu = t.wpc;
    if (ms) *ms = t.ms;
    HIP_TRY(hipStreamSynchronize(b->st));
    return BLCO_OK;
}

int blco_batch_mttkrp(blco_batch* b, int mode, int rank, int factor_dtype, void* const* factors,
                      void* const* outs, int tpw, int waves, int wgs_per_cu)
{
    int rc = check_call(b, mode, rank, factor_dtype);
    if (rc) return rc;
    if (!factors || !outs) return fail(BLCO_EINVAL, "null factor or output table");
    for (int t : members(b, mode)) {
        if (!outs[t]) return fail(BLCO_EINVAL, "output of tensor %d is null", t);
        for (int m = 0; m < b->devs[t]->nmodes; ++m)
            if (!factors[(size_t)t * BLCO_MAX_MODES + m])
                return fail(BLCO_EINVAL, "factor %d of tensor %d is null", m, t);
    }
    if (tpw < 0 || tpw > 64 || waves < 0 || waves > BATCH_MAX_WAVES || wgs_per_cu < 0 || wgs_per_cu > 32)
        return fail(BLCO_EINVAL, "bad launch shape (%d, %d, %d)", tpw, waves, wgs_per_cu);
    rc = upload_descriptors(b, descriptors(b, mode, factors, outs));
    if (rc) return rc;
    if (!tpw || !waves || !wgs_per_cu) {  // the tuned shape, else a default
        tpw = 4;
        waves = 4;
        wgs_per_cu = 4;
        for (const Tuned& t : b->tuned)
            if (t.mode == mode && t.rank == rank && t.ftype == factor_dtype) {
                tpw = t.tpw;
                waves = t.waves;
                wgs_per_cu = t.wpc;
            }
    }
    rc = launch_batch(b, b->d_tens, mode, rank, factor_dtype, tpw, waves, wgs_per_cu, b->st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->st));
    return BLCO_OK;
}

int blco_batch_cpals(blco_batch* b, const blco_cpals_opts* o, double* const* U, double* lambda, double* fits,
                     int* iters, double* iter_ms)
{
    if (!b || !o || !U) return fail(BLCO_EINVAL, "null argument");
    const int T = (int)b->devs.size(), R = o->rank;
    const int ft = o->precision;
    int rc = check_call(b, 0, R, ft);
    if (rc) return rc;
    blco_cpals_opts o2 = *o;
    o2.use_graph = 0;
    o2.mode_order = 0;  // modes 0..N-1 in every tensor
    o2.verbose = 0;
    std::vector<blco_cpals_state*> S(T, nullptr);
    std::vector<BatchTensor*> d_desc(b->maxN, nullptr);
    std::vector<double*> own_fitbuf(T, nullptr);
    double* d_fits = nullptr;
    double* h_fits = nullptr;
    hipEvent_t fork = nullptr, it0 = nullptr, it1 = nullptr;
    std::vector<hipEvent_t> join(T, nullptr);
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    auto cleanup = [&](int code) {
        if (b->st) (void)hipStreamSynchronize(b->st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (int t = 0; t < T; ++t) {
            if (S[t]) {
                S[t]->fitbuf = own_fitbuf[t];
                blco_cpals_free(S[t]);
            }
            if (join[t]) (void)hipEventDestroy(join[t]);
        }
        for (BatchTensor* p : d_desc)
            if (p) (void)hipFree(p);
        if (d_fits) (void)hipFree(d_fits);
        if (h_fits) (void)hipHostFree(h_fits);
        for (hipEvent_t e : {fork, it0, it1})
            if (e) (void)hipEventDestroy(e);
        return code;
    };
    for (int t = 0; t < T; ++t) {
        rc = blco_cpals_create(b->devs[t], &o2, (const double* const*)(U + (size_t)t * BLCO_MAX_MODES), &S[t]);
        if (rc) return cleanup(rc);
        own_fitbuf[t] = S[t]->fitbuf;
        if (hipEventCreateWithFlags(&join[t], hipEventDisableTiming) != hipSuccess)
            return cleanup(fail(BLCO_EHIP, "event creation failed"));
    }
    // every tensor's fit lands in one array: one read-back per iteration
    if (hipMalloc((void**)&d_fits, sizeof(double) * 4 * T) != hipSuccess ||
        hipHostMalloc((void**)&h_fits, sizeof(double) * 4 * T, hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(BLCO_ENOMEM, "fit buffers"));
    for (int t = 0; t < T; ++t) S[t]->fitbuf = d_fits + 4 * t;
    if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess || hipEventCreate(&it0) != hipSuccess ||
        hipEventCreate(&it1) != hipSuccess)
        return cleanup(fail(BLCO_EHIP, "event creation failed"));
    // per-mode descriptors (the solver buffers do not move) and launch shapes
    std::vector<Tuned> shape(b->maxN);
    for (int n = 0; n < b->maxN; ++n) {
        std::vector<void*> F((size_t)T * BLCO_MAX_MODES, nullptr), O(T, nullptr);
        for (int t = 0; t < T; ++t) {
            for (int m = 0; m < S[t]->N; ++m) F[(size_t)t * BLCO_MAX_MODES + m] = S[t]->U[m];
            O[t] = S[t]->Mbuf;
        }
        const std::vector<BatchTensor> v = descriptors(b, n, F.data(), O.data());
        if (hipMalloc(&d_desc[n], sizeof(BatchTensor) * std::max<size_t>(1, v.size())) != hipSuccess)
            return cleanup(fail(BLCO_ENOMEM, "descriptor allocation failed"));
        if (!v.empty() && hipMemcpy(d_desc[n], v.data(), sizeof(BatchTensor) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(BLCO_EHIP, "descriptor upload failed"));
        rc = tune(b, d_desc[n], n, R, ft, &shape[n]);  // scribbles on the outputs, cleared below
        if (rc) return cleanup(rc);
    }
    if (hipStreamSynchronize(b->st) != hipSuccess) return cleanup(fail(BLCO_EHIP, "tuning failed"));

    // the sweep: per mode clear outputs, one persistent MTTKRP, every ALS step
    auto fork_each = [&](const std::vector<int>& who, auto&& body) -> int {
        HIP_TRY(hipEventRecord(fork, b->st));
        for (int t : who) {
            HIP_TRY(hipStreamWaitEvent(S[t]->st, fork, 0));
            int r = body(t);
            if (r) return r;
            HIP_TRY(hipEventRecord(join[t], S[t]->st));
            HIP_TRY(hipStreamWaitEvent(b->st, join[t], 0));
        }
        return BLCO_OK;
    };
    std::vector<int> all(T);
    for (int t = 0; t < T; ++t) all[t] = t;
    auto sweep = [&](int which_norm) -> int {
        for (int n = 0; n < b->maxN; ++n) {
            const std::vector<int> mem = members(b, n);
            int r = fork_each(mem, [&](int t) {
                return launch_fill(S[t]->Mbuf, S[t]->d->dims[n] * (uint64_t)R, S[t]->mt, 0.0, S[t]->st);
            });
            if (!r) r = launch_batch(b, d_desc[n], n, R, ft, shape[n].tpw, shape[n].waves, shape[n].wpc, b->st);
            if (!r) r = fork_each(mem, [&](int t) { return enqueue_step(S[t], n, which_norm, S[t]->st); });
            if (r) return r;
        }
        return fork_each(all, [&](int t) { return enqueue_fit(S[t], S[t]->st); });
    };

    std::vector<double> fit_old(T, 0.0);
    std::vector<int> stopped_at(T, 0);
    int alive = T, it = 0;
    double ms_sum = 0.0;
    int ms_n = 0;
    for (; it < o->max_iters && alive > 0; ++it) {
        HIP_TRY(hipEventRecord(it0, b->st));
        if (it == 0) {
            rc = sweep(0);
            if (rc) return cleanup(rc);
        } else {
            if (!gexec) {
                HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                rc = sweep(1);
                hipError_t e = hipStreamEndCapture(b->st, &graph);
                if (rc) return cleanup(rc);
                if (e != hipSuccess) return cleanup(fail(BLCO_EHIP, "sweep capture failed: %s", hipGetErrorString(e)));
                if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess)
                    return cleanup(fail(BLCO_EHIP, "graph instantiation failed"));
            }
            HIP_TRY(hipGraphLaunch(gexec, b->st));
        }
        HIP_TRY(hipEventRecord(it1, b->st));
        HIP_TRY(hipMemcpyAsync(h_fits, d_fits, sizeof(double) * 4 * T, hipMemcpyDeviceToHost, b->st));
        HIP_TRY(hipStreamSynchronize(b->st));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, it0, it1));
        if (it > 0 || o->max_iters == 1) {
            ms_sum += ms;
            ++ms_n;
        }
        for (int t = 0; t < T; ++t) {
            if (stopped_at[t]) continue;
            const double f = h_fits[4 * t];
            const bool stop = (it > 0 && std::fabs(fit_old[t] - f) < o->tolerance) || std::isnan(f) ||
                              it + 1 == o->max_iters;
            fit_old[t] = f;
            if (stop) {
                // this tensor's result is final: keep it before later sweeps move it
                stopped_at[t] = it + 1;
                --alive;
                rc = blco_cpals_get(S[t], U + (size_t)t * BLCO_MAX_MODES, lambda ? lambda + (size_t)t * R : nullptr);
                if (rc) return cleanup(rc);
            }
        }
    }
    for (int t = 0; t < T; ++t) {
        if (!stopped_at[t]) {  // max_iters == 0
            rc = blco_cpals_get(S[t], U + (size_t)t * BLCO_MAX_MODES, lambda ? lambda + (size_t)t * R : nullptr);
            if (rc) return cleanup(rc);
        }
        if (fits) fits[t] = fit_old[t];
        if (iters) iters[t] = stopped_at[t];
    }
    if (iter_ms) *iter_ms = ms_n ? ms_sum / ms_n : 0.0;
    return cleanup(BLCO_OK);
}

}  // extern "C"
