// selector.hip -- per-mode MTTKRP kernel choice by a learned cost model.
//
// The reference picks a kernel per mode with a fixed rule (src/alto_dev.cu:
// 1975,1996-2001: lvl3 when the mode has at most 100 rows, else lvl1).  Here
// a small network (mlp.hip) predicts log2 of the time of every candidate
// kernel from features of the mode, and the fastest prediction among the
// candidates that can run is taken.  With no network installed the rule of
// kernel id 10 applies.
//
// Features of (tensor, mode n, rank R, factor type), BLCO_SEL_NFEAT floats:
//   0 log2 I_n            1 log2 nnz            2 log2 blocks
//   3 log2 (nnz / blocks) 4 log2 R              5 order N
//   6 log2 factor bytes   7 log2 sum_{m!=n} I_m
//   8..15 fraction of nonzeros in mode-n slices (output rows) holding
//         1, 2-3, 4-15, 16-63, 64-255, 256-1023, 1024-4095, >= 4096 nonzeros
//   16 log2 largest slice 17 fraction of nonzeros in the 64 largest slices
//   18..21 fraction of nonzeros in mode-n fibers (all other indices fixed)
//         of length 1, 2-3, 4-15, >= 16
//   22 log2 (nnz / I_n)   23 1 if the output fits the lvl3 LDS copy
// Slice and fiber statistics are exact when the mode is short enough to
// count, else estimated from the groups whose hashed key falls in a 1/S
// sample (whole groups are kept or dropped, so sizes are not biased).
#include <dlfcn.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "devcommon.hpp"
#include "device.hpp"

namespace blco {

namespace {

// candidate c -> (kernel id, lvl3 partition count, tiles of 64 nonzeros per
// wave: 0 for the kernel's default)
constexpr int CAND_KERNEL[BLCO_SEL_NCAND] = {BLCO_K_BATCHED2, BLCO_K_BATCHED, BLCO_K_ATOMIC, BLCO_K_LVL3,
                                             BLCO_K_LVL3,     BLCO_K_LVL3,    BLCO_K_LVL3,  BLCO_K_LVL3,
                                             BLCO_K_LVL3,     BLCO_K_LVL3,    BLCO_K_SORTED, BLCO_K_BATCHED,
                                             BLCO_K_BATCHED,  BLCO_K_SORTED,  BLCO_K_SORTED};
constexpr int CAND_NPRTN[BLCO_SEL_NCAND] = {1, 1, 1, 1, 2, 4, 8, 16, 32, 64, 1, 1, 1, 1, 1};
constexpr int CAND_TPW[BLCO_SEL_NCAND] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 16, 8, 32};

// The network's choice replaces the rule's only when it predicts at least
// 2^SEL_MARGIN times less time (7%; the training-set regrets of margins 0
// to 0.6 lie within 0.9 points of each other, 4.1-5.0%, and 0.1 is within
// 0.2 of the lowest), only for tensors of at most 10^7 nonzeros (log2 of nnz
// up to SEL_MAX_LOG2_NNZ: the range of 400 of the 430 training tensors;
// above it, with 30, the network's held-out log2-time error is 0.54, far
// beyond the rule's 3.6% regret there), and only for modes whose size features (log2
// of I_n, nnz, blocks, nnz per block, R, sum of the other lengths, largest
// slice, nnz / I_n) lie within SEL_SIZE_Z standard deviations of the
// training set's mean (the training data reach 3.6).
constexpr float SEL_MARGIN = 0.1f;
constexpr float SEL_MAX_LOG2_NNZ = 23.25f;
constexpr float SEL_SIZE_Z = 4.0f;
constexpr int SIZE_FEATS[] = {0, 1, 2, 3, 4, 7, 16, 22};

// The candidate of the rule's kernel (default tiles per wave; lvl3 with 16
// partitions, the rule's timing set-up), or -1.
int rule_candidate(int kernel)
{
    for (int c = 0; c < BLCO_SEL_NCAND; ++c)
        if (CAND_KERNEL[c] == kernel && CAND_TPW[c] == 0 && (kernel != BLCO_K_LVL3 || CAND_NPRTN[c] == 16)) return c;
    return -1;
}

struct Selector {
    std::mutex mu;
    std::atomic<uint64_t> gen{1};  // bumped by every install; stale decisions are redone
    bool tried_default = false;
    int nlayers = 0;
    int widths[BLCO_MLP_MAX_LAYERS + 1] = {};
    std::vector<float> params, xmean, xstd;
    blco_mlp* net = nullptr;  // built on first use
    int net_device = -1;
};

Selector& sel()
{
    static Selector s;
    return s;
}

int install_locked(Selector& s, int nlayers, const int* widths, const float* params, const float* xmean,
                   const float* xstd);

// Reads a selector file (magic "BLCOSEL1", int32 layer count L, int32
// widths[L+1], float feature mean[NFEAT], float feature scale[NFEAT], float
// parameters in the network's flat order) and installs it; s.mu is held.
int load_file(Selector& s, const char* path, std::string* why)
{
    FILE* f = fopen(path, "rb");
    if (!f) {
        *why = "cannot open";
        return BLCO_EINVAL;
    }
    char magic[8];
    int32_t nl = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && !memcmp(magic, "BLCOSEL1", 8) && fread(&nl, 4, 1, f) == 1 &&
              nl >= 1 && nl <= BLCO_MLP_MAX_LAYERS;
    int32_t w[BLCO_MLP_MAX_LAYERS + 1] = {};
    if (ok) ok = fread(w, 4, nl + 1, f) == (size_t)nl + 1;
    if (ok) ok = w[0] == BLCO_SEL_NFEAT && w[nl] == BLCO_SEL_NCAND;
    uint64_t n = 0;
    for (int l = 0; ok && l < nl; ++l) {
        ok = w[l] >= 1 && w[l] <= BLCO_MLP_MAX_WIDTH && w[l + 1] >= 1 && w[l + 1] <= BLCO_MLP_MAX_WIDTH;
        n += (uint64_t)w[l] * w[l + 1] + w[l + 1];
    }
    std::vector<float> xm(BLCO_SEL_NFEAT), xs(BLCO_SEL_NFEAT), p(n);
    if (ok)
        ok = fread(xm.data(), 4, xm.size(), f) == xm.size() && fread(xs.data(), 4, xs.size(), f) == xs.size() &&
             fread(p.data(), 4, n, f) == n;
    fclose(f);
    if (!ok) {
        *why = "not a selector file of this build";
        return BLCO_EINVAL;
    }
    int widths[BLCO_MLP_MAX_LAYERS + 1];
    for (int l = 0; l <= nl; ++l) widths[l] = w[l];
    return install_locked(s, nl, widths, p.data(), xm.data(), xs.data());
}

// The selector file next to the library (<lib>/../selector/selector.bin),
// or $BLCO_SELECTOR; tried once; s.mu is held.
void try_default(Selector& s)
{
    if (s.tried_default) return;
    s.tried_default = true;
    std::string path;
    if (const char* e = getenv("BLCO_SELECTOR")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr((void*)&blco_selector_load, &info) && info.dli_fname) {
            std::string lib = info.dli_fname;
            size_t k = lib.rfind('/');
            std::string dir = k == std::string::npos ? "." : lib.substr(0, k);
            path = dir + "/../selector/selector.bin";
        }
    }
    if (path.empty() || access(path.c_str(), R_OK) != 0) return;
    std::string why;
    (void)load_file(s, path.c_str(), &why);
}

int install_locked(Selector& s, int nlayers, const int* widths, const float* params, const float* xmean,
                   const float* xstd)
{
    s.tried_default = true;  // an explicit network wins over the default file
    s.gen.fetch_add(1);
    if (s.net) blco_mlp_free(s.net);
    s.net = nullptr;
    s.nlayers = 0;
    if (nlayers == 0) return BLCO_OK;  // uninstall
    if (nlayers < 1 || nlayers > BLCO_MLP_MAX_LAYERS || !widths || !params || !xmean || !xstd)
        return fail(BLCO_EINVAL, "bad selector network");
    if (widths[0] != BLCO_SEL_NFEAT || widths[nlayers] != BLCO_SEL_NCAND)
        return fail(BLCO_EINVAL, "selector network maps %d -> %d, expected %d -> %d", widths[0],
                    widths[nlayers], BLCO_SEL_NFEAT, BLCO_SEL_NCAND);
    uint64_t n = 0;
    for (int l = 0; l < nlayers; ++l) {
        if (widths[l + 1] < 1 || widths[l + 1] > BLCO_MLP_MAX_WIDTH) return fail(BLCO_EINVAL, "bad width");
        n += (uint64_t)widths[l] * widths[l + 1] + widths[l + 1];
    }
    for (int j = 0; j < BLCO_SEL_NFEAT; ++j)
        if (!(xstd[j] > 0.f)) return fail(BLCO_EINVAL, "feature scale %d is not positive", j);
    s.params.assign(params, params + n);
    s.xmean.assign(xmean, xmean + BLCO_SEL_NFEAT);
    s.xstd.assign(xstd, xstd + BLCO_SEL_NFEAT);
    for (int l = 0; l <= nlayers; ++l) s.widths[l] = widths[l];
    s.nlayers = nlayers;
    return BLCO_OK;
}

}  // namespace

int mode_features(const blco_dev* d, int mode, int rank, int ftype, float* out)
{
    std::memcpy(out, d->sfeat[mode], sizeof(float) * BLCO_SEL_NFEAT);
    const int es = dtype_size(ftype);
    out[4] = (float)std::log2((double)rank);
    out[6] = (float)std::log2((double)es);
    out[23] = lvl3_output_fits(d, mode, rank, ftype) ? 1.f : 0.f;
    return BLCO_OK;
}

bool candidate_valid(const blco_dev* d, int mode, int rank, int ftype, int c)
{
    if (CAND_KERNEL[c] == BLCO_K_LVL3) return lvl3_output_fits(d, mode, rank, ftype);
    if (CAND_KERNEL[c] == BLCO_K_SORTED) return !d->stream_data && !d->sorted_failed[mode];
    return true;
}

// The learned choice for (mode, rank, ftype); false when no network is
// installed.
bool select_kernel(blco_dev* d, int mode, int rank, int ftype, int* kernel, int* nprtn, int* tpw)
{
    Selector& s = sel();
    const uint64_t gen = s.gen.load();
    for (const auto& c : d->sel_cache)
        if (c.mode == mode && c.rank == rank && c.ftype == ftype && c.gen == gen) {
            *kernel = c.kernel;
            *nprtn = c.nprtn;
            *tpw = c.tpw;
            return c.kernel != 0;
        }
    std::lock_guard<std::mutex> lk(s.mu);
    try_default(s);
    SelCache entry{mode, rank, ftype, 0, 1, 0, s.gen.load()};
    if (s.nlayers > 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (!s.net || s.net_device != dev) {
            if (s.net) blco_mlp_free(s.net);
            s.net = nullptr;
            if (blco_mlp_create(s.nlayers, s.widths, 0, &s.net) == BLCO_OK &&
                blco_mlp_set_params(s.net, s.params.data()) == BLCO_OK)
                s.net_device = dev;
            else if (s.net) {
                blco_mlp_free(s.net);
                s.net = nullptr;
            }
        }
        float x[BLCO_SEL_NFEAT], y[BLCO_SEL_NCAND];
        mode_features(d, mode, rank, ftype, x);
        // a mode whose sizes lie outside what the network was trained on
        // (more than 10^7 nonzeros, or a size feature beyond SEL_SIZE_Z
        // training standard deviations) is left to the rule
        bool inside = x[1] <= SEL_MAX_LOG2_NNZ;
        for (int j = 0; j < BLCO_SEL_NFEAT; ++j) x[j] = (x[j] - s.xmean[j]) / s.xstd[j];
        for (int f : SIZE_FEATS) inside = inside && std::fabs(x[f]) <= SEL_SIZE_Z;
        if (inside && s.net && blco_mlp_forward(s.net, x, 1, y) == BLCO_OK) {
            int best = -1;
            for (int c = 0; c < BLCO_SEL_NCAND; ++c)
                if (candidate_valid(d, mode, rank, ftype, c) && (best < 0 || y[c] < y[best])) best = c;
            // the rule's candidate unless the network predicts a clear gain
            const int rc = rule_candidate(auto_kernel(d, mode, rank, ftype));
            if (best >= 0 && rc >= 0 && candidate_valid(d, mode, rank, ftype, rc) && y[rc] - y[best] <= SEL_MARGIN)
                best = rc;
            if (best >= 0) {
                entry.kernel = CAND_KERNEL[best];
                entry.nprtn = CAND_NPRTN[best];
                entry.tpw = CAND_TPW[best];
            }
        }
    }
    for (auto& c : d->sel_cache)
        if (c.mode == mode && c.rank == rank && c.ftype == ftype) {
            c = entry;
            *kernel = entry.kernel;
            *nprtn = entry.nprtn;
            *tpw = entry.tpw;
            return entry.kernel != 0;
        }
    d->sel_cache.push_back(entry);
    *kernel = entry.kernel;
    *nprtn = entry.nprtn;
    *tpw = entry.tpw;
    return entry.kernel != 0;
}

}  // namespace blco

using namespace blco;

extern "C" {

int blco_sel_candidate(int c, int* kernel, int* nprtn, int* thread_cf)
{
    if (c < 0 || c >= BLCO_SEL_NCAND || !kernel || !nprtn || !thread_cf)
        return fail(BLCO_EINVAL, "candidate %d outside 0..%d", c, BLCO_SEL_NCAND - 1);
    *kernel = CAND_KERNEL[c];
    *nprtn = CAND_NPRTN[c];
    *thread_cf = CAND_TPW[c];
    return BLCO_OK;
}

int blco_mode_features(const blco_dev* d, int mode, int rank, int factor_dtype, float* feat)
{
    if (!d || !feat) return fail(BLCO_EINVAL, "null tensor or feature buffer");
    if (mode < 0 || mode >= d->nmodes) return fail(BLCO_EINVAL, "mode %d out of range", mode);
    if (rank < 1) return fail(BLCO_EINVAL, "rank %d", rank);
    if (factor_dtype < BLCO_F64 || factor_dtype > BLCO_BF16) return fail(BLCO_EINVAL, "dtype %d", factor_dtype);
    return mode_features(d, mode, rank, factor_dtype, feat);
}

int blco_sel_valid(const blco_dev* d, int mode, int rank, int factor_dtype, int c)
{
    if (!d || mode < 0 || mode >= d->nmodes || c < 0 || c >= BLCO_SEL_NCAND) return 0;
    return candidate_valid(d, mode, rank, factor_dtype, c) ? 1 : 0;
}

int blco_selector_install(int nlayers, const int* widths, const float* params, const float* xmean,
                          const float* xstd)
{
    Selector& s = sel();
    std::lock_guard<std::mutex> lk(s.mu);
    return install_locked(s, nlayers, widths, params, xmean, xstd);
}

int blco_selector_load(const char* path)
{
    if (!path) return fail(BLCO_EINVAL, "null path");
    Selector& s = sel();
    std::lock_guard<std::mutex> lk(s.mu);
    std::string why;
    int rc = load_file(s, path, &why);
    if (rc && why.size()) return fail(rc, "%s: %s", path, why.c_str());
    return rc;
}

int blco_selector_installed(void)
{
    Selector& s = sel();
    std::lock_guard<std::mutex> lk(s.mu);
    try_default(s);
    return s.nlayers > 0 ? 1 : 0;
}

int blco_select_kernel(blco_dev* d, int mode, int rank, int factor_dtype, int* kernel, int* nprtn, int* thread_cf)
{
    if (!d || !kernel || !nprtn || !thread_cf) return fail(BLCO_EINVAL, "null argument");
    if (mode < 0 || mode >= d->nmodes) return fail(BLCO_EINVAL, "mode %d out of range", mode);
    if (rank < 1) return fail(BLCO_EINVAL, "rank %d", rank);
    if (!select_kernel(d, mode, rank, factor_dtype, kernel, nprtn, thread_cf)) {
        *kernel = auto_kernel(d, mode, rank, factor_dtype);
        *nprtn = 0;
        *thread_cf = 0;
    }
    return BLCO_OK;
}

}  // extern "C"
