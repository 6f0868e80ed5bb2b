// solve_noalloc.hip -- the normal-equation solve of one ALS mode update
// without any device allocation, for callers that cannot allocate (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  It runs the kernels of blco_solve; their scratch (the R x R
// factor, the Jacobi work area, the status flag and the per-workgroup column
// statistics) lives in one file-scope device array instead of per-call
// allocations.
//
// The array exists once per GPU the library runs on and is held until the
// process ends.  Every call on a GPU uses that GPU's copy, so calls that
// overlap on one GPU (two streams, two host threads) race on it; callers
// serialise them.  Nothing in the scratch has to start at zero: the
// factorisation writes the status flag and the factor on both of its paths,
// and the row solve writes every workgroup's statistics slot.
#include <hip/hip_runtime.h>

#include <mutex>

#include "als.hpp"
#include "devcommon.hpp"

namespace blco {

namespace {

constexpr int NOALLOC_MAX_RANK = 64;
// fac (R^2) + Jacobi work (2 R^2 + R) + column statistics (at most
// 2048 workgroups x R) in doubles, then the status flag
constexpr size_t NOALLOC_FAC = (size_t)NOALLOC_MAX_RANK * NOALLOC_MAX_RANK;
constexpr size_t NOALLOC_WORK = 2 * NOALLOC_FAC + NOALLOC_MAX_RANK;
constexpr size_t NOALLOC_STATS = 2048 * (size_t)NOALLOC_MAX_RANK;
constexpr size_t NOALLOC_DOUBLES = NOALLOC_FAC + NOALLOC_WORK + NOALLOC_STATS;
constexpr size_t NOALLOC_BYTES = NOALLOC_DOUBLES * sizeof(double) + 256;

constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace

// The scratch array.  In host code its name is a host-side stand-in, not a
// device address: the address of each GPU's copy is looked up once per
// device (scratch_ptr) and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// Device address of the current GPU's scratch; looked up on the first call
// on each device, with that device current.
int scratch_ptr(void** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BLCO_EUNSUPPORTED, "device %d beyond the %d the scratch table holds", dev,
                    NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BLCO_EHIP, "no device address for the solve scratch on device %d", dev);
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return BLCO_OK;
}

}  // namespace

}  // namespace blco

using namespace blco;

extern "C" uint64_t blco_solve_noalloc_scratch_bytes(void) { return (uint64_t)NOALLOC_BYTES; }

extern "C" int blco_solve_noalloc(double* const* grams, int nmodes, int mode, int rank, const void* M, void* U,
                                  int dtype, uint64_t rows, int* used_fallback, void* stream)
{
    if (!grams || !M || !U || nmodes < 2 || nmodes > BLCO_MAX_MODES || mode < 0 || mode >= nmodes)
        return fail(BLCO_EINVAL, "bad solve arguments");
    if (rank < 1 || rank > NOALLOC_MAX_RANK)
        return fail(BLCO_EUNSUPPORTED, "rank %d outside [1, %d]: the allocation-free solve's scratch holds "
                    "%llu bytes", rank, NOALLOC_MAX_RANK, (unsigned long long)NOALLOC_BYTES);
    if (dtype < BLCO_F64 || dtype > BLCO_BF16) return fail(BLCO_EINVAL, "factor dtype %d", dtype);
    const uint32_t np = solve_parts(rows);
    if ((size_t)np * rank > NOALLOC_STATS)
        return fail(BLCO_EUNSUPPORTED, "%u row-solve workgroups exceed the scratch", np);
    void* base = nullptr;
    int rc = scratch_ptr(&base);
    if (rc) return rc;
    double* fac = static_cast<double*>(base);
    double* work = fac + (size_t)rank * rank;
    double* stats = work + 2 * (size_t)rank * rank + rank;
    int* status = reinterpret_cast<int*>(static_cast<double*>(base) + NOALLOC_DOUBLES);
    hipStream_t st = (hipStream_t)stream;
    GramSet gs;
    for (int m = 0; m < nmodes; ++m) gs.g[m] = grams[m];
    rc = launch_factorize(gs, nmodes, mode, rank, gram_eps(BLCO_F64), fac, work, status, st);
    if (!rc) rc = launch_solve_rows(M, U, dtype, rows, rank, fac, 0, stats, st);
    if (rc) return rc;
    if (used_fallback) {
        int h_status = 0;
        HIP_TRY(hipMemcpyAsync(&h_status, status, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        *used_fallback = h_status;
    }
    return BLCO_OK;
}
