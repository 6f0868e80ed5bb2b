// mlp.hip -- a small fully connected network (ReLU hidden layers, linear
// output) run and trained on the GPU with its own kernels: forward, masked
// mean-squared-error loss, backward, Adam.  It is the model behind the
// learned MTTKRP kernel selector: inputs are per-mode tensor features,
// outputs are predicted log2 kernel times, one per candidate kernel.
//
// Layer l maps width w[l] to w[l+1]: Y = X W_l^T + b_l, W_l row-major
// (w[l+1] x w[l]).  Parameters are stored flat, layer by layer, W then b.
// Sizes are small (widths <= 256, batches of a few thousand rows), so every
// kernel is one lane per output element with a short serial loop: no atomics,
// and results do not depend on the launch shape.
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../host/rng.hpp"
#include "devcommon.hpp"

struct blco_mlp {
    int nlayers = 0;
    int w[BLCO_MLP_MAX_LAYERS + 1] = {};
    uint64_t woff[BLCO_MLP_MAX_LAYERS] = {}, boff[BLCO_MLP_MAX_LAYERS] = {};
    uint64_t nparams = 0;
    float* p = nullptr;  // parameters
    float* g = nullptr;  // gradients
    float* m = nullptr;  // Adam first moment
    float* v = nullptr;  // Adam second moment
    int step = 0;
    int cap = 0;                                 // batch capacity of the buffers
    float* act[BLCO_MLP_MAX_LAYERS + 1] = {};    // act[0] = input, act[l+1] = layer l output
    float* dact[BLCO_MLP_MAX_LAYERS + 1] = {};   // d loss / d act
    float* tgt = nullptr;
    float* msk = nullptr;
    float* loss = nullptr;  // [2]: loss, mask sum
    hipStream_t st = nullptr;
};

namespace blco {
namespace {

constexpr uint64_t STREAM_MLP = 0x500;

// Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o])
__global__ __launch_bounds__(WG) void k_dense_fwd(const float* __restrict__ X, const float* __restrict__ W,
                                                  const float* __restrict__ bias, float* __restrict__ Y,
                                                  int B, int nin, int nout, int relu)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= B * nout) return;
    const int b = t / nout, o = t % nout;
    const float* x = X + (uint64_t)b * nin;
    const float* wr = W + (uint64_t)o * nin;
    float s = bias[o];
    for (int i = 0; i < nin; ++i) s = fmaf(x[i], wr[i], s);
    Y[t] = relu && s < 0.f ? 0.f : s;
}

// dX[b][i] = (sum_o dY[b][o] W[o][i]) * (X[b][i] > 0 when X is a ReLU output)
__global__ __launch_bounds__(WG) void k_dense_bwd_x(const float* __restrict__ dY, const float* __restrict__ W,
                                                    const float* __restrict__ X, float* __restrict__ dX,
                                                    int B, int nin, int nout, int relu_in)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= B * nin) return;
    const int b = t / nin, i = t % nin;
    float s = 0.f;
    if (!relu_in || X[t] > 0.f) {
        const float* dy = dY + (uint64_t)b * nout;
        for (int o = 0; o < nout; ++o) s = fmaf(dy[o], W[(uint64_t)o * nin + i], s);
    }
    dX[t] = s;
}

// dW[o][i] = sum_b dY[b][o] X[b][i];  db[o] = sum_b dY[b][o]
__global__ __launch_bounds__(WG) void k_dense_bwd_w(const float* __restrict__ dY, const float* __restrict__ X,
                                                    float* __restrict__ dW, float* __restrict__ db,
                                                    int B, int nin, int nout)
{
    const int t = blockIdx.x * WG + threadIdx.x;
    if (t >= nout * (nin + 1)) return;
    const int o = t / (nin + 1), i = t % (nin + 1);
    float s = 0.f;
    if (i < nin)
        for (int b = 0; b < B; ++b) s = fmaf(dY[(uint64_t)b * nout + o], X[(uint64_t)b * nin + i], s);
    else
        for (int b = 0; b < B; ++b) s += dY[(uint64_t)b * nout + o];
    if (i < nin) dW[(uint64_t)o * nin + i] = s;
    else db[o] = s;
}

// loss = sum M (Y - T)^2 / sum M;  dY = 2 M (Y - T) / sum M.  One workgroup.
__global__ __launch_bounds__(WG) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                            const float* __restrict__ M, float* __restrict__ dY,
                                            float* __restrict__ out, int n)
{
    __shared__ float s_l[WG], s_m[WG];
    float l = 0.f, msum = 0.f;
    for (int i = threadIdx.x; i < n; i += WG) {
        const float d = Y[i] - T[i];
        l += M[i] * d * d;
        msum += M[i];
    }
    s_l[threadIdx.x] = l;
    s_m[threadIdx.x] = msum;
    __syncthreads();
    for (int h = WG / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s_l[threadIdx.x] += s_l[threadIdx.x + h];
            s_m[threadIdx.x] += s_m[threadIdx.x + h];
        }
        __syncthreads();
    }
    const float den = s_m[0] > 0.f ? s_m[0] : 1.f;
    for (int i = threadIdx.x; i < n; i += WG) dY[i] = 2.f * M[i] * (Y[i] - T[i]) / den;
    if (threadIdx.x == 0) {
        out[0] = s_l[0] / den;
        out[1] = s_m[0];
    }
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ __launch_bounds__(WG) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                             float* __restrict__ m, float* __restrict__ v, uint64_t n,
                                             float lr_t, float b1, float b2, float eps)
{
    const uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

inline uint32_t grid_for(uint64_t n) { return (uint32_t)((n + WG - 1) / WG); }

int ensure_batch(blco_mlp* h, int B)
{
    if (B <= h->cap) return BLCO_OK;
    for (int l = 0; l <= h->nlayers; ++l) {
        if (h->act[l]) (void)hipFree(h->act[l]);
        if (h->dact[l]) (void)hipFree(h->dact[l]);
        h->act[l] = h->dact[l] = nullptr;
    }
    if (h->tgt) (void)hipFree(h->tgt);
    if (h->msk) (void)hipFree(h->msk);
    h->tgt = h->msk = nullptr;
    h->cap = 0;
    for (int l = 0; l <= h->nlayers; ++l) {
        HIP_TRY(hipMalloc(&h->act[l], sizeof(float) * (size_t)B * h->w[l]));
        HIP_TRY(hipMalloc(&h->dact[l], sizeof(float) * (size_t)B * h->w[l]));
    }
    const int nout = h->w[h->nlayers];
    HIP_TRY(hipMalloc(&h->tgt, sizeof(float) * (size_t)B * nout));
    HIP_TRY(hipMalloc(&h->msk, sizeof(float) * (size_t)B * nout));
    h->cap = B;
    return BLCO_OK;
}

int forward(blco_mlp* h, int B)
{
    for (int l = 0; l < h->nlayers; ++l) {
        const int nin = h->w[l], nout = h->w[l + 1];
        hipLaunchKernelGGL(k_dense_fwd, dim3(grid_for((uint64_t)B * nout)), dim3(WG), 0, h->st, h->act[l],
                           h->p + h->woff[l], h->p + h->boff[l], h->act[l + 1], B, nin, nout,
                           l + 1 < h->nlayers ? 1 : 0);
        HIP_TRY(hipGetLastError());
    }
    return BLCO_OK;
}

int backward(blco_mlp* h, int B)
{
    for (int l = h->nlayers - 1; l >= 0; --l) {
        const int nin = h->w[l], nout = h->w[l + 1];
        hipLaunchKernelGGL(k_dense_bwd_w, dim3(grid_for((uint64_t)nout * (nin + 1))), dim3(WG), 0, h->st,
                           h->dact[l + 1], h->act[l], h->g + h->woff[l], h->g + h->boff[l], B, nin, nout);
        HIP_TRY(hipGetLastError());
        if (l > 0) {
            hipLaunchKernelGGL(k_dense_bwd_x, dim3(grid_for((uint64_t)B * nin)), dim3(WG), 0, h->st,
                               h->dact[l + 1], h->p + h->woff[l], h->act[l], h->dact[l], B, nin, nout, 1);
            HIP_TRY(hipGetLastError());
        }
    }
    return BLCO_OK;
}

int loss_grad(blco_mlp* h, const float* X, const float* T, const float* M, int B)
{
    int rc = ensure_batch(h, B);
    if (rc) return rc;
    const int nin = h->w[0], nout = h->w[h->nlayers];
    HIP_TRY(hipMemcpyAsync(h->act[0], X, sizeof(float) * (size_t)B * nin, hipMemcpyHostToDevice, h->st));
    HIP_TRY(hipMemcpyAsync(h->tgt, T, sizeof(float) * (size_t)B * nout, hipMemcpyHostToDevice, h->st));
    HIP_TRY(hipMemcpyAsync(h->msk, M, sizeof(float) * (size_t)B * nout, hipMemcpyHostToDevice, h->st));
    if ((rc = forward(h, B))) return rc;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(WG), 0, h->st, h->act[h->nlayers], h->tgt, h->msk,
                       h->dact[h->nlayers], h->loss, B * nout);
    HIP_TRY(hipGetLastError());
    return backward(h, B);
}

int adam(blco_mlp* h, float lr, float b1, float b2, float eps)
{
    h->step += 1;
    const double c1 = 1.0 - std::pow((double)b1, h->step), c2 = 1.0 - std::pow((double)b2, h->step);
    const float lr_t = (float)(lr * std::sqrt(c2) / c1);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(h->nparams)), dim3(WG), 0, h->st, h->p, h->g, h->m, h->v,
                       h->nparams, lr_t, b1, b2, eps);
    HIP_TRY(hipGetLastError());
    return BLCO_OK;
}

}  // namespace
}  // namespace blco

using namespace blco;

extern "C" {

int blco_mlp_create(int nlayers, const int* widths, uint64_t seed, blco_mlp** out)
{
    if (!out) return fail(BLCO_EINVAL, "null output handle");
    *out = nullptr;
    if (nlayers < 1 || nlayers > BLCO_MLP_MAX_LAYERS || !widths)
        return fail(BLCO_EINVAL, "layer count %d outside 1..%d", nlayers, BLCO_MLP_MAX_LAYERS);
    for (int l = 0; l <= nlayers; ++l)
        if (widths[l] < 1 || widths[l] > BLCO_MLP_MAX_WIDTH)
            return fail(BLCO_EINVAL, "width %d of layer %d outside 1..%d", widths[l], l, BLCO_MLP_MAX_WIDTH);
    if (access("/dev/kfd", F_OK) != 0) return fail(BLCO_ENODEV, "no GPU (/dev/kfd absent)");
    blco_mlp* h = new (std::nothrow) blco_mlp();
    if (!h) return fail(BLCO_ENOMEM, "out of host memory");
    h->nlayers = nlayers;
    uint64_t n = 0;
    for (int l = 0; l <= nlayers; ++l) h->w[l] = widths[l];
    for (int l = 0; l < nlayers; ++l) {
        h->woff[l] = n;
        n += (uint64_t)widths[l] * widths[l + 1];
        h->boff[l] = n;
        n += widths[l + 1];
    }
    h->nparams = n;
    // He-uniform weights U(-sqrt(6/fan_in), sqrt(6/fan_in)), zero biases,
    // from the counter generator: identical for a seed on every machine.
    std::vector<float> p(n, 0.f);
    for (int l = 0; l < nlayers; ++l) {
        const double a = std::sqrt(6.0 / widths[l]);
        for (uint64_t k = 0; k < (uint64_t)widths[l] * widths[l + 1]; ++k)
            p[h->woff[l] + k] = (float)((2.0 * rng_u01(seed, STREAM_MLP + l, k) - 1.0) * a);
    }
    auto bad = [&](int rc) { blco_mlp_free(h); return rc; };
    if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
        return bad(fail(BLCO_EHIP, "stream creation failed"));
    if (hipMalloc(&h->p, sizeof(float) * n) != hipSuccess || hipMalloc(&h->g, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&h->m, sizeof(float) * n) != hipSuccess || hipMalloc(&h->v, sizeof(float) * n) != hipSuccess ||
        hipMalloc(&h->loss, sizeof(float) * 2) != hipSuccess)
        return bad(fail(BLCO_ENOMEM, "device allocation failed (network)"));
    if (hipMemcpy(h->p, p.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->g, 0, sizeof(float) * n) != hipSuccess || hipMemset(h->m, 0, sizeof(float) * n) != hipSuccess ||
        hipMemset(h->v, 0, sizeof(float) * n) != hipSuccess)
        return bad(fail(BLCO_EHIP, "upload of the initial parameters failed"));
    *out = h;
    return BLCO_OK;
}

void blco_mlp_free(blco_mlp* h)
{
    if (!h) return;
    if (h->st) (void)hipStreamSynchronize(h->st);
    for (float* q : {h->p, h->g, h->m, h->v, h->tgt, h->msk, h->loss})
        if (q) (void)hipFree(q);
    for (int l = 0; l <= h->nlayers; ++l) {
        if (h->act[l]) (void)hipFree(h->act[l]);
        if (h->dact[l]) (void)hipFree(h->dact[l]);
    }
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

uint64_t blco_mlp_nparams(const blco_mlp* h) { return h ? h->nparams : 0; }

int blco_mlp_get_params(const blco_mlp* h, float* params, float* grads)
{
    if (!h) return fail(BLCO_EINVAL, "null network");
    HIP_TRY(hipStreamSynchronize(h->st));
    if (params) HIP_TRY(hipMemcpy(params, h->p, sizeof(float) * h->nparams, hipMemcpyDeviceToHost));
    if (grads) HIP_TRY(hipMemcpy(grads, h->g, sizeof(float) * h->nparams, hipMemcpyDeviceToHost));
    return BLCO_OK;
}

int blco_mlp_set_params(blco_mlp* h, const float* params)
{
    if (!h || !params) return fail(BLCO_EINVAL, "null network or parameters");
    HIP_TRY(hipStreamSynchronize(h->st));
    HIP_TRY(hipMemcpy(h->p, params, sizeof(float) * h->nparams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(h->m, 0, sizeof(float) * h->nparams));
    HIP_TRY(hipMemset(h->v, 0, sizeof(float) * h->nparams));
    h->step = 0;
    return BLCO_OK;
}

int blco_mlp_forward(blco_mlp* h, const float* X, int batch, float* Y)
{
    if (!h || !X || !Y || batch < 0) return fail(BLCO_EINVAL, "bad forward arguments");
    if (batch == 0) return BLCO_OK;
    int rc = ensure_batch(h, batch);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(h->act[0], X, sizeof(float) * (size_t)batch * h->w[0], hipMemcpyHostToDevice, h->st));
    if ((rc = forward(h, batch))) return rc;
    HIP_TRY(hipMemcpyAsync(Y, h->act[h->nlayers], sizeof(float) * (size_t)batch * h->w[h->nlayers],
                           hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    return BLCO_OK;
}

int blco_mlp_loss_grad(blco_mlp* h, const float* X, const float* T, const float* M, int batch, float* loss)
{
    if (!h || !X || !T || !M || batch < 1) return fail(BLCO_EINVAL, "bad loss arguments");
    int rc = loss_grad(h, X, T, M, batch);
    if (rc) return rc;
    float l[2];
    HIP_TRY(hipMemcpyAsync(l, h->loss, sizeof(l), hipMemcpyDeviceToHost, h->st));
    HIP_TRY(hipStreamSynchronize(h->st));
    if (loss) *loss = l[0];
    return BLCO_OK;
}

int blco_mlp_adam_step(blco_mlp* h, float lr, float beta1, float beta2, float eps)
{
    if (!h || !(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
        return fail(BLCO_EINVAL, "bad Adam arguments");
    return adam(h, lr, beta1, beta2, eps);
}

int blco_mlp_train(blco_mlp* h, const float* X, const float* T, const float* M, int n, int epochs, int batch,
                   float lr, uint64_t seed, float* loss_trace)
{
    if (!h || !X || !T || !M || n < 1 || epochs < 0 || batch < 1 || !(lr > 0.f))
        return fail(BLCO_EINVAL, "bad training arguments");
    const int nin = h->w[0], nout = h->w[h->nlayers];
    batch = batch > n ? n : batch;
    std::vector<int> order(n);
    std::vector<float> bx((size_t)batch * nin), bt((size_t)batch * nout), bm((size_t)batch * nout);
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < n; ++i) order[i] = i;
        for (int i = n - 1; i > 0; --i) {  // Fisher-Yates from the counter generator
            const int j = (int)(rng_u64(seed, STREAM_MLP + 0x80, (uint64_t)e * n + i) % (uint64_t)(i + 1));
            std::swap(order[i], order[j]);
        }
        double esum = 0.0;
        int nb = 0;
        for (int s = 0; s < n; s += batch) {
            const int B = std::min(batch, n - s);
            for (int k = 0; k < B; ++k) {
                const int r = order[s + k];
                std::memcpy(&bx[(size_t)k * nin], X + (size_t)r * nin, sizeof(float) * nin);
                std::memcpy(&bt[(size_t)k * nout], T + (size_t)r * nout, sizeof(float) * nout);
                std::memcpy(&bm[(size_t)k * nout], M + (size_t)r * nout, sizeof(float) * nout);
            }
            int rc = loss_grad(h, bx.data(), bt.data(), bm.data(), B);
            if (rc) return rc;
            if ((rc = adam(h, lr, 0.9f, 0.999f, 1e-8f))) return rc;
            if (loss_trace) {
                float l[2];
                HIP_TRY(hipMemcpyAsync(l, h->loss, sizeof(l), hipMemcpyDeviceToHost, h->st));
                HIP_TRY(hipStreamSynchronize(h->st));
                esum += l[0];
            } else {
                HIP_TRY(hipStreamSynchronize(h->st));  // the batch buffers are reused
            }
            ++nb;
        }
        if (loss_trace) loss_trace[e] = (float)(esum / nb);
    }
    HIP_TRY(hipStreamSynchronize(h->st));
    return BLCO_OK;
}

}  // extern "C"
