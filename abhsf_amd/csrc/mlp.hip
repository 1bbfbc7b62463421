// mlp.hip -- the tuner's cost model: a 2 x 64 hidden-layer ReLU MLP in FP32,
// with hand-written forward, backward and Adam kernels (no reference
// counterpart; the reference's block-size search is exhaustive over storage
// bits, block_statistics/mmf.h:77-102).
//
// Parameters theta (float, P = 64*F + 64 + 64*64 + 64 + 64 + 1):
//   W1[64][F], b1[64], W2[64][64], b2[64], W3[64], b3
// Forward per sample (one thread each): h1 = relu(W1 x + b1),
//   h2 = relu(W2 h1 + b2), out = W3 . h2 + b3.
// Loss: weighted mean squared error sum_b w_b (out_b - y_b)^2 / sum_b w_b.
// Gradients of the weights are sums over samples taken in sample order by one
// thread per parameter, so training is deterministic.
#include "common.hpp"

#include <cmath>
#include <mutex>
#include <vector>

namespace abhsf {
namespace {

constexpr int H = 64;
constexpr int kT = 256;
constexpr int kMaxF = 64;

struct Layout {
    int64_t W1, b1, W2, b2, W3, b3, P;
};

__host__ __device__ inline Layout layout(int F)
{
    Layout L;
    L.W1 = 0;
    L.b1 = (int64_t)H * F;
    L.W2 = L.b1 + H;
    L.b2 = L.W2 + (int64_t)H * H;
    L.W3 = L.b2 + H;
    L.b3 = L.W3 + H;
    L.P = L.b3 + 1;
    return L;
}

__global__ __launch_bounds__(kT) void k_mlp_fwd(const float *__restrict__ theta, int F, const float *__restrict__ X,
                                                int64_t B, float *__restrict__ H1, float *__restrict__ H2,
                                                float *__restrict__ out)
{
    __shared__ float sW[H * kMaxF + H + H * H + H + H + 1];
    const Layout L = layout(F);
    for (int64_t i = threadIdx.x; i < L.P; i += blockDim.x) sW[i] = theta[i];
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (b >= B) return;
    float h1[H], h2[H];
#pragma unroll
    for (int i = 0; i < H; i++) h1[i] = sW[L.b1 + i];
    for (int f = 0; f < F; f++) {
        const float xf = X[b * F + f];
#pragma unroll
        for (int i = 0; i < H; i++) h1[i] = fmaf(sW[L.W1 + (int64_t)i * F + f], xf, h1[i]);
    }
#pragma unroll
    for (int i = 0; i < H; i++) {
        h1[i] = h1[i] > 0.f ? h1[i] : 0.f;
        H1[b * H + i] = h1[i];
    }
#pragma unroll
    for (int i = 0; i < H; i++) {
        float a = sW[L.b2 + i];
#pragma unroll
        for (int j = 0; j < H; j++) a = fmaf(sW[L.W2 + i * H + j], h1[j], a);
        h2[i] = a > 0.f ? a : 0.f;
        H2[b * H + i] = h2[i];
    }
    float o = sW[L.b3];
#pragma unroll
    for (int j = 0; j < H; j++) o = fmaf(sW[L.W3 + j], h2[j], o);
    out[b] = o;
}

// Ensemble inference in one launch: out[b] = mean over the nm members of the
// forward pass of sample b (thetas: nm parameter vectors of P floats, one
// after the other).  Each member's parameters are staged in LDS in turn; the
// members' outputs are summed in FP64 in member order.
__global__ __launch_bounds__(kT) void k_mlp_ens_fwd(const float *__restrict__ thetas, int nm, int F,
                                                    const float *__restrict__ X, int64_t B, double *__restrict__ out)
{
    __shared__ float sW[H * kMaxF + H + H * H + H + H + 1];
    const Layout L = layout(F);
    const int64_t b = (int64_t)blockIdx.x * kT + threadIdx.x;
    double acc = 0.0;
    for (int mb = 0; mb < nm; mb++) {
        __syncthreads();  // the previous member's parameters are no longer read
        const float *theta = thetas + (int64_t)mb * L.P;
        for (int64_t i = threadIdx.x; i < L.P; i += blockDim.x) sW[i] = theta[i];
        __syncthreads();
        if (b >= B) continue;
        float h1[H];
#pragma unroll
        for (int i = 0; i < H; i++) h1[i] = sW[L.b1 + i];
        for (int f = 0; f < F; f++) {
            const float xf = X[b * F + f];
#pragma unroll
            for (int i = 0; i < H; i++) h1[i] = fmaf(sW[L.W1 + (int64_t)i * F + f], xf, h1[i]);
        }
#pragma unroll
        for (int i = 0; i < H; i++) h1[i] = h1[i] > 0.f ? h1[i] : 0.f;
        float o = sW[L.b3];
        for (int i = 0; i < H; i++) {
            float a = sW[L.b2 + i];
#pragma unroll
            for (int j = 0; j < H; j++) a = fmaf(sW[L.W2 + i * H + j], h1[j], a);
            o = fmaf(sW[L.W3 + i], a > 0.f ? a : 0.f, o);
        }
        acc += (double)o;
    }
    if (b < B) out[b] = acc / nm;
}

// Per-sample deltas: D3 = dL/dout, D2 = dL/dz2, D1 = dL/dz1.
// wsum_p (device, may be null) replaces wsum when set.
__global__ __launch_bounds__(kT) void k_mlp_delta(const float *__restrict__ theta, int F, int64_t B,
                                                  const float *__restrict__ H1, const float *__restrict__ H2,
                                                  const float *__restrict__ out, const float *__restrict__ y,
                                                  const float *__restrict__ w, float wsum,
                                                  const float *__restrict__ wsum_p, float *__restrict__ D1,
                                                  float *__restrict__ D2, float *__restrict__ D3)
{
    if (wsum_p) wsum = *wsum_p;
    __shared__ float sW2[H * H + H];
    const Layout L = layout(F);
    for (int i = threadIdx.x; i < H * H + H; i += blockDim.x) sW2[i] = theta[L.W2 + i];  // W2 and b2, then W3
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (b >= B) return;
    const float wb = w ? w[b] : 1.f;
    const float d3 = 2.f * wb * (out[b] - y[b]) / wsum;
    D3[b] = d3;
    float d2[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
        d2[i] = (H2[b * H + i] > 0.f) ? theta[L.W3 + i] * d3 : 0.f;
        D2[b * H + i] = d2[i];
    }
#pragma unroll
    for (int j = 0; j < H; j++) {
        float a = 0.f;
#pragma unroll
        for (int i = 0; i < H; i++) a = fmaf(sW2[i * H + j], d2[i], a);
        D1[b * H + j] = (H1[b * H + j] > 0.f) ? a : 0.f;
    }
}

// grad[p] = sum over samples, in sample order.
__global__ __launch_bounds__(kT) void k_mlp_grad(int F, const float *__restrict__ X, int64_t B,
                                                 const float *__restrict__ H1, const float *__restrict__ H2,
                                                 const float *__restrict__ D1, const float *__restrict__ D2,
                                                 const float *__restrict__ D3, float *__restrict__ grad)
{
    const Layout L = layout(F);
    const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (p >= L.P) return;
    float a = 0.f;
    if (p < L.b1) {
        const int i = (int)(p / F), f = (int)(p % F);
        for (int64_t b = 0; b < B; b++) a = fmaf(D1[b * H + i], X[b * F + f], a);
    } else if (p < L.W2) {
        const int i = (int)(p - L.b1);
        for (int64_t b = 0; b < B; b++) a += D1[b * H + i];
    } else if (p < L.b2) {
        const int i = (int)((p - L.W2) / H), j = (int)((p - L.W2) % H);
        for (int64_t b = 0; b < B; b++) a = fmaf(D2[b * H + i], H1[b * H + j], a);
    } else if (p < L.W3) {
        const int i = (int)(p - L.b2);
        for (int64_t b = 0; b < B; b++) a += D2[b * H + i];
    } else if (p < L.b3) {
        const int j = (int)(p - L.W3);
        for (int64_t b = 0; b < B; b++) a = fmaf(D3[b], H2[b * H + j], a);
    } else {
        for (int64_t b = 0; b < B; b++) a += D3[b];
    }
    grad[p] = a;
}

__global__ __launch_bounds__(kT) void k_adam(float *__restrict__ theta, const float *__restrict__ grad,
                                             float *__restrict__ m, float *__restrict__ v, int64_t P, float lr,
                                             float b1, float b2, float eps, float c1, float c2)
{
    const int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (p >= P) return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.f - b1) * g;
    const float vp = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    theta[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// Weighted squared error sum, one workgroup, fixed order.
__global__ __launch_bounds__(kT) void k_loss(const float *__restrict__ out, const float *__restrict__ y,
                                             const float *__restrict__ w, int64_t B, float wsum,
                                             const float *__restrict__ wsum_p, float *__restrict__ loss)
{
    if (wsum_p) wsum = *wsum_p;
    __shared__ float red[kT];
    float a = 0.f;
    for (int64_t b = threadIdx.x; b < B; b += kT) {
        const float d = out[b] - y[b];
        a += (w ? w[b] : 1.f) * d * d;
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = kT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0] / wsum;
}

// Sum of the sample weights in sample order, in double, rounded to float:
// the same value the host computes for abhsf_mlp_train.  One thread.
__global__ void k_wsum(const float *__restrict__ w, int64_t B, float *__restrict__ wsum)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int64_t b = 0; b < B; b++) s += (double)w[b];
    *wsum = (float)s;
}

__global__ __launch_bounds__(kT) void k_fill0(float *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

inline unsigned g(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// Training work buffer of abhsf_mlp_train_noalloc: a file-scope device array
// (one copy per GPU, held until the process ends) instead of a hipMalloc per
// call.  Layout per call: H1, H2, D1, D2 [B][64], out, D3 [B], grad, m, v [P],
// loss [steps], wsum [1].
constexpr int64_t kNoAllocMaxSamples = 4096;
constexpr int64_t kNoAllocMaxSteps = 16384;
constexpr int64_t kMaxParams = (int64_t)H * kMaxF + H + (int64_t)H * H + H + H + 1;
constexpr int64_t kScratchFloats =
    kNoAllocMaxSamples * (4 * H + 2) + 3 * kMaxParams + kNoAllocMaxSteps + 1;
static_assert(kScratchFloats * 4 < (8ll << 20), "scratch stays under 8 MiB per GPU");

}  // namespace
}  // namespace abhsf

using namespace abhsf;

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {
// Device address of `scratch` on the current device, looked up once per
// device with that device current (each GPU holds its own copy at its own
// address; the host-side name is not a device address).
std::mutex g_scratch_mu;
float *g_scratch_dev[64] = {};

float *scratch_on_current_device()
{
    int dev = 0;
    ABHSF_HIP(hipGetDevice(&dev));
    ABHSF_REQUIRE(dev >= 0 && dev < 64, ABHSF_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        ABHSF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ABHSF_REQUIRE(p != nullptr, ABHSF_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = static_cast<float *>(p);
    }
    return g_scratch_dev[dev];
}
}  // namespace

extern "C" int64_t abhsf_mlp_num_params(int32_t F) { return (F >= 1 && F <= kMaxF) ? layout(F).P : -1; }

extern "C" abhsf_status abhsf_mlp_forward_impl(const float *theta, int32_t F, const float *X, int64_t B, float *H1,
                                               float *H2, float *out, void *stream, std::string *err)
{
    try {
        ABHSF_REQUIRE(theta && X && H1 && H2 && out && B >= 0, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(F >= 1 && F <= kMaxF, ABHSF_ERR_UNSUPPORTED, "feature count must be in [1, 64]");
        if (B == 0) return ABHSF_OK;
        k_mlp_fwd<<<g(B), kT, 0, static_cast<hipStream_t>(stream)>>>(theta, F, X, B, H1, H2, out);
        ABHSF_HIP(hipGetLastError());
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}

extern "C" abhsf_status abhsf_mlp_predict_impl(const float *thetas, int32_t nm, int32_t F, const float *X,
                                               int64_t B, double *out, void *stream, std::string *err)
{
    try {
        ABHSF_REQUIRE(thetas && X && out && B >= 0, ABHSF_ERR_INVALID, "null argument");
        ABHSF_REQUIRE(F >= 1 && F <= kMaxF, ABHSF_ERR_UNSUPPORTED, "feature count must be in [1, 64]");
        ABHSF_REQUIRE(nm >= 1, ABHSF_ERR_INVALID, "at least one ensemble member");
        if (B == 0) return ABHSF_OK;
        k_mlp_ens_fwd<<<g(B), kT, 0, static_cast<hipStream_t>(stream)>>>(thetas, nm, F, X, B, out);
        ABHSF_HIP(hipGetLastError());
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}

extern "C" abhsf_status abhsf_mlp_backward_impl(const float *theta, int32_t F, const float *X, int64_t B,
                                                const float *H1, const float *H2, const float *out, const float *y,
                                                const float *w, float wsum, float *D1, float *D2, float *D3,
                                                float *grad, void *stream, std::string *err)
{
    try {
        ABHSF_REQUIRE(theta && X && H1 && H2 && out && y && D1 && D2 && D3 && grad && B >= 1, ABHSF_ERR_INVALID,
                      "null argument");
        ABHSF_REQUIRE(F >= 1 && F <= kMaxF, ABHSF_ERR_UNSUPPORTED, "feature count must be in [1, 64]");
        ABHSF_REQUIRE(wsum > 0.f, ABHSF_ERR_INVALID, "weights must sum to a positive value");
        hipStream_t st = static_cast<hipStream_t>(stream);
        k_mlp_delta<<<g(B), kT, 0, st>>>(theta, F, B, H1, H2, out, y, w, wsum, nullptr, D1, D2, D3);
        ABHSF_HIP(hipGetLastError());
        k_mlp_grad<<<g(layout(F).P), kT, 0, st>>>(F, X, B, H1, H2, D1, D2, D3, grad);
        ABHSF_HIP(hipGetLastError());
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}

extern "C" abhsf_status abhsf_mlp_train_impl(float *theta, int32_t F, const float *X, const float *y, const float *w,
                                             int64_t B, int64_t steps, float lr, float *loss_host, void *stream,
                                             std::string *err)
{
    try {
        ABHSF_REQUIRE(theta && X && y && B >= 1 && steps >= 0, ABHSF_ERR_INVALID, "invalid argument");
        ABHSF_REQUIRE(F >= 1 && F <= kMaxF, ABHSF_ERR_UNSUPPORTED, "feature count must be in [1, 64]");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const Layout L = layout(F);
        float wsum = (float)B;
        if (w) {
            std::vector<float> hw((size_t)B);
            ABHSF_HIP(hipMemcpyAsync(hw.data(), w, (size_t)B * 4, hipMemcpyDeviceToHost, st));
            ABHSF_HIP(hipStreamSynchronize(st));
            double s = 0;
            for (float v : hw) s += v;
            wsum = (float)s;
        }
        float *buf = nullptr;
        const size_t nfloat = (size_t)B * (4 * H + 2) + 3 * (size_t)L.P + (size_t)steps + 1;
        ABHSF_HIP(hipMalloc(&buf, nfloat * 4));
        float *H1 = buf, *H2 = H1 + B * H, *D1 = H2 + B * H, *D2 = D1 + B * H, *out = D2 + B * H, *D3 = out + B;
        float *grad = D3 + B, *m = grad + L.P, *v = m + L.P, *loss = v + L.P;
        try {
            ABHSF_HIP(hipMemsetAsync(m, 0, 2 * (size_t)L.P * 4, st));
            const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
            for (int64_t t = 1; t <= steps; t++) {
                k_mlp_fwd<<<g(B), kT, 0, st>>>(theta, F, X, B, H1, H2, out);
                k_loss<<<1, kT, 0, st>>>(out, y, w, B, wsum, nullptr, loss + (t - 1));
                k_mlp_delta<<<g(B), kT, 0, st>>>(theta, F, B, H1, H2, out, y, w, wsum, nullptr, D1, D2, D3);
                k_mlp_grad<<<g(L.P), kT, 0, st>>>(F, X, B, H1, H2, D1, D2, D3, grad);
                const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
                k_adam<<<g(L.P), kT, 0, st>>>(theta, grad, m, v, L.P, lr, b1, b2, eps, c1, c2);
                ABHSF_HIP(hipGetLastError());
            }
            if (loss_host && steps)
                ABHSF_HIP(hipMemcpyAsync(loss_host, loss, (size_t)steps * 4, hipMemcpyDeviceToHost, st));
            ABHSF_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(buf);
            throw;
        }
        ABHSF_HIP(hipFree(buf));
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}

// abhsf_mlp_train without allocation: the same kernels in the same order on
// the same data, so theta and the losses equal abhsf_mlp_train's bit for bit;
// the work buffer is the device array `scratch`.  Asynchronous on stream (no
// host synchronisation, no memset: m and v are cleared by a fill kernel), so
// it can be captured into a graph.  Calls that overlap on one GPU share the
// array and race.
extern "C" abhsf_status abhsf_mlp_train_noalloc_impl(float *theta, int32_t F, const float *X, const float *y,
                                                     const float *w, int64_t B, int64_t steps, float lr,
                                                     float *loss_dev, void *stream, std::string *err)
{
    try {
        ABHSF_REQUIRE(theta && X && y && B >= 1 && steps >= 0, ABHSF_ERR_INVALID, "invalid argument");
        ABHSF_REQUIRE(F >= 1 && F <= kMaxF, ABHSF_ERR_UNSUPPORTED, "feature count must be in [1, 64]");
        ABHSF_REQUIRE(B <= kNoAllocMaxSamples, ABHSF_ERR_UNSUPPORTED,
                      "abhsf_mlp_train_noalloc: nsamples " + std::to_string(B) + " exceeds the scratch limit of " +
                          std::to_string(kNoAllocMaxSamples) + " samples");
        ABHSF_REQUIRE(steps <= kNoAllocMaxSteps, ABHSF_ERR_UNSUPPORTED,
                      "abhsf_mlp_train_noalloc: steps " + std::to_string(steps) + " exceeds the scratch limit of " +
                          std::to_string(kNoAllocMaxSteps) + " steps");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const Layout L = layout(F);
        float *buf = scratch_on_current_device();
        float *H1 = buf, *H2 = H1 + B * H, *D1 = H2 + B * H, *D2 = D1 + B * H, *out = D2 + B * H, *D3 = out + B;
        float *grad = D3 + B, *m = grad + L.P, *v = m + L.P, *loss = v + L.P, *wsum_p = loss + steps;
        if (loss_dev) loss = loss_dev;
        const float wsum = (float)B;  // unit weights; with w the device sum replaces it
        if (w) {
            k_wsum<<<1, 64, 0, st>>>(w, B, wsum_p);
            ABHSF_HIP(hipGetLastError());
        } else {
            wsum_p = nullptr;
        }
        k_fill0<<<g(2 * L.P), kT, 0, st>>>(m, 2 * L.P);
        ABHSF_HIP(hipGetLastError());
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int64_t t = 1; t <= steps; t++) {
            k_mlp_fwd<<<g(B), kT, 0, st>>>(theta, F, X, B, H1, H2, out);
            k_loss<<<1, kT, 0, st>>>(out, y, w, B, wsum, wsum_p, loss + (t - 1));
            k_mlp_delta<<<g(B), kT, 0, st>>>(theta, F, B, H1, H2, out, y, w, wsum, wsum_p, D1, D2, D3);
            k_mlp_grad<<<g(L.P), kT, 0, st>>>(F, X, B, H1, H2, D1, D2, D3, grad);
            const float c1 = 1.f - std::pow(b1, (float)t), c2 = 1.f - std::pow(b2, (float)t);
            k_adam<<<g(L.P), kT, 0, st>>>(theta, grad, m, v, L.P, lr, b1, b2, eps, c1, c2);
            ABHSF_HIP(hipGetLastError());
        }
        return ABHSF_OK;
    } catch (const Error &e) {
        *err = e.what();
        return e.code;
    }
}
