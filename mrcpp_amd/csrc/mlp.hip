// Refinement predictor: a 12 -> 64 -> 64 -> 1 multilayer perceptron (ReLU,
// FP32) with hand-written inference, gradient and Adam kernels.
//
// Parameters, one float array (MWG_MLP_PARAMS floats):
//   W1[12][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3
// A sample's logit is z = w3 . relu(W2^T relu(W1^T x + b1) + b2) + b3 and
// p = sigmoid(z).  Loss: mean over samples of
//   wpos * y * softplus(-z) + (1 - y) * softplus(z).
//
// One wave per sample at a time, lane j = hidden unit j.  The gradient kernel
// runs a fixed number of waves; each accumulates its samples' gradients in
// registers and writes one partial vector, and the partials are summed in
// wave order, so gradients do not depend on scheduling.
#include "device_common.h"

#include <mutex>

namespace mwgdev {

namespace {

constexpr int F = MWG_MLP_FEATURES, H = MWG_MLP_HIDDEN;
constexpr int OW1 = 0, OB1 = OW1 + F * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H;
static_assert(OB3 + 1 == MWG_MLP_PARAMS, "parameter layout");
constexpr int kGradWaves = 512;

__device__ inline float softplus(float z) { return z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z)); }
__device__ inline float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

__device__ inline float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Hidden activations of one sample: a1, a2 pre-activations of lane j; h1 of
// every unit left in sh (64 floats of this wave's LDS).
__device__ inline void forward_one(const float *P, const float *x, float *sh, float &a1, float &a2, float &z) {
    const int j = threadIdx.x & 63;
    a1 = P[OB1 + j];
#pragma unroll
    for (int f = 0; f < F; f++) a1 += x[f] * P[OW1 + f * H + j];
    sh[j] = fmaxf(a1, 0.0f);
    __builtin_amdgcn_wave_barrier();
    a2 = P[OB2 + j];
#pragma unroll 8
    for (int k = 0; k < H; k++) a2 += sh[k] * P[OW2 + k * H + j];
    z = wave_sum(fmaxf(a2, 0.0f) * P[OW3 + j]) + P[OB3];
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *P, const float *X, int n, float *prob) {
    __shared__ float sh[4][H];
    const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
    for (int i = blockIdx.x * 4 + w; i < n; i += gridDim.x * 4) {
        float x[F];
#pragma unroll
        for (int f = 0; f < F; f++) x[f] = X[(size_t)i * F + f];
        float a1, a2, z;
        forward_one(P, x, sh[w], a1, a2, z);
        if (j == 0) prob[i] = sigmoid(z);
        __builtin_amdgcn_wave_barrier();
    }
}

// Per-wave gradient partials (MWG_MLP_PARAMS + 1 floats each, the last one
// the wave's loss sum) for samples i = wave, wave + kGradWaves, ...
__global__ __launch_bounds__(256) void k_mlp_grad_partial(const float *P, const float *X, const float *Y, int n, float wpos, float *partial) {
    __shared__ float sh1[4][H], sd2[4][H];
    const int w = threadIdx.x >> 6, j = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + w;
    float gW1[F], gW2[H], gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f, loss = 0.0f;
#pragma unroll
    for (int f = 0; f < F; f++) gW1[f] = 0.0f;
#pragma unroll
    for (int k = 0; k < H; k++) gW2[k] = 0.0f;
    for (int i = wave; i < n; i += kGradWaves) {
        float x[F];
#pragma unroll
        for (int f = 0; f < F; f++) x[f] = X[(size_t)i * F + f];
        const float y = Y[i];
        float a1, a2, z;
        forward_one(P, x, sh1[w], a1, a2, z);
        loss += wpos * y * softplus(-z) + (1.0f - y) * softplus(z);
        // dL/dz for this sample
        const float s = sigmoid(z);
        const float dz = wpos * y * (s - 1.0f) + (1.0f - y) * s;
        const float h2 = fmaxf(a2, 0.0f);
        gw3 += dz * h2;
        gb3 += dz;
        const float d2 = a2 > 0.0f ? dz * P[OW3 + j] : 0.0f;
        gb2 += d2;
#pragma unroll
        for (int k = 0; k < H; k++) gW2[k] += sh1[w][k] * d2; // d/dW2[k][j]
        sd2[w][j] = d2;
        __builtin_amdgcn_wave_barrier();
        float dh1 = 0.0f; // lane j as hidden unit k = j of layer 1
#pragma unroll 8
        for (int m = 0; m < H; m++) dh1 += P[OW2 + j * H + m] * sd2[w][m];
        const float d1 = a1 > 0.0f ? dh1 : 0.0f;
        gb1 += d1;
#pragma unroll
        for (int f = 0; f < F; f++) gW1[f] += x[f] * d1;
        __builtin_amdgcn_wave_barrier();
    }
    float *out = partial + (size_t)wave * (MWG_MLP_PARAMS + 1);
#pragma unroll
    for (int f = 0; f < F; f++) out[OW1 + f * H + j] = gW1[f];
    out[OB1 + j] = gb1;
#pragma unroll
    for (int k = 0; k < H; k++) out[OW2 + k * H + j] = gW2[k];
    out[OB2 + j] = gb2;
    out[OW3 + j] = gw3;
    const float g3 = wave_sum(j == 0 ? gb3 : 0.0f);
    const float ls = wave_sum(j == 0 ? loss : 0.0f);
    if (j == 0) {
        out[OB3] = g3;
        out[MWG_MLP_PARAMS] = ls;
    }
}

// grad[p] = sum over waves (in order) of partial[w][p] / n; grad[PARAMS] = mean loss.
__global__ __launch_bounds__(256) void k_mlp_grad_reduce(const float *partial, int n, float *grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p > MWG_MLP_PARAMS) return;
    float s = 0.0f;
    for (int w = 0; w < kGradWaves; w++) s += partial[(size_t)w * (MWG_MLP_PARAMS + 1) + p];
    grad[p] = n > 0 ? s / (float)n : 0.0f;
}

__global__ __launch_bounds__(256) void k_mlp_adam(float *P, const float *grad, float *m, float *v, int step, float lr, float b1, float b2, float eps) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= MWG_MLP_PARAMS) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float mh = mm / (1.0f - powf(b1, (float)step));
    const float vh = vv / (1.0f - powf(b2, (float)step));
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

Buffer g_partial;

} // namespace

} // namespace mwgdev

// Scratch of mwg_mlp_gradient_noalloc: the kGradWaves per-wave partial
// vectors, 512 x 5058 floats = 10,358,784 bytes per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in, never the
// device memory: kernels get the address hipGetSymbolAddress returns for the
// current device (scratch_dev, filled by mwgdev::mlp_scratch_setup).
static __device__ __align__(256) float scratch[mwgdev::kGradWaves * (MWG_MLP_PARAMS + 1)];

namespace mwgdev {

namespace {
constexpr int kScratchDevices = 64;
float *scratch_dev[kScratchDevices] = {};
std::mutex scratch_mutex;
} // namespace

// Device address of `scratch` on `device`, which must be current: looked up
// once per device (called by mwg_init).
mwg_status mlp_scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return fail(MWG_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(scratch_mutex);
    if (scratch_dev[device]) return MWG_OK;
    void *p = nullptr;
    MWG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    scratch_dev[device] = static_cast<float *>(p);
    return MWG_OK;
}

} // namespace mwgdev

using namespace mwgdev;

extern "C" {

mwg_status mwg_mlp_forward(const float *params, const float *features, int n, float *prob) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (n < 0) return fail(MWG_ERR_ARG, "negative sample count");
    if (n == 0) return MWG_OK;
    if (!params || !features || !prob) return fail(MWG_ERR_ARG, "null argument");
    const int blocks = std::min((n + 3) / 4, 4096);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(256), 0, 0, params, features, n, prob);
    MWG_LAUNCHED();
    return MWG_OK;
}

mwg_status mwg_mlp_gradient(const float *params, const float *features, const float *labels, int n, float pos_weight, float *grad) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (n < 0) return fail(MWG_ERR_ARG, "negative sample count");
    if (!params || !grad || (n > 0 && (!features || !labels))) return fail(MWG_ERR_ARG, "null argument");
    if (g_partial.ensure(sizeof(float) * (size_t)kGradWaves * (MWG_MLP_PARAMS + 1)) != MWG_OK) return MWG_ERR_NOMEM;
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3(kGradWaves / 4), dim3(256), 0, 0, params, features, labels, n, pos_weight, g_partial.as<float>());
    MWG_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((MWG_MLP_PARAMS + 1 + 255) / 256), dim3(256), 0, 0, g_partial.as<float>(), n, grad);
    MWG_LAUNCHED();
    return MWG_OK;
}

mwg_status mwg_mlp_gradient_noalloc(const float *params, const float *features, const float *labels, int n, float pos_weight, float *grad) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (n < 0) return fail(MWG_ERR_ARG, "negative sample count");
    if (!params || !grad || (n > 0 && (!features || !labels))) return fail(MWG_ERR_ARG, "null argument");
    int dev = -1;
    MWG_HIP(hipGetDevice(&dev));
    float *part = nullptr;
    {
        std::lock_guard<std::mutex> lk(scratch_mutex);
        if (dev >= 0 && dev < kScratchDevices) part = scratch_dev[dev];
    }
    if (!part) return fail(MWG_ERR_ARG, "mwg_init has not been called for the current device");
    // every partial vector is written whole by each call: nothing to clear
    hipLaunchKernelGGL(k_mlp_grad_partial, dim3(kGradWaves / 4), dim3(256), 0, 0, params, features, labels, n, pos_weight, part);
    MWG_LAUNCHED();
    hipLaunchKernelGGL(k_mlp_grad_reduce, dim3((MWG_MLP_PARAMS + 1 + 255) / 256), dim3(256), 0, 0, part, n, grad);
    MWG_LAUNCHED();
    return MWG_OK;
}

mwg_status mwg_mlp_adam(float *params, const float *grad, float *m, float *v, int step, float lr, float beta1, float beta2, float eps) {
    if (need_init() != MWG_OK) return MWG_ERR_ARG;
    if (!params || !grad || !m || !v || step < 1) return fail(MWG_ERR_ARG, "bad Adam arguments");
    hipLaunchKernelGGL(k_mlp_adam, dim3((MWG_MLP_PARAMS + 255) / 256), dim3(256), 0, 0, params, grad, m, v, step, lr, beta1, beta2, eps);
    MWG_LAUNCHED();
    return MWG_OK;
}

} // extern "C"
