// Width-controller MLP on gfx950: forward, gradients and Adam, FP32.
//
// Network: 12 inputs -> 64 ReLU -> 64 ReLU -> 1 logit (probability of
// widening = sigmoid(logit)); loss = mean binary cross-entropy on the logits.
// Parameters are one flat FP32 array in the order
//   W1 [64 x 12] row-major, b1 [64], W2 [64 x 64] row-major, b2 [64],
//   W3 [64], b3 [1]            (kMlpParams = 5057 values).
// Inputs are normalized on the device with per-feature (mean, inverse std).
//
// The batch is split into chunks; a block walks its chunk sample by sample
// with the whole network in LDS, and each thread owns a fixed set of
// parameters whose gradient contributions it accumulates in registers in
// sample order.  Chunk partials are added in chunk order by a second kernel,
// so gradients are deterministic.
#include "kernels.hpp"

#include <algorithm>

namespace amflsx {
namespace {

constexpr int kIn = kMlpInputs, kH = kMlpHidden;
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH, ob3 = oW3 + kH;
static_assert(ob3 + 1 == kMlpParams, "parameter layout");
constexpr int kThreads = 256;
constexpr int kPerThread = (kMlpParams + kThreads - 1) / kThreads;  // 20

struct Sample {
    float x[kIn];
    float h1[kH], h2[kH], d1[kH], d2[kH];
    float z, dz;
};

// Forward pass of one sample held in shared memory (all threads call).
__device__ void forward_one(const float* P, Sample& s) {
    const int t = threadIdx.x;
    if (t < kH) {
        float a = P[ob1 + t];
        for (int i = 0; i < kIn; ++i) a = fmaf(P[oW1 + t * kIn + i], s.x[i], a);
        s.h1[t] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    if (t < kH) {
        float a = P[ob2 + t];
        for (int i = 0; i < kH; ++i) a = fmaf(P[oW2 + t * kH + i], s.h1[i], a);
        s.h2[t] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    if (t < kWave) {
        float a = P[oW3 + t] * s.h2[t];
#pragma unroll
        for (int off = kWave / 2; off > 0; off >>= 1) a += __shfl_xor(a, off);
        if (t == 0) s.z = a + P[ob3];
    }
    __syncthreads();
}

__device__ void load_sample(const float* X, int64_t i, const float* norm, Sample& s) {
    if (threadIdx.x < kIn) s.x[threadIdx.x] = (X[i * kIn + threadIdx.x] - norm[threadIdx.x]) * norm[kIn + threadIdx.x];
    __syncthreads();
}

__global__ __launch_bounds__(kThreads) void mlp_forward_kernel(const float* __restrict__ params,
                                                               const float* __restrict__ norm, int64_t n,
                                                               const float* __restrict__ X, float* __restrict__ logits) {
    __shared__ float P[kMlpParams];
    __shared__ Sample s;
    for (int i = threadIdx.x; i < kMlpParams; i += kThreads) P[i] = params[i];
    __syncthreads();
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        load_sample(X, i, norm, s);
        forward_one(P, s);
        if (threadIdx.x == 0) logits[i] = s.z;
        __syncthreads();
    }
}

// Per-chunk gradient partials and loss partials.
__global__ __launch_bounds__(kThreads) void mlp_grad_kernel(const float* __restrict__ params,
                                                            const float* __restrict__ norm, int64_t n,
                                                            int64_t chunk, const float* __restrict__ X,
                                                            const float* __restrict__ y, float* __restrict__ gpart,
                                                            float* __restrict__ lpart) {
    __shared__ float P[kMlpParams];
    __shared__ Sample s;
    const int t = threadIdx.x;
    for (int i = t; i < kMlpParams; i += kThreads) P[i] = params[i];
    __syncthreads();
    float g[kPerThread];
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) g[u] = 0.0f;
    float loss = 0.0f;
    const float inv_n = 1.0f / static_cast<float>(n);
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * chunk;
    const int64_t i1 = std::min<int64_t>(n, i0 + chunk);
    for (int64_t i = i0; i < i1; ++i) {
        load_sample(X, i, norm, s);
        forward_one(P, s);
        if (t == 0) {
            const float z = s.z, yy = y[i];
            // stable BCE with logits: max(z,0) - y z + log(1 + exp(-|z|))
            loss += fmaxf(z, 0.0f) - yy * z + log1pf(expf(-fabsf(z)));
            s.dz = (1.0f / (1.0f + expf(-z)) - yy) * inv_n;
        }
        __syncthreads();
        if (t < kH) s.d2[t] = s.h2[t] > 0.0f ? s.dz * P[oW3 + t] : 0.0f;
        __syncthreads();
        if (t < kH) {
            float a = 0.0f;
            for (int j = 0; j < kH; ++j) a = fmaf(P[oW2 + j * kH + t], s.d2[j], a);
            s.d1[t] = s.h1[t] > 0.0f ? a : 0.0f;
        }
        __syncthreads();
        // parameter p = t + u * kThreads
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            const int p = t + u * kThreads;
            float d = 0.0f;
            if (p < ob1) d = s.d1[p / kIn] * s.x[p % kIn];
            else if (p < oW2) d = s.d1[p - ob1];
            else if (p < ob2) d = s.d2[(p - oW2) / kH] * s.h1[(p - oW2) % kH];
            else if (p < oW3) d = s.d2[p - ob2];
            else if (p < ob3) d = s.dz * s.h2[p - oW3];
            else if (p == ob3) d = s.dz;
            g[u] += d;
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kPerThread; ++u) {
        const int p = t + u * kThreads;
        if (p < kMlpParams) gpart[static_cast<int64_t>(blockIdx.x) * kMlpParams + p] = g[u];
    }
    if (t == 0) lpart[blockIdx.x] = loss;
}

// grad[p] = sum over chunks (chunk order); loss = sum / n (thread 0 of block 0).
__global__ void mlp_reduce_kernel(int nchunks, int64_t n, const float* __restrict__ gpart,
                                  const float* __restrict__ lpart, float* __restrict__ grad, float* __restrict__ loss) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < kMlpParams) {
        float a = 0.0f;
        for (int c = 0; c < nchunks; ++c) a += gpart[static_cast<int64_t>(c) * kMlpParams + p];
        grad[p] = a;
    }
    if (p == 0) {
        float l = 0.0f;
        for (int c = 0; c < nchunks; ++c) l += lpart[c];
        *loss = l / static_cast<float>(n);
    }
}

// Adam (Kingma and Ba, 2015) with bias correction; step = 1, 2, ...
__global__ void mlp_adam_kernel(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                                const float* __restrict__ grad, float lr, float beta1, float beta2, float eps,
                                float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float mm = beta1 * m[p] + (1.0f - beta1) * g;
    const float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

}  // namespace

void mlp_forward(hipStream_t s, const float* params, const float* norm, int64_t n, const float* X, float* logits) {
    if (n <= 0) return;
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(n, 1024));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(blocks), dim3(kThreads), 0, s, params, norm, n, X, logits);
    AMFLS_LAUNCH_CHECK();
}

void mlp_gradient(hipStream_t s, const float* params, const float* norm, int64_t n, const float* X, const float* y,
                  float* grad, float* loss, Scratch& scratch) {
    if (n <= 0) throw std::invalid_argument("empty training batch");
    const int64_t nchunks = std::min<int64_t>(n, 128);
    const int64_t chunk = ceil_div(n, nchunks);
    const int64_t nc = ceil_div(n, chunk);
    scratch.ensure(static_cast<size_t>(nc * (kMlpParams + 1)) * sizeof(float));
    float* gpart = static_cast<float*>(scratch.ptr);
    float* lpart = gpart + nc * kMlpParams;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(static_cast<unsigned>(nc)), dim3(kThreads), 0, s, params, norm, n, chunk,
                       X, y, gpart, lpart);
    AMFLS_LAUNCH_CHECK();
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3(static_cast<unsigned>(ceil_div(kMlpParams, kThreads))), dim3(kThreads),
                       0, s, static_cast<int>(nc), n, gpart, lpart, grad, loss);
    AMFLS_LAUNCH_CHECK();
}

void mlp_adam(hipStream_t s, float* params, float* m, float* v, const float* grad, float lr, int step) {
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(static_cast<unsigned>(ceil_div(kMlpParams, kThreads))), dim3(kThreads), 0,
                       s, params, m, v, grad, lr, beta1, beta2, eps, c1, c2);
    AMFLS_LAUNCH_CHECK();
}

}  // namespace amflsx
