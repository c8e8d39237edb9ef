// The per-level BFS variant predictor: an FP32 MLP 8 -> 32 -> 32 -> K
// (ReLU hidden layers, linear outputs = predicted log2 level time of each
// variant) with hand-written inference, gradient and training kernels.
//
// The role is that of the reference's per-level model lookup in
// SwitchImplementation (ImplementationTemplate.hpp:545-908), whose models
// are decision trees trained offline in Haskell (benchmark-analysis); here
// the model is trained on the GPU from exhaustive per-level timings and
// evaluated on the GPU inside the BFS level loop.
//
// Packed parameter layout (floats, row-major):
//   mean[8] std[8] W1[32][8] b1[32] W2[32][32] b2[32] W3[K][32] b3[K]
// mean/std normalise the features and are not trained.  Training is
// minibatch Adam on L = mean over the batch and the K outputs of
// (out - y)^2, run as ONE workgroup of 1024 threads that keeps the
// parameters, both Adam moments and the batch activations in LDS for the
// whole run; every sum has a fixed order, so results are reproducible bit
// for bit.

#include "common.hpp"

#include <algorithm>

namespace bw {
namespace {

constexpr int kF = BW_BFS_FEATURES;
constexpr int kH = 32;
constexpr int kMaxK = 16;
constexpr int kMaxBatch = BW_MLP_MAX_BATCH;
constexpr int kThreads = 1024;

__host__ __device__ constexpr int packed_len(int K) { return 2 * kF + kH * kF + kH + kH * kH + kH + K * kH + K; }
constexpr int kMaxParams = packed_len(kMaxK);

struct Layout {
    int W1, b1, W2, b2, W3, b3, end;
    __host__ __device__ explicit Layout(int K)
    {
        W1 = 2 * kF;
        b1 = W1 + kH * kF;
        W2 = b1 + kH;
        b2 = W2 + kH * kH;
        W3 = b2 + kH;
        b3 = W3 + K * kH;
        end = b3 + K;
    }
};

// Thread per sample; parameters read through the caches.
__global__ __launch_bounds__(256) void mlp_forward_kernel(const float *__restrict__ w, int K,
                                                          const float *__restrict__ X, uint64_t n,
                                                          float *__restrict__ out)
{
    const Layout L(K);
    for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < n;
         s += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        float x[kF], h1[kH], h2[kH];
        for (int f = 0; f < kF; f++) x[f] = (X[s * kF + f] - w[f]) / w[kF + f];
        for (int j = 0; j < kH; j++) {
            float a = w[L.b1 + j];
            for (int f = 0; f < kF; f++) a += w[L.W1 + j * kF + f] * x[f];
            h1[j] = fmaxf(a, 0.0f);
        }
        for (int j = 0; j < kH; j++) {
            float a = w[L.b2 + j];
            for (int i = 0; i < kH; i++) a += w[L.W2 + j * kH + i] * h1[i];
            h2[j] = fmaxf(a, 0.0f);
        }
        for (int k = 0; k < K; k++) {
            float a = w[L.b3 + k];
            for (int j = 0; j < kH; j++) a += w[L.W3 + k * kH + j] * h2[j];
            out[s * K + k] = a;
        }
    }
}

// LDS working set of one workgroup.
struct Shared {
    float P[kMaxParams];     // packed parameters
    float M1[kMaxParams];    // Adam first moment
    float M2[kMaxParams];    // Adam second moment
    float XN[kMaxBatch][kF]; // normalised inputs
    float H1[kMaxBatch][kH];
    float H2[kMaxBatch][kH];
    float G[kMaxBatch][kMaxK];  // dL/dout
    float D2[kMaxBatch][kH];    // dL/dpre-activation of layer 2
    float D1[kMaxBatch][kH];    // dL/dpre-activation of layer 1
    float partial[kThreads / kWave];
};

// Forward and backward over the batch `rows` (sample ids) of size B;
// returns the batch loss sum of (out - y)^2 in thread 0.  Gradients are
// left in G, D2, D1 (scaled by 1 / (B K)).
__device__ float batch_pass(Shared &S, const Layout &L, int K, const float *X, const float *Y,
                            const uint32_t *rows, int B)
{
    const int t = threadIdx.x;
    for (int i = t; i < B * kF; i += kThreads) {
        const int s = i / kF, f = i % kF;
        S.XN[s][f] = (X[static_cast<uint64_t>(rows[s]) * kF + f] - S.P[f]) / S.P[kF + f];
    }
    __syncthreads();
    for (int i = t; i < B * kH; i += kThreads) {
        const int s = i / kH, j = i % kH;
        float a = S.P[L.b1 + j];
        for (int f = 0; f < kF; f++) a += S.P[L.W1 + j * kF + f] * S.XN[s][f];
        S.H1[s][j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    for (int i = t; i < B * kH; i += kThreads) {
        const int s = i / kH, j = i % kH;
        float a = S.P[L.b2 + j];
        for (int q = 0; q < kH; q++) a += S.P[L.W2 + j * kH + q] * S.H1[s][q];
        S.H2[s][j] = fmaxf(a, 0.0f);
    }
    __syncthreads();
    const float scale = 2.0f / static_cast<float>(B * K);
    float lsum = 0.0f;
    for (int i = t; i < B * K; i += kThreads) {
        const int s = i / K, k = i % K;
        float a = S.P[L.b3 + k];
        for (int j = 0; j < kH; j++) a += S.P[L.W3 + k * kH + j] * S.H2[s][j];
        const float r = a - Y[static_cast<uint64_t>(rows[s]) * K + k];
        lsum += r * r;
        S.G[s][k] = scale * r;
    }
    lsum = wave_sum_f32(lsum);
    if ((t & (kWave - 1)) == 0) S.partial[t / kWave] = lsum;
    __syncthreads();
    for (int i = t; i < B * kH; i += kThreads) {
        const int s = i / kH, j = i % kH;
        float a = 0.0f;
        for (int k = 0; k < K; k++) a += S.G[s][k] * S.P[L.W3 + k * kH + j];
        S.D2[s][j] = S.H2[s][j] > 0.0f ? a : 0.0f;
    }
    __syncthreads();
    for (int i = t; i < B * kH; i += kThreads) {
        const int s = i / kH, q = i % kH;
        float a = 0.0f;
        for (int j = 0; j < kH; j++) a += S.D2[s][j] * S.P[L.W2 + j * kH + q];
        S.D1[s][q] = S.H1[s][q] > 0.0f ? a : 0.0f;
    }
    __syncthreads();
    float loss = 0.0f;
    if (t == 0)
        for (int w = 0; w < kThreads / kWave; w++) loss += S.partial[w];
    return loss;
}

// Gradient of trainable parameter p (packed index >= 2 * kF) over a batch.
__device__ float param_grad(const Shared &S, const Layout &L, int p, int B)
{
    float g = 0.0f;
    if (p < L.b1) {
        const int j = (p - L.W1) / kF, f = (p - L.W1) % kF;
        for (int s = 0; s < B; s++) g += S.D1[s][j] * S.XN[s][f];
    } else if (p < L.W2) {
        const int j = p - L.b1;
        for (int s = 0; s < B; s++) g += S.D1[s][j];
    } else if (p < L.b2) {
        const int j = (p - L.W2) / kH, q = (p - L.W2) % kH;
        for (int s = 0; s < B; s++) g += S.D2[s][j] * S.H1[s][q];
    } else if (p < L.W3) {
        const int j = p - L.b2;
        for (int s = 0; s < B; s++) g += S.D2[s][j];
    } else if (p < L.b3) {
        const int k = (p - L.W3) / kH, j = (p - L.W3) % kH;
        for (int s = 0; s < B; s++) g += S.G[s][k] * S.H2[s][j];
    } else {
        const int k = p - L.b3;
        for (int s = 0; s < B; s++) g += S.G[s][k];
    }
    return g;
}

// Gradient of L over the n <= kMaxBatch samples, for checking.
__global__ __launch_bounds__(kThreads) void mlp_grad_kernel(const float *w, int K, const float *X, const float *Y,
                                                            uint32_t n, float *grad, float *loss)
{
    __shared__ Shared S;
    __shared__ uint32_t rows[kMaxBatch];
    const Layout L(K);
    for (int i = threadIdx.x; i < L.end; i += kThreads) S.P[i] = w[i];
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) rows[i] = i;
    __syncthreads();
    const int B = static_cast<int>(n);
    const float lsum = batch_pass(S, L, K, X, Y, rows, B);
    for (int p = 2 * kF + threadIdx.x; p < L.end; p += kThreads) grad[p] = param_grad(S, L, p, B);
    for (int p = threadIdx.x; p < 2 * kF; p += kThreads) grad[p] = 0.0f;
    if (threadIdx.x == 0) *loss = lsum / static_cast<float>(B * K);
}

struct TrainArgs {
    float *w;
    int K;
    const float *X;
    const float *Y;
    uint32_t n;
    uint32_t epochs;
    uint32_t batch;
    uint32_t stride;  // coprime to n: epoch order is (offset + i * stride) mod n
    uint32_t seed;
    float lr;
    float beta1, beta2, eps;
    float *epoch_loss;  // optional, [epochs]
};

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(kThreads) void mlp_train_kernel(TrainArgs a)
{
    __shared__ Shared S;
    __shared__ uint32_t rows[kMaxBatch];
    const Layout L(a.K);
    const int t = threadIdx.x;
    for (int i = t; i < L.end; i += kThreads) {
        S.P[i] = a.w[i];
        S.M1[i] = 0.0f;
        S.M2[i] = 0.0f;
    }
    __syncthreads();
    uint32_t step = 0;
    for (uint32_t e = 0; e < a.epochs; e++) {
        const uint64_t offset = mix64(static_cast<uint64_t>(a.seed) * 0x9e3779b97f4a7c15ull + e) % a.n;
        float esum = 0.0f;
        for (uint32_t b0 = 0; b0 < a.n; b0 += a.batch) {
            const int B = static_cast<int>(min(a.batch, a.n - b0));
            for (int i = t; i < B; i += kThreads)
                rows[i] = static_cast<uint32_t>((offset + static_cast<uint64_t>(b0 + i) * a.stride) % a.n);
            __syncthreads();
            const float lsum = batch_pass(S, L, a.K, a.X, a.Y, rows, B);
            esum += lsum;  // meaningful in thread 0 only
            step++;
            const float c1 = 1.0f - powf(a.beta1, static_cast<float>(step));
            const float c2 = 1.0f - powf(a.beta2, static_cast<float>(step));
            // Every gradient reads activations only, so parameters can be
            // updated in place.
            for (int p = 2 * kF + t; p < L.end; p += kThreads) {
                const float g = param_grad(S, L, p, B);
                const float m1 = a.beta1 * S.M1[p] + (1.0f - a.beta1) * g;
                const float m2 = a.beta2 * S.M2[p] + (1.0f - a.beta2) * g * g;
                S.M1[p] = m1;
                S.M2[p] = m2;
                S.P[p] -= a.lr * (m1 / c1) / (sqrtf(m2 / c2) + a.eps);
            }
            __syncthreads();
        }
        if (t == 0 && a.epoch_loss) a.epoch_loss[e] = esum / static_cast<float>(a.n * a.K);
    }
    for (int i = t; i < L.end; i += kThreads) a.w[i] = S.P[i];
}

int check_model(int K)
{
    if (K < 1 || K > kMaxK) return fail(BW_HIP_ERR_ARG, "MLP output count must be in 1.." + std::to_string(kMaxK));
    return BW_HIP_OK;
}

uint32_t gcd32(uint32_t x, uint32_t y)
{
    while (y) {
        const uint32_t r = x % y;
        x = y;
        y = r;
    }
    return x;
}

}  // namespace
}  // namespace bw

using namespace bw;

extern "C" int32_t bw_mlp_params(int32_t outputs) { return packed_len(outputs); }

extern "C" int bw_mlp_forward(const float *weights, int32_t outputs, const float *X, uint64_t n, float *out,
                              void *stream)
{
    BW_TRY(check_model(outputs));
    if (!weights || (n && (!X || !out))) return fail(BW_HIP_ERR_ARG, "bw_mlp_forward: NULL pointer");
    if (n == 0) return BW_HIP_OK;
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), weights,
                       outputs, X, n, out);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_mlp_grad(const float *weights, int32_t outputs, const float *X, const float *Y, uint32_t n,
                           float *grad, float *loss, void *stream)
{
    BW_TRY(check_model(outputs));
    if (!weights || !X || !Y || !grad || !loss) return fail(BW_HIP_ERR_ARG, "bw_mlp_grad: NULL pointer");
    if (n < 1 || n > kMaxBatch)
        return fail(BW_HIP_ERR_ARG, "bw_mlp_grad takes 1.." + std::to_string(kMaxBatch) + " samples");
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(1), dim3(kThreads), 0, static_cast<hipStream_t>(stream), weights,
                       outputs, X, Y, n, grad, loss);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}

extern "C" int bw_mlp_train(float *weights, int32_t outputs, const float *X, const float *Y, uint32_t n,
                            uint32_t epochs, uint32_t batch, float lr, uint32_t seed, float *epoch_loss,
                            void *stream)
{
    BW_TRY(check_model(outputs));
    if (!weights || !X || !Y) return fail(BW_HIP_ERR_ARG, "bw_mlp_train: NULL pointer");
    if (n < 1) return fail(BW_HIP_ERR_ARG, "bw_mlp_train needs at least one sample");
    if (batch < 1 || batch > static_cast<uint32_t>(kMaxBatch))
        return fail(BW_HIP_ERR_ARG, "batch must be in 1.." + std::to_string(kMaxBatch));
    if (!(lr > 0.0f)) return fail(BW_HIP_ERR_ARG, "learning rate must be positive");
    TrainArgs a;
    a.w = weights;
    a.K = outputs;
    a.X = X;
    a.Y = Y;
    a.n = n;
    a.epochs = epochs;
    a.batch = batch;
    // A stride near n * 0.618 coprime to n spreads consecutive batches.
    uint32_t stride = std::max<uint32_t>(1, static_cast<uint32_t>(static_cast<double>(n) * 0.6180339887));
    while (gcd32(stride, n) != 1) stride++;
    a.stride = stride % n ? stride % n : 1;
    a.seed = seed;
    a.lr = lr;
    a.beta1 = 0.9f;
    a.beta2 = 0.999f;
    a.eps = 1e-7f;
    a.epoch_loss = epoch_loss;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kThreads), 0, static_cast<hipStream_t>(stream), a);
    BW_LAUNCH_CHECK();
    return BW_HIP_OK;
}
