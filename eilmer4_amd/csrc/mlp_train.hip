// mlp_train.hip -- training of the temperature-guess network on the device
// with this library's own forward and backward kernels (FP32).
//
// Network (inputs normalised on the host): z[8] (7 features + 1 zero pad)
//   h1 = silu(W1 z + b1)   W1 [64][8]
//   h2 = silu(W2 h1 + b2)  W2 [64][64]
//   h3 = silu(W3 h2 + b3)  W3 [64][64]
//   y  = w4 . h3 + b4      w4 [64]
// Flat parameter layout (floats): W1, b1, W2, b2, W3, b3, w4, b4 (8961 values).
// Loss: mean squared error of y against the normalised Newton temperature.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/e4hip.h"

namespace {

constexpr int NIN = 8, NH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH, OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH;
constexpr int NPARAM = OFF_B4 + 1;
static_assert(NPARAM == E4HIP_MLP_NPARAM, "parameter count");

__device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }
__device__ __forceinline__ float dsilu(float x)
{
    float s = 1.0f / (1.0f + expf(-x));
    return s * (1.0f + x * (1.0f - s));
}

// pre[i][j] = b[j] + sum_k X[i][k] W[j][k]; act[i][j] = silu(pre)
template <int K>
__global__ __launch_bounds__(256) void fwd_layer(const float* __restrict__ X, const float* __restrict__ W,
                                                 const float* __restrict__ b, float* __restrict__ pre,
                                                 float* __restrict__ act, int n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * NH) return;
    const long i = t / NH;
    const int j = (int)(t % NH);
    float acc = b[j];
    const float* x = X + i * K;
    const float* w = W + j * K;
#pragma unroll 8
    for (int k = 0; k < K; ++k) acc += x[k] * w[k];
    pre[t] = acc;
    act[t] = silu(acc);
}

// y[i] = b4 + sum_j h3[i][j] w4[j]; dy = 2 (y - t) / n; loss accumulates sum (y - t)^2
__global__ __launch_bounds__(256) void fwd_out(const float* __restrict__ h3, const float* __restrict__ w4,
                                               const float* __restrict__ b4, const float* __restrict__ target,
                                               float* __restrict__ y, float* __restrict__ dy, float* loss, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (i < n) {
        float acc = b4[0];
        for (int j = 0; j < NH; ++j) acc += h3[(long)i * NH + j] * w4[j];
        y[i] = acc;
        if (target) {
            float e = acc - target[i];
            dy[i] = 2.0f * e / (float)n;
            e2 = e * e;
        }
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off);
        if (__lane_id() == 0) atomicAdd(loss, e2);
    }
}

// output layer backward: g3[i][j] = dy[i] w4[j] silu'(pre3[i][j])
__global__ __launch_bounds__(256) void bwd_out(const float* __restrict__ dy, const float* __restrict__ w4,
                                               const float* __restrict__ pre3, float* __restrict__ g3, int n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * NH) return;
    const long i = t / NH;
    const int j = (int)(t % NH);
    g3[t] = dy[i] * w4[j] * dsilu(pre3[t]);
}

// hidden backward: gprev[i][k] = (sum_j g[i][j] W[j][k]) silu'(preprev[i][k])
__global__ __launch_bounds__(256) void bwd_hidden(const float* __restrict__ g, const float* __restrict__ W,
                                                  const float* __restrict__ preprev, float* __restrict__ gprev, int n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n * NH) return;
    const long i = t / NH;
    const int k = (int)(t % NH);
    float acc = 0.0f;
    for (int j = 0; j < NH; ++j) acc += g[i * NH + j] * W[j * NH + k];
    gprev[t] = acc * dsilu(preprev[t]);
}

// weight gradients of one layer: dW[j][k] = sum_i g[i][j] X[i][k], db[j] = sum_i g[i][j]
// one block per output unit j, threads over k and a slice of samples
template <int K>
__global__ __launch_bounds__(256) void bwd_weights(const float* __restrict__ g, const float* __restrict__ X,
                                                   float* __restrict__ dW, float* __restrict__ db, int n)
{
    __shared__ float red[256];
    const int j = blockIdx.x;
    const int tid = threadIdx.x;
    const int k = tid % K;            // K divides 256
    const int slice = tid / K, nslice = 256 / K;
    float acc = 0.0f, accb = 0.0f;
    for (int i = slice; i < n; i += nslice) {
        float gi = g[(long)i * NH + j];
        acc += gi * X[(long)i * K + k];
        if (k == 0) accb += gi;
    }
    red[tid] = acc;
    __syncthreads();
    if (tid < K) {
        float s = 0.0f;
        for (int q = 0; q < nslice; ++q) s += red[q * K + tid];
        dW[j * K + tid] = s;
    }
    __syncthreads();
    red[tid] = (k == 0) ? accb : 0.0f;
    __syncthreads();
    if (tid == 0) {
        float s = 0.0f;
        for (int q = 0; q < nslice; ++q) s += red[q * K];
        db[j] = s;
    }
}

// output weights: dw4[j] = sum_i dy[i] h3[i][j], db4 = sum_i dy[i]
__global__ __launch_bounds__(256) void bwd_out_weights(const float* __restrict__ dy, const float* __restrict__ h3,
                                                       float* __restrict__ dw4, float* __restrict__ db4, int n)
{
    __shared__ float red[256];
    const int j = blockIdx.x;          // 0..NH (NH = bias)
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        acc += dy[i] * ((j < NH) ? h3[(long)i * NH + j] : 1.0f);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (j < NH) dw4[j] = red[0]; else db4[0] = red[0];
    }
}

__global__ void adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                            float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float gt = g[t];
    float mt = b1 * m[t] + (1.0f - b1) * gt;
    float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

struct Work {
    int cap = 0;
    float *X = nullptr, *T = nullptr, *pre1 = nullptr, *h1 = nullptr, *pre2 = nullptr, *h2 = nullptr, *pre3 = nullptr,
          *h3 = nullptr, *y = nullptr, *dy = nullptr, *g1 = nullptr, *g2 = nullptr, *g3 = nullptr, *loss = nullptr;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    int alloc(int n)
    {
        if (n <= cap) return 0;
        release();
        size_t a = (size_t)n * NH * sizeof(float);
        if (hipMalloc(&X, (size_t)n * NIN * sizeof(float)) || hipMalloc(&T, n * sizeof(float)) ||
            hipMalloc(&pre1, a) || hipMalloc(&h1, a) || hipMalloc(&pre2, a) || hipMalloc(&h2, a) ||
            hipMalloc(&pre3, a) || hipMalloc(&h3, a) || hipMalloc(&g1, a) || hipMalloc(&g2, a) || hipMalloc(&g3, a) ||
            hipMalloc(&y, n * sizeof(float)) || hipMalloc(&dy, n * sizeof(float)) || hipMalloc(&loss, sizeof(float)) ||
            hipMalloc(&params, NPARAM * sizeof(float)) || hipMalloc(&grads, NPARAM * sizeof(float)) ||
            hipMalloc(&m, NPARAM * sizeof(float)) || hipMalloc(&v, NPARAM * sizeof(float)))
            return -1;
        cap = n;
        return 0;
    }
    void release()
    {
        float** all[] = {&X, &T, &pre1, &h1, &pre2, &h2, &pre3, &h3, &y, &dy, &g1, &g2, &g3, &loss, &params, &grads, &m, &v};
        for (float** p : all) if (*p) { hipFree(*p); *p = nullptr; }
        cap = 0;
    }
};

// forward + backward of one batch already on the device; grads accumulate into w.grads
void fwd_bwd(Work& w, const float* X, const float* T, int n, hipStream_t s)
{
    const float* P = w.params;
    float* G = w.grads;
    const dim3 b(256);
    const dim3 gh((unsigned)(((long)n * NH + 255) / 256)), gn((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(fwd_layer<NIN>, gh, b, 0, s, X, P + OFF_W1, P + OFF_B1, w.pre1, w.h1, n);
    hipLaunchKernelGGL(fwd_layer<NH>, gh, b, 0, s, w.h1, P + OFF_W2, P + OFF_B2, w.pre2, w.h2, n);
    hipLaunchKernelGGL(fwd_layer<NH>, gh, b, 0, s, w.h2, P + OFF_W3, P + OFF_B3, w.pre3, w.h3, n);
    hipMemsetAsync(w.loss, 0, sizeof(float), s);
    hipLaunchKernelGGL(fwd_out, gn, b, 0, s, w.h3, P + OFF_W4, P + OFF_B4, T, w.y, w.dy, w.loss, n);
    hipLaunchKernelGGL(bwd_out, gh, b, 0, s, w.dy, P + OFF_W4, w.pre3, w.g3, n);
    hipLaunchKernelGGL(bwd_out_weights, dim3(NH + 1), b, 0, s, w.dy, w.h3, G + OFF_W4, G + OFF_B4, n);
    hipLaunchKernelGGL(bwd_weights<NH>, dim3(NH), b, 0, s, w.g3, w.h2, G + OFF_W3, G + OFF_B3, n);
    hipLaunchKernelGGL(bwd_hidden, gh, b, 0, s, w.g3, P + OFF_W3, w.pre2, w.g2, n);
    hipLaunchKernelGGL(bwd_weights<NH>, dim3(NH), b, 0, s, w.g2, w.h1, G + OFF_W2, G + OFF_B2, n);
    hipLaunchKernelGGL(bwd_hidden, gh, b, 0, s, w.g2, P + OFF_W2, w.pre1, w.g1, n);
    hipLaunchKernelGGL(bwd_weights<NIN>, dim3(NH), b, 0, s, w.g1, X, G + OFF_W1, G + OFF_B1, n);
}

Work g_work;

} // namespace

extern "C" {

int e4hip_mlp_grad(const float* params, const float* X, const float* T, int n, float* loss, float* grads)
{
    if (!params || !X || !T || n <= 0 || !grads) return E4HIP_ERR_ARG;
    if (g_work.alloc(n)) return E4HIP_ERR_HIP;
    hipStream_t s = (hipStream_t)e4hip_stream();
    hipMemcpyAsync(g_work.params, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(g_work.X, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(g_work.T, T, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s);
    fwd_bwd(g_work, g_work.X, g_work.T, n, s);
    float l = 0.0f;
    hipMemcpyAsync(&l, g_work.loss, sizeof(float), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(grads, g_work.grads, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) return E4HIP_ERR_HIP;
    if (loss) *loss = l / (float)n;
    return E4HIP_OK;
}

int e4hip_mlp_train(float* params, const float* X, const float* T, int n, int epochs, int batch, float lr,
                    unsigned long long seed, float* loss_per_epoch)
{
    if (!params || !X || !T || n <= 0 || epochs <= 0 || batch <= 0 || batch > n) return E4HIP_ERR_ARG;
    // the whole data set lives on the device; batches are contiguous slices of a
    // host-shuffled copy, reshuffled every epoch
    float *dX = nullptr, *dT = nullptr;
    if (hipMalloc(&dX, (size_t)n * NIN * sizeof(float)) || hipMalloc(&dT, (size_t)n * sizeof(float)))
        return E4HIP_ERR_HIP;
    if (g_work.alloc(batch)) return E4HIP_ERR_HIP;
    hipStream_t s = (hipStream_t)e4hip_stream();
    hipMemcpyAsync(g_work.params, params, NPARAM * sizeof(float), hipMemcpyHostToDevice, s);
    hipMemsetAsync(g_work.m, 0, NPARAM * sizeof(float), s);
    hipMemsetAsync(g_work.v, 0, NPARAM * sizeof(float), s);
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::vector<float> hX((size_t)n * NIN), hT(n);
    unsigned long long st = seed ? seed : 88172645463325252ull;
    auto rnd = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
    const int nb = n / batch;
    long step = 0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = n - 1; i > 0; --i) { int j = (int)(rnd() % (unsigned long long)(i + 1)); std::swap(perm[i], perm[j]); }
        for (int i = 0; i < n; ++i) {
            memcpy(&hX[(size_t)i * NIN], &X[(size_t)perm[i] * NIN], NIN * sizeof(float));
            hT[i] = T[perm[i]];
        }
        hipMemcpyAsync(dX, hX.data(), hX.size() * sizeof(float), hipMemcpyHostToDevice, s);
        hipMemcpyAsync(dT, hT.data(), hT.size() * sizeof(float), hipMemcpyHostToDevice, s);
        double ep_loss = 0.0;
        for (int bi = 0; bi < nb; ++bi) {
            fwd_bwd(g_work, dX + (size_t)bi * batch * NIN, dT + (size_t)bi * batch, batch, s);
            ++step;
            // cosine decay of the learning rate over the run
            float frac = (float)step / (float)(epochs * nb);
            float lr_t = lr * 0.5f * (1.0f + cosf(3.14159265f * frac)) + 1e-5f;
            float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
            hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, g_work.params, g_work.grads,
                               g_work.m, g_work.v, NPARAM, lr_t, b1, b2, eps, c1, c2);
            if (loss_per_epoch && bi % 16 == 0) {
                float l = 0.0f;
                hipMemcpyAsync(&l, g_work.loss, sizeof(float), hipMemcpyDeviceToHost, s);
                hipStreamSynchronize(s);
                ep_loss += l / batch;
            }
        }
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(ep_loss / ((nb + 15) / 16));
    }
    hipMemcpyAsync(params, g_work.params, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    hipFree(dX);
    hipFree(dT);
    return (e == hipSuccess) ? E4HIP_OK : E4HIP_ERR_HIP;
}

} // extern "C"
