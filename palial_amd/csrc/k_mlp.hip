// k_mlp.hip -- the task-time cost model: a 3-layer perceptron, 64 hidden
// units, FP32 (no reference counterpart).
//
//   h1 = relu(x W1 + b1),  h2 = relu(h1 W2 + b2),  y = h2 . w3 + b3
//
// Parameters, one float array: W1 [F x 64] (row f holds the 64 weights of
// input f), b1 [64], W2 [64 x 64] (row i: weights from h1_i), b2 [64],
// w3 [64], b3 [1].  Loss: mean squared error over the samples.
//
// mlp_forward_kernel : one thread per sample, parameters staged in LDS.
// mlp_train_kernel   : one workgroup; samples in chunks of MB = 128 (in
//                      order), per-sample activations in LDS, parameter
//                      gradients summed over the chunk in a fixed order (so
//                      a gradient is bitwise reproducible), Adam step per
//                      chunk.  epochs = 0 returns the full-batch gradient.
#include "internal.hpp"

namespace palial {

namespace {

constexpr int HID = 64;
constexpr int MAXF = 16;
constexpr int MB = 128;          // samples per chunk
constexpr int TT = 256;          // threads of the training workgroup

__host__ __device__ constexpr int nparams(int F) { return F * HID + HID + HID * HID + HID + HID + 1; }

__global__ __launch_bounds__(256) void mlp_forward_kernel(int n, int F, const float *X, const float *P, float *Y)
{
    __shared__ float sp[nparams(MAXF)];
    const int np = nparams(F);
    for (int e = threadIdx.x; e < np; e += blockDim.x) sp[e] = P[e];
    __syncthreads();
    const float *W1 = sp, *b1 = W1 + F * HID, *W2 = b1 + HID, *b2 = W2 + HID * HID, *w3 = b2 + HID, *b3 = w3 + HID;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float x[MAXF];
#pragma unroll
    for (int f = 0; f < MAXF; f++) x[f] = (f < F) ? X[(size_t)s * F + f] : 0.f;
    float h1[HID];
#pragma unroll
    for (int h = 0; h < HID; h++) {
        float a = b1[h];
#pragma unroll
        for (int f = 0; f < MAXF; f++)
            if (f < F) a += x[f] * W1[f * HID + h];
        h1[h] = fmaxf(a, 0.f);
    }
    float y = b3[0];
    for (int j = 0; j < HID; j++) {
        float a = b2[j];
#pragma unroll
        for (int i = 0; i < HID; i++) a += h1[i] * W2[i * HID + j];
        y += fmaxf(a, 0.f) * w3[j];
    }
    Y[s] = y;
}

// LDS: parameters, their gradient, one chunk of inputs, h1, h2 and the
// back-propagated dh2 (h2's slot is reused for dh1).
struct TrainLDS {
    float p[nparams(MAXF)];
    float gsum[nparams(MAXF)];
    float x[MB * MAXF];
    float h1[MB * HID];
    float h2[MB * HID];
    float dh2[MB * HID];
    float dout[MB];
    float red[TT];
};

__global__ __launch_bounds__(TT) void mlp_train_kernel(int n, int F, const float *X, const float *Y, float *P,
                                                       float *adam_m, float *adam_v, int epochs, float lr,
                                                       float *loss_out, float *grad_out)
{
    extern __shared__ float smem_f[];
    TrainLDS &L = *reinterpret_cast<TrainLDS *>(smem_f);
    const int t = threadIdx.x;
    const int np = nparams(F);
    for (int e = t; e < np; e += TT) L.p[e] = P[e];
    __syncthreads();
    float *W1 = L.p, *b1 = W1 + F * HID, *W2 = b1 + HID, *b2 = W2 + HID * HID, *w3 = b2 + HID, *b3 = w3 + HID;
    float *gW1 = L.gsum, *gb1 = gW1 + F * HID, *gW2 = gb1 + HID, *gb2 = gW2 + HID * HID, *gw3 = gb2 + HID,
          *gb3 = gw3 + HID;
    const bool grad_only = (epochs == 0);
    const int nep = grad_only ? 1 : epochs;
    const float b1m = 0.9f, b2m = 0.999f, eps = 1e-8f;
    int step = 0;
    for (int ep = 0; ep < nep; ep++) {
        float loss_acc = 0.f;   // thread 0 keeps the running sum (fixed order)
        if (grad_only)
            for (int e = t; e < np; e += TT) L.gsum[e] = 0.f;
        for (int c0 = 0; c0 < n; c0 += MB) {
            const int m = min(MB, n - c0);
            const float scale = grad_only ? 1.f / (float)n : 1.f / (float)m;   // mean over the batch
            if (!grad_only)
                for (int e = t; e < np; e += TT) L.gsum[e] = 0.f;
            for (int e = t; e < m * F; e += TT) L.x[(e / F) * MAXF + e % F] = X[(size_t)(c0 + e / F) * F + e % F];
            __syncthreads();
            // forward: thread t < m handles sample t of the chunk
            if (t < m) {
                const float *x = L.x + t * MAXF;
                for (int h = 0; h < HID; h++) {
                    float a = b1[h];
                    for (int f = 0; f < F; f++) a += x[f] * W1[f * HID + h];
                    L.h1[t * HID + h] = fmaxf(a, 0.f);
                }
                float y = b3[0];
                for (int j = 0; j < HID; j++) {
                    float a = b2[j];
                    for (int i = 0; i < HID; i++) a += L.h1[t * HID + i] * W2[i * HID + j];
                    const float h2 = fmaxf(a, 0.f);
                    L.h2[t * HID + j] = h2;
                    y += h2 * w3[j];
                }
                const float r = y - Y[c0 + t];
                L.dout[t] = 2.f * r * scale;
                L.red[t] = r * r;
                for (int j = 0; j < HID; j++) {
                    const float h2 = L.h2[t * HID + j];
                    L.dh2[t * HID + j] = (h2 > 0.f) ? L.dout[t] * w3[j] : 0.f;
                }
            }
            __syncthreads();
            if (t == 0)
                for (int s = 0; s < m; s++) loss_acc += L.red[s];
            // gradients of layers 3 and 2 (sums over the chunk in sample order)
            for (int e = t; e < HID * HID; e += TT) {
                const int i = e / HID, j = e % HID;
                float a = 0.f;
                for (int s = 0; s < m; s++) a += L.h1[s * HID + i] * L.dh2[s * HID + j];
                gW2[e] += a;
            }
            if (t < HID) {
                float a = 0.f, c = 0.f;
                for (int s = 0; s < m; s++) {
                    a += L.dout[s] * L.h2[s * HID + t];
                    c += L.dh2[s * HID + t];
                }
                gw3[t] += a;
                gb2[t] += c;
            } else if (t == HID) {
                float a = 0.f;
                for (int s = 0; s < m; s++) a += L.dout[s];
                gb3[0] += a;
            }
            __syncthreads();
            // dh1 = (dh2 W2^T) * (h1 > 0), into the h2 slot
            if (t < m) {
                for (int i = 0; i < HID; i++) {
                    float a = 0.f;
                    for (int j = 0; j < HID; j++) a += L.dh2[t * HID + j] * W2[i * HID + j];
                    L.h2[t * HID + i] = (L.h1[t * HID + i] > 0.f) ? a : 0.f;
                }
            }
            __syncthreads();
            for (int e = t; e < F * HID; e += TT) {
                const int f = e / HID, h = e % HID;
                float a = 0.f;
                for (int s = 0; s < m; s++) a += L.x[s * MAXF + f] * L.h2[s * HID + h];
                gW1[e] += a;
            }
            if (t < HID) {
                float a = 0.f;
                for (int s = 0; s < m; s++) a += L.h2[s * HID + t];
                gb1[t] += a;
            }
            __syncthreads();
            if (!grad_only) {
                // Adam step on this chunk's mean gradient
                step++;
                const float c1 = 1.f - powf(b1m, (float)step), c2 = 1.f - powf(b2m, (float)step);
                for (int e = t; e < np; e += TT) {
                    const float gv = L.gsum[e];
                    const float mv = b1m * adam_m[e] + (1.f - b1m) * gv;
                    const float vv = b2m * adam_v[e] + (1.f - b2m) * gv * gv;
                    adam_m[e] = mv;
                    adam_v[e] = vv;
                    L.p[e] -= lr * (mv / c1) / (sqrtf(vv / c2) + eps);
                }
                __syncthreads();
            }
        }
        if (t == 0 && loss_out) loss_out[ep] = loss_acc / (float)n;
    }
    for (int e = t; e < np; e += TT) {
        if (grad_only) {
            if (grad_out) grad_out[e] = L.gsum[e];
        } else {
            P[e] = L.p[e];
        }
    }
}

}  // namespace

int mlp_nparams(int F) { return nparams(F); }

hipError_t launch_mlp_forward(int n, int F, const float *X, const float *P, float *Y, hipStream_t s)
{
    if (F < 1 || F > MAXF || n < 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, F, X, P, Y);
    return hipGetLastError();
}

void mlp_init()
{
    (void)hipFuncSetAttribute((const void *)mlp_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)sizeof(TrainLDS));
}

hipError_t launch_mlp_train(int n, int F, const float *X, const float *Y, float *P, float *m, float *v, int epochs,
                            float lr, float *loss, float *grad, hipStream_t s)
{
    if (F < 1 || F > MAXF || n < 1 || epochs < 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(TT), sizeof(TrainLDS), s, n, F, X, Y, P, m, v, epochs, lr,
                       loss, grad);
    return hipGetLastError();
}

}  // namespace palial
