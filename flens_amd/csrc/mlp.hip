// mlp.hip -- the precision selector's network: an FP32 MLP
//   x (MLP_IN) -> relu(W1 x + b1) (64) -> relu(W2 h1 + b2) (64) -> W3 h2 + b3 (MLP_OUT)
// with its own forward, backward (softmax cross-entropy) and Adam kernels.
// Parameters live in one flat FP32 array in this order:
//   W1 [64][MLP_IN], b1 [64], W2 [64][64], b2 [64], W3 [MLP_OUT][64], b3 [MLP_OUT]
// The same layout is read by the in-kernel inference of the batched solver
// (batched.hip).  No counterpart in the reference.
#include <cmath>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

constexpr int MH = MLP_HIDDEN;
constexpr int O_W1 = 0;
constexpr int O_B1 = O_W1 + MH * MLP_IN;
constexpr int O_W2 = O_B1 + MH;
constexpr int O_B2 = O_W2 + MH * MH;
constexpr int O_W3 = O_B2 + MH;
constexpr int O_B3 = O_W3 + MLP_OUT * MH;
static_assert(O_B3 + MLP_OUT == MLP_PARAMS, "parameter layout");

// one sample per thread: all weights staged in LDS
constexpr int MS = 64;   // samples per workgroup (= threads)

__global__ __launch_bounds__(MS) void mlp_forward_kernel(const float *__restrict__ P,
                                                         const float *__restrict__ X, int nb,
                                                         float *__restrict__ logits)
{
    __shared__ float w[MLP_PARAMS];
    for (int q = threadIdx.x; q < MLP_PARAMS; q += MS) w[q] = P[q];
    __syncthreads();
    const int s = blockIdx.x * MS + threadIdx.x;
    if (s >= nb) return;
    float x[MLP_IN], h1[MH], h2[MH];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) x[i] = X[(long)s * MLP_IN + i];
    for (int j = 0; j < MH; ++j) {
        float a = w[O_B1 + j];
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) a = fmaf(w[O_W1 + j * MLP_IN + i], x[i], a);
        h1[j] = fmaxf(a, 0.0f);
    }
    for (int j = 0; j < MH; ++j) {
        float a = w[O_B2 + j];
        for (int i = 0; i < MH; ++i) a = fmaf(w[O_W2 + j * MH + i], h1[i], a);
        h2[j] = fmaxf(a, 0.0f);
    }
    for (int o = 0; o < MLP_OUT; ++o) {
        float a = w[O_B3 + o];
        for (int i = 0; i < MH; ++i) a = fmaf(w[O_W3 + o * MH + i], h2[i], a);
        logits[(long)s * MLP_OUT + o] = a;
    }
}

// Gradient of the mean softmax cross-entropy over nb samples (labels in
// [0, MLP_OUT); weight[s] scales sample s, may be null).  Each workgroup
// writes the summed gradient of its samples to part[blk][MLP_PARAMS] and
// its summed loss to part_loss[blk]; mlp_reduce_kernel adds the blocks in
// a fixed order (deterministic).
__global__ __launch_bounds__(MS) void mlp_grad_kernel(const float *__restrict__ P,
                                                      const float *__restrict__ X,
                                                      const int *__restrict__ label,
                                                      const float *__restrict__ weight, int nb,
                                                      float *__restrict__ part,
                                                      float *__restrict__ part_loss)
{
    __shared__ float w[MLP_PARAMS];
    __shared__ float g[MLP_PARAMS];
    __shared__ float sl[MS];
    for (int q = threadIdx.x; q < MLP_PARAMS; q += MS) { w[q] = P[q]; g[q] = 0.0f; }
    __syncthreads();
    const int s = blockIdx.x * MS + threadIdx.x;
    const bool ok = s < nb;
    float x[MLP_IN], h1[MH], h2[MH], z[MLP_OUT], d3[MLP_OUT], d2[MH];
    float loss = 0.0f;
    if (ok) {
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) x[i] = X[(long)s * MLP_IN + i];
        for (int j = 0; j < MH; ++j) {
            float a = w[O_B1 + j];
#pragma unroll
            for (int i = 0; i < MLP_IN; ++i) a = fmaf(w[O_W1 + j * MLP_IN + i], x[i], a);
            h1[j] = fmaxf(a, 0.0f);
        }
        for (int j = 0; j < MH; ++j) {
            float a = w[O_B2 + j];
            for (int i = 0; i < MH; ++i) a = fmaf(w[O_W2 + j * MH + i], h1[i], a);
            h2[j] = fmaxf(a, 0.0f);
        }
        float zmax = -INFINITY;
        for (int o = 0; o < MLP_OUT; ++o) {
            float a = w[O_B3 + o];
            for (int i = 0; i < MH; ++i) a = fmaf(w[O_W3 + o * MH + i], h2[i], a);
            z[o] = a;
            zmax = fmaxf(zmax, a);
        }
        float se = 0.0f;
        for (int o = 0; o < MLP_OUT; ++o) se += expf(z[o] - zmax);
        const int y = label[s];
        const float ws = weight ? weight[s] : 1.0f;
        loss = ws * (logf(se) + zmax - z[y]);
        for (int o = 0; o < MLP_OUT; ++o)
            d3[o] = ws * (expf(z[o] - zmax) / se - (o == y ? 1.0f : 0.0f));
        for (int i = 0; i < MH; ++i) {
            float a = 0.0f;
            for (int o = 0; o < MLP_OUT; ++o) a = fmaf(w[O_W3 + o * MH + i], d3[o], a);
            d2[i] = h2[i] > 0.0f ? a : 0.0f;
        }
    }
    // accumulate the block's parameter gradients: one sample at a time per
    // parameter row keeps the summation order fixed (threads own rows)
    sl[threadIdx.x] = loss;
    __shared__ float sx[MS][MLP_IN], sh1[MS][MH + 1], sh2[MS][MH + 1], sd3[MS][MLP_OUT], sd2[MS][MH + 1];
    __shared__ float sd1[MS][MH + 1];
    if (ok) {
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) sx[threadIdx.x][i] = x[i];
        for (int j = 0; j < MH; ++j) { sh1[threadIdx.x][j] = h1[j]; sh2[threadIdx.x][j] = h2[j]; sd2[threadIdx.x][j] = d2[j]; }
        for (int o = 0; o < MLP_OUT; ++o) sd3[threadIdx.x][o] = d3[o];
        for (int j = 0; j < MH; ++j) {
            float a = 0.0f;
            for (int i = 0; i < MH; ++i) a = fmaf(w[O_W2 + i * MH + j], d2[i], a);
            sd1[threadIdx.x][j] = h1[j] > 0.0f ? a : 0.0f;
        }
    } else {
        for (int i = 0; i < MLP_IN; ++i) sx[threadIdx.x][i] = 0.0f;
        for (int j = 0; j < MH; ++j) { sh1[threadIdx.x][j] = 0.0f; sh2[threadIdx.x][j] = 0.0f; sd2[threadIdx.x][j] = 0.0f; sd1[threadIdx.x][j] = 0.0f; }
        for (int o = 0; o < MLP_OUT; ++o) sd3[threadIdx.x][o] = 0.0f;
    }
    __syncthreads();
    const int ns = min(MS, nb - blockIdx.x * MS);
    // thread j owns hidden unit j: W1 row j, b1[j], W2 row j, b2[j]; threads
    // < MLP_OUT also own W3 row o and b3[o]
    {
        const int j = threadIdx.x;
        float gw1[MLP_IN] = {}, gb1 = 0.0f, gb2 = 0.0f;
        for (int q = 0; q < ns; ++q) {
            const float d1 = sd1[q][j];
#pragma unroll
            for (int i = 0; i < MLP_IN; ++i) gw1[i] = fmaf(d1, sx[q][i], gw1[i]);
            gb1 += d1;
            gb2 += sd2[q][j];
        }
#pragma unroll
        for (int i = 0; i < MLP_IN; ++i) g[O_W1 + j * MLP_IN + i] = gw1[i];
        g[O_B1 + j] = gb1;
        g[O_B2 + j] = gb2;
        for (int i = 0; i < MH; ++i) {
            float a = 0.0f;
            for (int q = 0; q < ns; ++q) a = fmaf(sd2[q][j], sh1[q][i], a);
            g[O_W2 + j * MH + i] = a;
        }
        if (j < MLP_OUT) {
            float gb3 = 0.0f;
            for (int q = 0; q < ns; ++q) gb3 += sd3[q][j];
            g[O_B3 + j] = gb3;
            for (int i = 0; i < MH; ++i) {
                float a = 0.0f;
                for (int q = 0; q < ns; ++q) a = fmaf(sd3[q][j], sh2[q][i], a);
                g[O_W3 + j * MH + i] = a;
            }
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < MLP_PARAMS; q += MS) part[(long)blockIdx.x * MLP_PARAMS + q] = g[q];
    if (threadIdx.x == 0) {
        float a = 0.0f;
        for (int q = 0; q < ns; ++q) a += sl[q];
        part_loss[blockIdx.x] = a;
    }
}

__global__ void mlp_reduce_kernel(const float *__restrict__ part, const float *__restrict__ part_loss,
                                  int nblk, float inv_nb, float *__restrict__ grad,
                                  float *__restrict__ loss)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < MLP_PARAMS) {
        float a = 0.0f;
        for (int b = 0; b < nblk; ++b) a += part[(long)b * MLP_PARAMS + q];
        grad[q] = a * inv_nb;
    }
    if (q == 0 && loss) {
        float a = 0.0f;
        for (int b = 0; b < nblk; ++b) a += part_loss[b];
        *loss = a * inv_nb;
    }
}

__global__ void adam_kernel(float *__restrict__ P, const float *__restrict__ grad,
                            float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2, float wd)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= MLP_PARAMS) return;
    const float gq = grad[q] + wd * P[q];
    const float mq = b1 * m[q] + (1.0f - b1) * gq;
    const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
    m[q] = mq;
    v[q] = vq;
    P[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

size_t mlp_part_floats(int nb) { return (size_t)((nb + MS - 1) / MS) * (MLP_PARAMS + 1); }

hipError_t mlp_forward(hipStream_t s, const float *P, const float *X, int nb, float *logits)
{
    if (nb <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((nb + MS - 1) / MS), dim3(MS), 0, s, P, X, nb,
                       logits);
    return hipGetLastError();
}

hipError_t mlp_grad(hipStream_t s, const float *P, const float *X, const int *label,
                    const float *weight, int nb, float *part, float *grad, float *loss)
{
    if (nb <= 0) return hipSuccess;
    const int nblk = (nb + MS - 1) / MS;
    float *part_loss = part + (size_t)nblk * MLP_PARAMS;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk), dim3(MS), 0, s, P, X, label, weight, nb, part,
                       part_loss);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s,
                       (const float *)part, (const float *)part_loss, nblk, 1.0f / nb, grad, loss);
    return hipGetLastError();
}

hipError_t mlp_adam(hipStream_t s, float *P, const float *grad, float *m, float *v, int step,
                    float lr, float wd)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((MLP_PARAMS + 255) / 256), dim3(256), 0, s, P, grad, m, v,
                       lr, b1, b2, 1e-8f, c1, c2, wd);
    return hipGetLastError();
}

} // namespace flx
