// On-device peak probes: the dense matrix-core rate of one MFMA shape per
// operand type, and the bandwidth of a streaming copy through HBM, each
// timed with HIP events on the handle's stream.  They give the achievable
// peaks that roofline fractions can be quoted against beside the
// published ones.
#include <hip/hip_runtime.h>

#include "context.h"

namespace flx {
namespace {

typedef double pf64x4 __attribute__((ext_vector_type(4)));
typedef float pf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 pbf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 pf16x8 __attribute__((ext_vector_type(8)));

constexpr int PROBE_CHAINS = 8;      // independent accumulators per wave
constexpr int PROBE_THREADS = 256;
constexpr int PROBE_GRID = 1024;     // four workgroups per CU

// flops of one instruction of each probed shape (per wave)
constexpr double probe_flops(int kind)
{
    return kind <= 1 ? 2.0 * 16 * 16 * 4 : 2.0 * 16 * 16 * 32;
}

// kind 0: v_mfma_f64_16x16x4f64, 1: v_mfma_f32_16x16x4f32,
// 2: v_mfma_f32_16x16x32_bf16, 3: v_mfma_f32_16x16x32_f16
template <int KIND>
__global__ __launch_bounds__(PROBE_THREADS) void mfma_probe_kernel(int iters, int write,
                                                                   float *sink)
{
    const int lane = threadIdx.x & 63;
    float out = 0.0f;
    if constexpr (KIND == 0) {
        const double a = 1.0 + 1e-9 * lane, b = 1.0 - 1e-9 * lane;
        pf64x4 acc[PROBE_CHAINS];
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) acc[c] = pf64x4{0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int c = 0; c < PROBE_CHAINS; ++c)
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) out += (float)(acc[c][0] + acc[c][3]);
    } else if constexpr (KIND == 1) {
        const float a = 1.0f + 1e-6f * lane, b = 1.0f - 1e-6f * lane;
        pf32x4 acc[PROBE_CHAINS];
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) acc[c] = pf32x4{0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int c = 0; c < PROBE_CHAINS; ++c)
                acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[c], 0, 0, 0);
        }
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) out += acc[c][0] + acc[c][3];
    } else {
        typedef typename std::conditional<KIND == 2, pbf16x8, pf16x8>::type v8;
        typedef typename std::conditional<KIND == 2, __bf16, _Float16>::type e;
        v8 a, b;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a[k] = (e)(1.0f + 0.001f * ((lane + k) & 7));
            b[k] = (e)(1.0f - 0.001f * ((lane + 3 * k) & 7));
        }
        pf32x4 acc[PROBE_CHAINS];
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) acc[c] = pf32x4{0.f, 0.f, 0.f, 0.f};
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int c = 0; c < PROBE_CHAINS; ++c) {
                if constexpr (KIND == 2)
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[c], 0, 0, 0);
                else
                    acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[c], 0, 0, 0);
            }
        }
#pragma unroll
        for (int c = 0; c < PROBE_CHAINS; ++c) out += acc[c][0] + acc[c][3];
    }
    // the caller passes write = 0: the results only have to look used
    if (write) sink[blockIdx.x * PROBE_THREADS + threadIdx.x] = out;
}

__global__ __launch_bounds__(256) void copy_probe_kernel(const pf64x4 *__restrict__ src,
                                                         pf64x4 *__restrict__ dst, size_t n4)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        dst[i] = __builtin_nontemporal_load(&src[i]);
}

// best of reps timed runs of launch() on s (after one untimed run), in ms
template <typename F>
hipError_t best_ms(hipStream_t s, int reps, F launch, float *ms)
{
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) { (void)hipEventDestroy(e0); return e; }
    float best = 1e30f;
    for (int r = 0; r <= reps && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && r > 0 && t < best) best = t;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return e;
}

template <int KIND>
hipError_t launch_mfma_probe(hipStream_t s, int iters)
{
    mfma_probe_kernel<KIND><<<PROBE_GRID, PROBE_THREADS, 0, s>>>(iters, 0, nullptr);
    return hipGetLastError();
}

}  // namespace
}  // namespace flx

using namespace flx;

extern "C" int flx_probe_mfma(flx_handle h, int kind, int iters, double *tflops)
{
    if (!h || kind < 0 || kind > 3 || iters <= 0 || !tflops) return FLX_INVALID_ARG;
    hipStream_t s = h->stream;
    auto launch = [&]() -> hipError_t {
        switch (kind) {
        case 0: return launch_mfma_probe<0>(s, iters);
        case 1: return launch_mfma_probe<1>(s, iters);
        case 2: return launch_mfma_probe<2>(s, iters);
        default: return launch_mfma_probe<3>(s, iters);
        }
    };
    float ms = 0.0f;
    FLX_TRY(best_ms(s, 3, launch, &ms));
    const double waves = (double)PROBE_GRID * (PROBE_THREADS / 64);
    const double flops = waves * iters * PROBE_CHAINS * probe_flops(kind);
    *tflops = flops / (ms * 1e-3) / 1e12;
    return FLX_SUCCESS;
}

extern "C" int flx_probe_copy(flx_handle h, const void *src, void *dst, size_t bytes,
                              double *gbps)
{
    if (!h || !src || !dst || !gbps || bytes < 32 || bytes % 32 != 0) return FLX_INVALID_ARG;
    if ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) % 32 != 0)
        return FLX_INVALID_ARG;
    hipStream_t s = h->stream;
    const size_t n4 = bytes / 32;
    auto launch = [&]() -> hipError_t {
        copy_probe_kernel<<<4096, 256, 0, s>>>(static_cast<const pf64x4 *>(src),
                                               static_cast<pf64x4 *>(dst), n4);
        return hipGetLastError();
    };
    float ms = 0.0f;
    FLX_TRY(best_ms(s, 5, launch, &ms));
    *gbps = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    return FLX_SUCCESS;
}
