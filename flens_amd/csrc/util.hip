// util.hip -- device-side input generation.
//
// Counter-based generator: element idx of stream `seed` is
// splitmix64(seed + (idx+1)*golden), mapped exactly onto the dyadic grid
// k*2^-52 - 1 in [-1,1).  The same integer recipe runs on the host in the
// test suite, so device-generated inputs are reproducible bit for bit.
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

FLX_DEV uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void fill_uniform_kernel(double *__restrict__ A, int m, int n, long lda,
                                    uint64_t seed, uint64_t offset)
{
    const long total = (long)m * n;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long j = t / m, i = t % m;
        const uint64_t idx = offset + (uint64_t)t;
        const uint64_t z = mix64(seed + (idx + 1) * 0x9E3779B97F4A7C15ull);
        const long long k = (long long)(z >> 11) - (1ll << 52);
        A[j * lda + i] = (double)k * 0x1p-52;
    }
}

hipError_t fill_uniform_f64(hipStream_t s, double *A, int m, int n, long lda, uint64_t seed,
                            uint64_t offset)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    long total = (long)m * n;
    int grid = (int)((total + 255) / 256);
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid), dim3(256), 0, s, A, m, n, lda, seed,
                       offset);
    return hipGetLastError();
}

} // namespace flx
