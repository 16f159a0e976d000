// getri.hip -- inverse of a matrix from its LU factor on the device.
//
// Replaces flens::lapack::tri for GeMatrix (flens/lapack/ge/tri.tcc:64-187:
// inv(U) by trtri, flens/lapack/tr/tri.tcc, then inv(A) L = inv(U) and the
// column interchanges) and the Fortran seam dgetri_
// (cxxlapack/netlib/interface/lapack.in.h:6758-6766).
//
// MI355X form: inv(A) = inv(U) inv(L) P^T with both triangular inverses
// built by recursive doubling -- 64 x 64 diagonal blocks inverted directly
// (one workgroup each), then at sizes s = 64, 128, ... the off-diagonal
// block of each 2s x 2s diagonal block from two GEMMs,
//   lower:  X21 = -inv(T22) (T21 inv(T11)),   upper:  X12 = -(inv(T11) T12) inv(T22),
// and one final GEMM inv(U) inv(L) followed by a column gather.  Every
// O(n^3) step is the MFMA DGEMM.  The singularity test comes first (first
// exact zero on U's diagonal, as trtri), leaving A untouched when it fails.
#include <cfloat>
#include <climits>
#include <vector>
#include "flx_common.h"
#include "context.h"

namespace flx {

constexpr int TI_BS = 64;

// Inverse of diagonal block b of the unit-lower (UPPER = 0) or non-unit
// upper triangle of LU into out (its TI_BS x TI_BS block, the other
// triangle of the block zeroed).  Thread c builds column c.
template <bool UPPER>
__global__ __launch_bounds__(64) void tri_block_inv_kernel(const double *__restrict__ LU, long lda,
                                                           int n, double *__restrict__ out,
                                                           long ldo)
{
    __shared__ double T[TI_BS][TI_BS + 1];
    __shared__ double X[TI_BS][TI_BS + 1];   // X[row][col]
    const int c = threadIdx.x;
    const int k0 = blockIdx.x * TI_BS;
    const int bs = min(TI_BS, n - k0);
    for (int i = 0; i < bs; ++i) T[i][c] = (c < bs) ? LU[(long)(k0 + c) * lda + k0 + i] : 0.0;
    __syncthreads();
    if (c < bs) {
        if (UPPER) {
            for (int i = bs - 1; i >= 0; --i) {
                double v = (i == c) ? 1.0 : 0.0;
                if (i < c) {
                    for (int k = i + 1; k <= c; ++k) v -= T[i][k] * X[k][c];
                } else if (i > c) {
                    v = 0.0;
                }
                X[i][c] = (i <= c) ? v / T[i][i] : 0.0;
            }
        } else {
            for (int i = 0; i < bs; ++i) {
                double v = (i == c) ? 1.0 : 0.0;
                if (i > c)
                    for (int k = c; k < i; ++k) v -= T[i][k] * X[k][c];
                X[i][c] = (i >= c) ? v : 0.0;
            }
        }
    }
    __syncthreads();
    for (int i = 0; i < bs; ++i)
        if (c < bs) out[(long)(k0 + c) * ldo + k0 + i] = X[i][c];
}

__global__ void diag_first_zero_kernel(const double *__restrict__ LU, long lda, int n, int *first)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && LU[(long)i * lda + i] == 0.0) atomicMin(first, i + 1);
}

__global__ void fill_zero_kernel(double *__restrict__ p, long count)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (long)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

// out(:, p) = W(:, cols[p])
__global__ void gather_cols_kernel(const double *__restrict__ W, long ldw, int n,
                                   const int *__restrict__ cols, double *__restrict__ out,
                                   long ldo)
{
    for (int p = blockIdx.x; p < n; p += gridDim.x) {
        const double *src = W + (long)cols[p] * ldw;
        double *dst = out + (long)p * ldo;
        for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
    }
}

// inverse of the unit-lower (upper = 0) or upper triangle of LU into Tinv
// (n x n, ld n, the other triangle zero); tmp holds n x n scratch
static hipError_t tri_inverse(hipStream_t s, int upper, int n, const double *LU, long lda,
                              double *Tinv, double *tmp)
{
    const long ld = n;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(1024), dim3(256), 0, s, Tinv, (long)n * n);
    const int nblk = (n + TI_BS - 1) / TI_BS;
    if (upper)
        hipLaunchKernelGGL(tri_block_inv_kernel<true>, dim3(nblk), dim3(64), 0, s, LU, lda, n,
                           Tinv, ld);
    else
        hipLaunchKernelGGL(tri_block_inv_kernel<false>, dim3(nblk), dim3(64), 0, s, LU, lda, n,
                           Tinv, ld);
    FLX_CHECK(hipGetLastError());
    for (int sz = TI_BS; sz < n; sz *= 2) {
        for (int k0 = 0; k0 + sz < n; k0 += 2 * sz) {
            const int s1 = sz, s2 = (n - (k0 + sz) < sz) ? n - (k0 + sz) : sz;
            const int k1 = k0 + s1;
            if (!upper) {
                // tmp (s2 x s1) = T21 inv(T11);  X21 = -inv(T22) tmp
                FLX_CHECK(gemm_f64(s, 0, 0, s2, s1, s1, 1.0, LU + (long)k0 * lda + k1, lda,
                                   Tinv + (long)k0 * ld + k0, ld, 0.0, tmp, s2, 1));
                FLX_CHECK(gemm_f64(s, 0, 0, s2, s1, s2, -1.0, Tinv + (long)k1 * ld + k1, ld, tmp,
                                   s2, 0.0, Tinv + (long)k0 * ld + k1, ld, 1));
            } else {
                // tmp (s1 x s2) = inv(T11) T12;  X12 = -tmp inv(T22)
                FLX_CHECK(gemm_f64(s, 0, 0, s1, s2, s1, 1.0, Tinv + (long)k0 * ld + k0, ld,
                                   LU + (long)k1 * lda + k0, lda, 0.0, tmp, s1, 1));
                FLX_CHECK(gemm_f64(s, 0, 0, s1, s2, s2, -1.0, tmp, s1, Tinv + (long)k1 * ld + k1,
                                   ld, 0.0, Tinv + (long)k1 * ld + k0, ld, 1));
            }
        }
    }
    return hipSuccess;
}

} // namespace flx

using namespace flx;

extern "C" int flx_dgetri(flx_handle h, int n, double *A, int lda, const int *ipiv, int *info)
{
    ProfBind pb_(h);
    if (!h || !info || n < 0 || lda < (n > 1 ? n : 1)) return FLX_INVALID_ARG;
    *info = 0;
    if (n == 0) return FLX_SUCCESS;
    hipStream_t s = h->stream;
    // singularity first, as trtri: A is left untouched when U(i,i) = 0
    FLX_TRY(ensure(h, h->aux.slot, sizeof(unsigned long long) * 8));
    int *first = (int *)h->aux.slot.p;
    FLX_TRY(fill_int(s, first, 1, INT_MAX));
    hipLaunchKernelGGL(diag_first_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const double *)A, (long)lda, n, first);
    FLX_TRY(hipGetLastError());
    int fz = INT_MAX;
    FLX_TRY(hipMemcpyAsync(&fz, first, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<int> piv(n);
    FLX_TRY(hipMemcpyAsync(piv.data(), ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    FLX_TRY(hipStreamSynchronize(s));
    if (fz != INT_MAX) {
        *info = fz;
        return FLX_SUCCESS;
    }
    // column order of inv(A): the interchanges j = n-1 .. 1 of tri.tcc:176-182
    std::vector<int> cols(n);
    for (int j = 0; j < n; ++j) cols[j] = j;
    for (int j = n - 1; j >= 1; --j) {
        const int jp = piv[j - 1];
        if (jp < 1 || jp > n) return FLX_INVALID_ARG;
        if (jp != j) std::swap(cols[j - 1], cols[jp - 1]);
    }
    const size_t nn = (size_t)n * n;
    FLX_TRY(ensure(h, h->W, sizeof(double) * nn));
    FLX_TRY(ensure(h, h->C, sizeof(double) * nn));
    FLX_TRY(ensure(h, h->X, sizeof(double) * nn));
    FLX_TRY(ensure(h, h->perm, sizeof(int) * (size_t)n));
    double *Linv = (double *)h->W.p, *Uinv = (double *)h->C.p, *T = (double *)h->X.p;
    FLX_TRY(tri_inverse(s, 0, n, A, lda, Linv, T));
    FLX_TRY(tri_inverse(s, 1, n, A, lda, Uinv, T));
    // T = inv(U) inv(L), then inv(A)(:, p) = T(:, cols[p])
    FLX_TRY(gemm_f64(s, 0, 0, n, n, n, 1.0, Uinv, n, Linv, n, 0.0, T, n, 1));
    FLX_TRY(hipMemcpyAsync(h->perm.p, cols.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gather_cols_kernel, dim3(n < 1024 ? n : 1024), dim3(256), 0, s,
                       (const double *)T, (long)n, n, (const int *)h->perm.p, A, (long)lda);
    FLX_TRY(hipGetLastError());
    FLX_TRY(hipStreamSynchronize(s));
    return FLX_SUCCESS;
}
