// gemm.hip -- column-major GEMM on CDNA4 matrix cores.
//
//   C <- beta*C + alpha*op(A)*op(B)
//
// Replaces the cxxblas level-3 GEMM (generic path cxxblas/level3/gemm.tcc:
// 149-299, CBLAS seam gemm.tcc:327-349, playground intrinsics GEMM
// playground/cxxblas/intrinsics/level3/gemm.tcc:60-327).  Semantics kept
// from the reference: beta scales C before the product is added, so beta=0
// still computes 0*C (NaN/Inf in C propagate, gemm.tcc:175), and alpha=0
// returns after the beta scaling (gemm.tcc:176-178).
//
// Design: 256 threads = 4 wave64s in a 2x2 grid; each wave owns a 64x64
// sub-tile built from 4x4 blocks of the 16x16x4 MFMA (f64: v_mfma_f64_
// 16x16x4_f64, f32: v_mfma_f32_16x16x4_f32 -- same operand layout).  A is
// staged in LDS as [k][m] (m contiguous, the MFMA A-operand lane order), B
// as [n][k] padded by one element so the 16 lanes that read different n hit
// distinct banks.  Global->register prefetch of tile t+1 overlaps the MFMAs
// of tile t; LDS is double-buffered, one barrier per K tile.
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

// Accumulator type and output layout of the 16x16x4 MFMA.  Item q of lane
// l holds D[row][l%16] with row = l/16 + 4q for f64 and 4*(l/16) + q for
// f32 (measured on gfx950; the two instructions differ).
template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<double> {
    typedef double4_t type;
    static constexpr int row(int lk, int q) { return lk + 4 * q; }
};
template <> struct MfmaAcc<float> {
    typedef float4_t type;
    static constexpr int row(int lk, int q) { return 4 * lk + q; }
};

FLX_DEV double4_t mfma16x16x4(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
FLX_DEV float4_t mfma16x16x4(float a, float b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int GBM = 128, GBN = 128, GBK = 16;
constexpr int APAD = 16;   // [k][m] rows shifted by 32 banks
constexpr int BPAD = 1;    // [n][k] rows of 17 elements

// Loads one thread's 8-element slice of an A or B tile into registers.
// NT selects which index is contiguous in memory for this operand.
//  - "row run": 8 consecutive elements along the contiguous dimension.
template <typename T>
FLX_DEV void load_run8(T (&r)[8], const T *__restrict__ base, long stride_c,
                       int i0, int j, int imax, int jmax, bool full, bool vec)
{
    // elements base[(i0+q) + j*stride_c], q = 0..7, valid if i0+q<imax, j<jmax
    const T *p = base + (long)j * stride_c + i0;
    if (full && vec) {
        if constexpr (sizeof(T) == 8) {
            const double2 *q = reinterpret_cast<const double2 *>(p);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double2 v = q[u];
                r[2 * u] = v.x;
                r[2 * u + 1] = v.y;
            }
        } else {
            const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                float4 v = q[u];
                r[4 * u] = v.x; r[4 * u + 1] = v.y;
                r[4 * u + 2] = v.z; r[4 * u + 3] = v.w;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q)
            r[q] = (j < jmax && i0 + q < imax) ? p[q] : T(0);
    }
}

template <typename T, int TA, int TB>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(
    int M, int N, int K, T alpha, const T *__restrict__ A, long lda,
    const T *__restrict__ B, long ldb, T beta, T *__restrict__ C, long ldc,
    int vecA, int vecB, int blas_beta0)
{
    typedef typename MfmaAcc<T>::type acc_t;
    __shared__ T As[2][GBK][GBM + APAD];
    __shared__ T Bs[2][GBN][GBK + BPAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int m0 = blockIdx.x * GBM;
    const int n0 = blockIdx.y * GBN;

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const bool mfull = (m0 + GBM <= M);
    const bool nfull = (n0 + GBN <= N);

    // per-thread slice coordinates
    // A NoTrans: run along m (8 elems), one k column per thread
    //   thread -> k = tid/16, m = (tid%16)*8
    // A Trans  : run along k (8 elems), one m per thread
    //   thread -> m = tid/2, k = (tid%2)*8
    // B NoTrans: run along k, one n per thread: n = tid/2, k = (tid%2)*8
    // B Trans  : run along n, one k per thread: k = tid/16, n = (tid%16)*8
    T ra[8], rb[8];

    auto load_tiles = [&](int k0) {
        const bool kfull = (k0 + GBK <= K);
        if (!TA) {
            int kk = tid >> 4, mm = (tid & 15) * 8;
            load_run8(ra, A + m0, lda, mm, k0 + kk, M - m0, K, mfull && kfull, vecA);
        } else {
            int mm = tid >> 1, kk = (tid & 1) * 8;
            load_run8(ra, A + k0, lda, kk, m0 + mm, K - k0, M, mfull && kfull, vecA);
        }
        if (!TB) {
            int nn = tid >> 1, kk = (tid & 1) * 8;
            load_run8(rb, B + k0, ldb, kk, n0 + nn, K - k0, N, nfull && kfull, vecB);
        } else {
            int kk = tid >> 4, nn = (tid & 15) * 8;
            load_run8(rb, B + n0, ldb, nn, k0 + kk, N - n0, K, nfull && kfull, vecB);
        }
    };
    auto store_tiles = [&](int buf) {
        if (!TA) {
            int kk = tid >> 4, mm = (tid & 15) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) As[buf][kk][mm + q] = ra[q];
        } else {
            int mm = tid >> 1, kk = (tid & 1) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) As[buf][kk + q][mm] = ra[q];
        }
        if (!TB) {
            int nn = tid >> 1, kk = (tid & 1) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[buf][nn][kk + q] = rb[q];
        } else {
            int kk = tid >> 4, nn = (tid & 15) * 8;
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[buf][nn + q][kk] = rb[q];
        }
    };

    const int nkt = (alpha == T(0)) ? 0 : (K + GBK - 1) / GBK;
    if (nkt > 0) {
        load_tiles(0);
        store_tiles(0);
        __syncthreads();
    }
    const int lrow = lane & 15, lk = lane >> 4;
    for (int kt = 0; kt < nkt; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nkt) load_tiles((kt + 1) * GBK);
#pragma unroll
        for (int s = 0; s < GBK / 4; ++s) {
            T af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                af[i] = As[buf][s * 4 + lk][wm * 64 + i * 16 + lrow];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                bf[j] = Bs[buf][wn * 64 + j * 16 + lrow][s * 4 + lk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = mfma16x16x4(af[i], bf[j], acc[i][j]);
        }
        if (kt + 1 < nkt) {
            store_tiles(buf ^ 1);
            __syncthreads();
        }
    }

    // Epilogue: see MfmaAcc for which row each accumulator item holds.
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + lrow;
            if (col >= N) continue;
            T *cp = C + (long)col * ldc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm * 64 + i * 16 + MfmaAcc<T>::row(lk, q);
                if (row < M) {
                    // generic FLENS order: C*beta first (0*C for beta=0);
                    // BLAS convention (blas_beta0): beta=0 never reads C.
                    T cv = (blas_beta0 && beta == T(0)) ? T(0) : cp[row] * beta;
                    cp[row] = (nkt > 0) ? cv + alpha * acc[i][j][q] : cv;
                }
            }
        }
}

template <typename T>
static hipError_t launch_gemm(hipStream_t s, int ta, int tb, int M, int N,
                              int K, T alpha, const T *A, long lda, const T *B,
                              long ldb, T beta, T *C, long ldc, int blas_beta0)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    dim3 grid((M + GBM - 1) / GBM, (N + GBN - 1) / GBN);
    const uintptr_t al = 16;
    int vecA = ((uintptr_t)A % al == 0) && (lda * (long)sizeof(T)) % al == 0;
    int vecB = ((uintptr_t)B % al == 0) && (ldb * (long)sizeof(T)) % al == 0;
#define FLX_LG(a_, b_)                                                        \
    hipLaunchKernelGGL((gemm_mfma_kernel<T, a_, b_>), grid, dim3(256), 0, s,  \
                       M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vecA, vecB, \
                       blas_beta0)
    if (!ta && !tb) FLX_LG(0, 0);
    else if (!ta && tb) FLX_LG(0, 1);
    else if (ta && !tb) FLX_LG(1, 0);
    else FLX_LG(1, 1);
#undef FLX_LG
    return hipGetLastError();
}

hipError_t gemm_f64(hipStream_t s, int ta, int tb, int M, int N, int K,
                    double alpha, const double *A, long lda, const double *B,
                    long ldb, double beta, double *C, long ldc, int blas_beta0)
{
    return launch_gemm<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta,
                               C, ldc, blas_beta0);
}

hipError_t gemm_f32(hipStream_t s, int ta, int tb, int M, int N, int K,
                    float alpha, const float *A, long lda, const float *B,
                    long ldb, float beta, float *C, long ldc, int blas_beta0)
{
    return launch_gemm<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta,
                              C, ldc, blas_beta0);
}

} // namespace flx
