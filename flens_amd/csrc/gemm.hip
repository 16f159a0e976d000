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
// Design: a 128x128 C tile per 512-thread workgroup = 8 wave64s in a 4x2
// grid; each wave owns a 32x64 sub-tile built from 2x4 blocks of the
// 16x16x4 MFMA (f64: v_mfma_f64_16x16x4_f64, f32: v_mfma_f32_16x16x4_f32 --
// same operand layout).  The accumulators (64 registers per lane) fit the
// architectural VGPRs, so no accumulator copies and two waves per SIMD.  A
// is staged in LDS as [k][m] (m contiguous, the MFMA A-operand lane order),
// B as [n][k] padded by one element so the 16 lanes that read different n
// hit distinct banks.  Global->register prefetch of tile t+1 overlaps the
// MFMAs of tile t; LDS is double-buffered, one barrier per K tile; interior
// tiles run a predicate-free path with 16-byte loads.
#include <climits>
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
constexpr int GNT = 512;              // 8 waves
constexpr int GWM = 4, GWN = 2;       // wave grid: 4 along m, 2 along n
constexpr int GMI = GBM / GWM / 16;   // 16x16 blocks per wave along m (2)
constexpr int GNJ = GBN / GWN / 16;   // 16x16 blocks per wave along n (4)
constexpr int APAD = 16;              // [k][m] rows shifted by 32 banks
constexpr int BPAD = 1;               // [n][k] rows of 17 elements
constexpr int GRUN = GBM * GBK / GNT; // elements per thread per operand tile (4)
static_assert(GRUN == 4, "tile loader assumes 4 elements per thread");

// 4 consecutive elements base[(i0+q) + j*stride_c], q = 0..3, valid if
// i0+q < imax and j < jmax (zero otherwise); FAST = in bounds and aligned.
template <typename T, bool FAST>
FLX_DEV void load_run4(T (&r)[4], const T *__restrict__ base, long stride_c, int i0, int j,
                       int imax, int jmax)
{
    const T *p = base + (long)j * stride_c + i0;
    if (FAST) {
        if constexpr (sizeof(T) == 8) {
            const double2 *q = reinterpret_cast<const double2 *>(p);
            double2 v0 = q[0], v1 = q[1];
            r[0] = v0.x; r[1] = v0.y; r[2] = v1.x; r[3] = v1.y;
        } else {
            float4 v = *reinterpret_cast<const float4 *>(p);
            r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) r[q] = (j < jmax && i0 + q < imax) ? p[q] : T(0);
    }
}

// Per-thread slice of an A or B tile (4 consecutive elements along the
// operand's contiguous dimension):
//   A NoTrans: k = tid/32, m = (tid%32)*4      A Trans: m = tid/4, k = (tid%4)*4
//   B NoTrans: n = tid/4,  k = (tid%4)*4       B Trans: k = tid/32, n = (tid%32)*4
template <typename T, int TA, int TB, bool FAST>
struct TileIO {
    T ra[GRUN], rb[GRUN];

    __device__ __forceinline__ void load(const T *__restrict__ A, long lda,
                                         const T *__restrict__ B, long ldb, int M, int N,
                                         int K, int m0, int n0, int k0, int tid)
    {
        if (!TA) load_run4<T, FAST>(ra, A + m0, lda, (tid & 31) * 4, k0 + (tid >> 5), M - m0, K);
        else     load_run4<T, FAST>(ra, A + k0, lda, (tid & 3) * 4, m0 + (tid >> 2), K - k0, M);
        if (!TB) load_run4<T, FAST>(rb, B + k0, ldb, (tid & 3) * 4, n0 + (tid >> 2), K - k0, N);
        else     load_run4<T, FAST>(rb, B + n0, ldb, (tid & 31) * 4, k0 + (tid >> 5), N - n0, K);
    }

    __device__ __forceinline__ void store(T (*As)[GBM + APAD], T (*Bs)[GBK + BPAD], int tid)
    {
        if (!TA) {
            const int kk = tid >> 5, mm = (tid & 31) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[kk][mm + q] = ra[q];
        } else {
            const int mm = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) As[kk + q][mm] = ra[q];
        }
        if (!TB) {
            const int nn = tid >> 2, kk = (tid & 3) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[nn][kk + q] = rb[q];
        } else {
            const int kk = tid >> 5, nn = (tid & 31) * 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) Bs[nn + q][kk] = rb[q];
        }
    }
};

// GMI x GNJ blocks of 16x16x4 MFMAs over one K tile held in LDS.
template <typename T, typename acc_t>
FLX_DEV void mma_tile(acc_t (&acc)[GMI][GNJ], const T (*As)[GBM + APAD],
                      const T (*Bs)[GBK + BPAD], int wm, int wn, int lrow, int lk)
{
#pragma unroll
    for (int s = 0; s < GBK / 4; ++s) {
        T af[GMI], bf[GNJ];
#pragma unroll
        for (int i = 0; i < GMI; ++i) af[i] = As[s * 4 + lk][wm * (GMI * 16) + i * 16 + lrow];
#pragma unroll
        for (int j = 0; j < GNJ; ++j) bf[j] = Bs[wn * (GNJ * 16) + j * 16 + lrow][s * 4 + lk];
#pragma unroll
        for (int i = 0; i < GMI; ++i)
#pragma unroll
            for (int j = 0; j < GNJ; ++j) acc[i][j] = mfma16x16x4(af[i], bf[j], acc[i][j]);
    }
}

// K loop: register prefetch GPF tiles ahead (GPF + 1 register sets used in
// rotation, the loop unrolled by GPF + 1 so every set index is static), so a
// tile's loads have GPF iterations of MFMAs to arrive; LDS double buffer;
// one barrier per tile.  FAST = interior block with 16-byte aligned operands
// and K a multiple of the tile (no predicates at all).  GPF = 3 for short K
// loops on small grids (the LU's critical-path updates: few tiles in flight,
// exposed load latency); GPF = 1, one set, otherwise (measured faster for
// long K and large grids).
template <typename T, int TA, int TB, bool FAST, int GPF, typename acc_t>
FLX_DEV void gemm_mainloop(acc_t (&acc)[GMI][GNJ], T (*As)[GBK][GBM + APAD],
                           T (*Bs)[GBN][GBK + BPAD], const T *__restrict__ A, long lda,
                           const T *__restrict__ B, long ldb, int M, int N, int K, int m0,
                           int n0, int nkt, int tid, int wm, int wn, int lrow, int lk)
{
    if constexpr (GPF == 1) {
        TileIO<T, TA, TB, FAST> io;
        io.load(A, lda, B, ldb, M, N, K, m0, n0, 0, tid);
        io.store(As[0], Bs[0], tid);
        __syncthreads();
        for (int kt = 0; kt < nkt; ++kt) {
            const int buf = kt & 1;
            const bool more = kt + 1 < nkt;
            if (more) io.load(A, lda, B, ldb, M, N, K, m0, n0, (kt + 1) * GBK, tid);
            mma_tile<T>(acc, As[buf], Bs[buf], wm, wn, lrow, lk);
            if (more) {
                io.store(As[buf ^ 1], Bs[buf ^ 1], tid);
                __syncthreads();
            }
        }
        return;
    }
    constexpr int NS = GPF + 1;
    TileIO<T, TA, TB, FAST> io[NS];
#pragma unroll
    for (int p = 0; p < GPF; ++p)
        if (p < nkt) io[p].load(A, lda, B, ldb, M, N, K, m0, n0, p * GBK, tid);
    io[0].store(As[0], Bs[0], tid);
    __syncthreads();
    for (int kt0 = 0; kt0 < nkt; kt0 += NS) {
#pragma unroll
        for (int u = 0; u < NS; ++u) {
            const int kt = kt0 + u;
            if (kt >= nkt) break;
            // tile t lives in set t % NS: set (u + GPF) % NS held tile kt - 1
            // (in LDS since the previous iteration) and takes tile kt + GPF
            if (kt + GPF < nkt)
                io[(u + GPF) % NS].load(A, lda, B, ldb, M, N, K, m0, n0, (kt + GPF) * GBK, tid);
            mma_tile<T>(acc, As[kt & 1], Bs[kt & 1], wm, wn, lrow, lk);
            if (kt + 1 < nkt) {
                io[(u + 1) % NS].store(As[(kt + 1) & 1], Bs[(kt + 1) & 1], tid);
                __syncthreads();
            }
        }
    }
}

// One 128x128 tile of C.  The caller separates two tiles of one workgroup
// by a barrier (the LDS buffers are reused).
template <typename T, int TA, int TB, int GPF>
FLX_DEV void gemm_tile(T (*As)[GBK][GBM + APAD], T (*Bs)[GBN][GBK + BPAD], int m0, int n0,
                       int M, int N, int K, T alpha, const T *__restrict__ A, long lda,
                       const T *__restrict__ B, long ldb, T beta, T *__restrict__ C, long ldc,
                       int vecA, int vecB, int blas_beta0)
{
    typedef typename MfmaAcc<T>::type acc_t;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % GWM, wn = wave / GWM;
    const int lrow = lane & 15, lk = lane >> 4;

    acc_t acc[GMI][GNJ];
#pragma unroll
    for (int i = 0; i < GMI; ++i)
#pragma unroll
        for (int j = 0; j < GNJ; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    const int nkt = (alpha == T(0)) ? 0 : (K + GBK - 1) / GBK;
    if (nkt > 0) {
        const bool fast = (m0 + GBM <= M) && (n0 + GBN <= N) && (K % GBK == 0) && vecA && vecB;
        if (fast)
            gemm_mainloop<T, TA, TB, true, GPF>(acc, As, Bs, A, lda, B, ldb, M, N, K, m0, n0,
                                                nkt, tid, wm, wn, lrow, lk);
        else
            gemm_mainloop<T, TA, TB, false, GPF>(acc, As, Bs, A, lda, B, ldb, M, N, K, m0, n0,
                                                 nkt, tid, wm, wn, lrow, lk);
    }

    // Epilogue: see MfmaAcc for which row each accumulator item holds.  The
    // tile's C entries are all requested before any is used (one memory
    // latency, not one per entry); generic FLENS order: C*beta first (0*C
    // for beta=0); BLAS convention (blas_beta0): beta=0 never reads C.
    const bool readc = !(blas_beta0 && beta == T(0));
    T cv[GMI][GNJ][4];
#pragma unroll
    for (int i = 0; i < GMI; ++i)
#pragma unroll
        for (int j = 0; j < GNJ; ++j) {
            const int col = n0 + wn * (GNJ * 16) + j * 16 + lrow;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm * (GMI * 16) + i * 16 + MfmaAcc<T>::row(lk, q);
                const bool ok = readc && col < N && row < M;
                cv[i][j][q] = ok ? C[(long)col * ldc + row] : T(0);
            }
        }
#pragma unroll
    for (int i = 0; i < GMI; ++i)
#pragma unroll
        for (int j = 0; j < GNJ; ++j) {
            const int col = n0 + wn * (GNJ * 16) + j * 16 + lrow;
            if (col >= N) continue;
            T *cp = C + (long)col * ldc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = m0 + wm * (GMI * 16) + i * 16 + MfmaAcc<T>::row(lk, q);
                if (row < M) {
                    const T c = readc ? cv[i][j][q] * beta : T(0);
                    cp[row] = (nkt > 0) ? c + alpha * acc[i][j][q] : c;
                }
            }
        }
}

// Tile order.  The hardware hands workgroup b to XCD b mod 8, and each XCD
// has its own L2, so the dispatch sequence t (workgroup b takes t = b,
// b + grid, ...) is mapped to tiles such that the tiles of one XCD form a
// contiguous stretch of a grouped order: bands of GBAND tile-rows, each band
// walked column by column.  The tiles an XCD runs at once then share a few
// A row-panels and B column-panels in its L2 instead of touching eight
// times as many.  Both maps are bijections on [0, tiles).
constexpr int GXCD = 8, GBAND = 8;
FLX_DEV void gemm_tile_of(int t, int mt, int nt, int tiles, int &mi, int &ni)
{
    const int per = tiles / GXCD, rem = tiles % GXCD;
    const int x = t % GXCD, s = t / GXCD;
    const int L = x * per + min(x, rem) + s;          // contiguous stretch per XCD
    const int band = L / (GBAND * nt);
    const int h = min(GBAND, mt - band * GBAND);      // rows in this band
    const int loc = L - band * GBAND * nt;
    mi = band * GBAND + loc % h;
    ni = loc / h;
}

// Workgroup b takes dispatch slots b, b + grid, ... (one tile each unless
// the launch caps the grid so that concurrent work on another stream keeps
// some compute units).
template <typename T, int TA, int TB, int GPF>
__global__ __launch_bounds__(GNT) void gemm_mfma_kernel(
    int M, int N, int K, T alpha, const T *__restrict__ A, long lda,
    const T *__restrict__ B, long ldb, T beta, T *__restrict__ C, long ldc,
    int vecA, int vecB, int blas_beta0, int mt, int tiles)
{
    __shared__ T As[2][GBK][GBM + APAD];
    __shared__ T Bs[2][GBN][GBK + BPAD];
    const int nt = tiles / mt;
    for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
        if (t != (int)blockIdx.x) __syncthreads();
        int mi, ni;
        gemm_tile_of(t, mt, nt, tiles, mi, ni);
        gemm_tile<T, TA, TB, GPF>(As, Bs, mi * GBM, ni * GBN, M, N, K, alpha, A, lda, B,
                                  ldb, beta, C, ldc, vecA, vecB, blas_beta0);
    }
}

// Narrow C with short K (the LU's critical-path updates: M columns of the
// row-major matrix, K = nb): the tiled kernel would leave most compute units
// idle and, for M < 128, most of each tile's MFMAs padding.  One wave per
// 16 x 16 block of C: every operand element the block needs is requested
// from global memory at once (no LDS, no barrier), then K / 4 chained
// 16x16x4 MFMAs in the same k order as the tiled kernel, so the result is
// that kernel's, bit for bit, for K a multiple of 16.  FP64: M, K <= 128;
// FP32: M, K <= 256.
template <typename T>
struct SkinnyLimits;
template <> struct SkinnyLimits<double> { static constexpr int M = 128, K = 128; };
template <> struct SkinnyLimits<float> { static constexpr int M = 256, K = 256; };

template <typename T>
__global__ __launch_bounds__(64) void gemm_skinny_kernel(int M, int N, int K, T alpha,
                                                         const T *__restrict__ A, long lda,
                                                         const T *__restrict__ B, long ldb,
                                                         T beta, T *__restrict__ C, long ldc,
                                                         int blas_beta0, int mt)
{
    typedef typename MfmaAcc<T>::type acc_t;
    constexpr int KS = SkinnyLimits<T>::K / 4;   // MFMA steps at most
    const int lane = threadIdx.x, lrow = lane & 15, lk = lane >> 4;
    const int m0 = (blockIdx.x % mt) * 16, n0 = (blockIdx.x / mt) * 16;
    const int nks = (alpha == T(0)) ? 0 : (K + 3) / 4;
    const int m = m0 + lrow, n = n0 + lrow;
    T a[KS], b[KS];
#pragma unroll
    for (int st = 0; st < KS; ++st) {
        const int k = 4 * st + lk;
        a[st] = (st < nks && m < M && k < K) ? A[m + (long)k * lda] : T(0);
        b[st] = (st < nks && n < N && k < K) ? B[k + (long)n * ldb] : T(0);
    }
    const bool readc = !(blas_beta0 && beta == T(0));
    T cv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = m0 + MfmaAcc<T>::row(lk, q);
        cv[q] = (readc && row < M && n < N) ? C[row + (long)n * ldc] : T(0);
    }
    acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int st = 0; st < KS; ++st)
        if (st < nks) acc = mfma16x16x4(a[st], b[st], acc);
    if (n >= N) return;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int row = m0 + MfmaAcc<T>::row(lk, q);
        if (row < M) {
            const T c = readc ? cv[q] * beta : T(0);
            C[row + (long)n * ldc] = (nks > 0) ? c + alpha * acc[q] : c;
        }
    }
}

template <typename T>
static bool launch_skinny(hipStream_t s, int ta, int tb, int M, int N, int K, T alpha,
                          const T *A, long lda, const T *B, long ldb, T beta, T *C, long ldc,
                          int blas_beta0, int max_groups, hipError_t &err)
{
    if (ta || tb || M <= 0 || N <= 0 || M > SkinnyLimits<T>::M || K > SkinnyLimits<T>::K ||
        max_groups > 0)
        return false;
    const int mt = (M + 15) / 16, nt = (N + 15) / 16;
    if ((long)mt * nt > INT_MAX) return false;
    hipLaunchKernelGGL(gemm_skinny_kernel<T>, dim3(mt * nt), dim3(64), 0, s, M, N, K, alpha, A,
                       lda, B, ldb, beta, C, ldc, blas_beta0, mt);
    err = hipGetLastError();
    return true;
}

template <typename T>
static hipError_t launch_gemm(hipStream_t s, int ta, int tb, int M, int N, int K, T alpha,
                              const T *A, long lda, const T *B, long ldb, T beta, T *C,
                              long ldc, int blas_beta0, int max_groups)
{
    if (M <= 0 || N <= 0) return hipSuccess;
    const int mt = (M + GBM - 1) / GBM, nt = (N + GBN - 1) / GBN;
    if ((long)mt * nt > INT_MAX) return hipErrorInvalidValue;
    const int tiles = mt * nt;
    dim3 grid(max_groups > 0 && max_groups < tiles ? max_groups : tiles);
    const uintptr_t al = 16;
    int vecA = ((uintptr_t)A % al == 0) && (lda * (long)sizeof(T)) % al == 0;
    int vecB = ((uintptr_t)B % al == 0) && (ldb * (long)sizeof(T)) % al == 0;
    // short K loops on at most one workgroup per compute unit (the LU's
    // critical-path updates, K = nb) get the deep prefetch
    const bool deep = K <= 256 && (int)grid.x <= 256;
#define FLX_LG(a_, b_, p_)                                                               \
    hipLaunchKernelGGL((gemm_mfma_kernel<T, a_, b_, p_>), grid, dim3(GNT), 0, s, M, N, K, \
                       alpha, A, lda, B, ldb, beta, C, ldc, vecA, vecB, blas_beta0, mt, tiles)
    if (!ta && !tb) { if (deep) FLX_LG(0, 0, 3); else FLX_LG(0, 0, 1); }
    else if (!ta && tb) { if (deep) FLX_LG(0, 1, 3); else FLX_LG(0, 1, 1); }
    else if (ta && !tb) { if (deep) FLX_LG(1, 0, 3); else FLX_LG(1, 0, 1); }
    else { if (deep) FLX_LG(1, 1, 3); else FLX_LG(1, 1, 1); }
#undef FLX_LG
    return hipGetLastError();
}

hipError_t gemm_f64(hipStream_t s, int ta, int tb, int M, int N, int K, double alpha,
                    const double *A, long lda, const double *B, long ldb, double beta,
                    double *C, long ldc, int blas_beta0, int max_groups)
{
    hipError_t err = hipSuccess;
    if (launch_skinny<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                              blas_beta0, max_groups, err))
        return err;
    return launch_gemm<double>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                               blas_beta0, max_groups);
}

hipError_t gemm_f32(hipStream_t s, int ta, int tb, int M, int N, int K, float alpha,
                    const float *A, long lda, const float *B, long ldb, float beta, float *C,
                    long ldc, int blas_beta0, int max_groups)
{
    hipError_t err = hipSuccess;
    if (launch_skinny<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                             blas_beta0, max_groups, err))
        return err;
    return launch_gemm<float>(s, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                              blas_beta0, max_groups);
}

} // namespace flx
