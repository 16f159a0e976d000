// getrs.hip -- solves with the LU factors (FP64, column-major).
//
// Replaces flens::lapack::trs (flens/lapack/ge/trs.tcc:59-109): NoTrans =
// laswp(B,piv), L-unit solve, U solve; Trans = U' solve, L' solve,
// laswp(B, piv.reverse()).  The two triangular solves are the
// blas::sm -> cxxblas::trsm -> trsv chain of cxxblas/level3/trsm.tcc:40-68.
//
// Row interchanges are applied as one gather with the composed permutation
// (perm[i] = original row that ends up in row i), and the triangular solves
// are blocked by 64 rows: one launch per diagonal block solves the block in
// LDS and folds it into the remaining rows (GEMM-shaped, coalesced reads of
// the factor columns).
#include <algorithm>
#include <climits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

// ---------------------------------------------------------------------------
// perm from LAPACK ipiv (1-based, row i swapped with ipiv[i]); reverse=1
// composes the swaps last to first.  Sequential by nature; one lane walks
// the swaps in LDS (n <= 16384) or in global memory.
// ---------------------------------------------------------------------------
constexpr int PERM_LDS = 16384;

__global__ void ipiv_to_perm_kernel(const int *__restrict__ ipiv, int *__restrict__ perm,
                                    int n, int reverse)
{
    __shared__ int s[PERM_LDS];
    __shared__ int sp[PERM_LDS];   // the pivots, staged so the serial loop reads LDS only
    const bool lds = n <= PERM_LDS;
    int *p = lds ? s : perm;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        p[i] = i;
        if (lds) sp[i] = ipiv[i];
    }
    __syncthreads();
    const int *pv = lds ? sp : ipiv;
    if (threadIdx.x == 0) {
        if (!reverse) {
            for (int i = 0; i < n; ++i) {
                int q = pv[i] - 1;
                if (q != i && q >= 0 && q < n) { int t = p[i]; p[i] = p[q]; p[q] = t; }
            }
        } else {
            for (int i = n - 1; i >= 0; --i) {
                int q = pv[i] - 1;
                if (q != i && q >= 0 && q < n) { int t = p[i]; p[i] = p[q]; p[q] = t; }
            }
        }
    }
    __syncthreads();
    if (lds)
        for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = s[i];
}

hipError_t ipiv_to_perm(hipStream_t s, const int *ipiv, int *perm, int n, int reverse)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ipiv_to_perm_kernel, dim3(1), dim3(256), 0, s, ipiv, perm, n, reverse);
    return hipGetLastError();
}

__global__ void gather_rows_kernel(const double *__restrict__ B, long ldb,
                                   double *__restrict__ out, long ldo,
                                   const int *__restrict__ perm, int n, int nrhs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < n) out[(long)c * ldo + i] = B[(long)c * ldb + perm[i]];
}

hipError_t gather_rows_cm(hipStream_t s, const double *B, long ldb, double *out, long ldo,
                          const int *perm, int n, int nrhs)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, B,
                       ldb, out, ldo, perm, n, nrhs);
    return hipGetLastError();
}

__global__ void copy_cm_kernel(const double *__restrict__ in, long ldi,
                               double *__restrict__ out, long ldo, int m)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int c = blockIdx.y;
    if (i < m) out[(long)c * ldo + i] = in[(long)c * ldi + i];
}

hipError_t copy_cm(hipStream_t s, const double *in, long ldi, double *out, long ldo, int m,
                   int n)
{
    if (m <= 0 || n <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_cm_kernel, dim3((m + 255) / 256, n), dim3(256), 0, s, in, ldi, out,
                       ldo, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Blocked triangular solve.  T(i,q) is the triangular operand:
//   FWD (forward order, T lower): L (unit, NoTrans) or U^T (non-unit, Trans)
//   BWD (backward order, T upper): U (non-unit, NoTrans) or L^T (unit, Trans)
// TR selects the element address: NoTrans T(i,q) = LU[i + q*lda],
// Trans T(i,q) = LU[q + i*lda].
//
// 1. tri_inv_kernel inverts every 64x64 diagonal block of T at once (one
//    workgroup per block, one column of the inverse per thread, registers).
// 2. One launch per diagonal block: X_k = inv(T_kk) * B_k (no sequential
//    chain), block 0 stores X_k, and every workgroup folds X_k into its 256
//    trailing rows of B (coalesced reads of the factor's columns).
// The solution goes to a separate array so no workgroup can read a block
// row another one already overwrote.
// ---------------------------------------------------------------------------
constexpr int TS_BS = 64;
constexpr int TS_CG = 16;   // right-hand sides per pass

template <int TR, int UNIT, int FWD>
__global__ __launch_bounds__(64) void tri_inv_kernel(const double *__restrict__ LU, long lda,
                                                     int n, double *__restrict__ inv)
{
    __shared__ double ts[TS_BS][TS_BS + 1];
    __shared__ double xs[TS_BS][TS_BS + 1];   // xs[i][t]: column t of inv(T)
    const int b = blockIdx.x, kb = b * TS_BS, bs = min(TS_BS, n - kb);
    const int t = threadIdx.x;
    // T(i,q) for the block, identity outside the matrix
    for (int i = 0; i < TS_BS; ++i) {
        double v;
        if (i < bs && t < bs) v = TR ? LU[(long)(kb + i) * lda + kb + t] : LU[(long)(kb + t) * lda + kb + i];
        else v = (i == t) ? 1.0 : 0.0;
        ts[i][t] = v;
    }
    __syncthreads();
    // column t of inv(T): T x = e_t, by substitution (rows only this thread touches)
    for (int s = 0; s < TS_BS; ++s) {
        const int i = FWD ? s : TS_BS - 1 - s;
        double v = (i == t) ? 1.0 : 0.0;
        if (FWD) {
#pragma unroll 8
            for (int q = 0; q < i; ++q) v = fma(-ts[i][q], xs[q][t], v);
        } else {
#pragma unroll 8
            for (int q = i + 1; q < TS_BS; ++q) v = fma(-ts[i][q], xs[q][t], v);
        }
        xs[i][t] = UNIT ? v : v / ts[i][i];
    }
    double *o = inv + (size_t)b * TS_BS * TS_BS;   // row-major 64x64
    for (int i = 0; i < TS_BS; ++i) o[i * TS_BS + t] = xs[i][t];
}

template <int TR, int FWD>
__global__ __launch_bounds__(256) void trs_block_kernel(const double *__restrict__ LU, long lda,
                                                        const double *__restrict__ inv,
                                                        double *__restrict__ B, long ldb,
                                                        double *__restrict__ X, long ldx,
                                                        int n, int nrhs, int kb, int bs)
{
    __shared__ double xs[TS_BS][TS_CG + 1];
    __shared__ double is[TS_BS][TS_BS + 1];   // inv(T_kk)
    __shared__ double ys[TS_BS][TS_CG + 1];   // X_k
    const int tid = threadIdx.x;
    // thread -> (row ri of 64, right-hand sides 4*cq .. 4*cq+3): a wave
    // covers 64 consecutive rows, so factor-column reads coalesce
    const int ri = tid & 63, cq = tid >> 6;
    auto T = [&](int i, int q) -> double {
        return TR ? LU[(long)i * lda + q] : LU[(long)q * lda + i];
    };
    const double *ib = inv + (size_t)(kb / TS_BS) * TS_BS * TS_BS;
    for (int q = tid; q < TS_BS * TS_BS; q += 256) is[q / TS_BS][q % TS_BS] = ib[q];
    // trailing rows: FWD -> [kb+bs, n), BWD -> [0, kb)
    const int t0 = FWD ? kb + bs : 0;
    const int tn = FWD ? n - (kb + bs) : kb;
    const int r = t0 + blockIdx.x * 64 + ri;
    const bool has_row = (blockIdx.x * 64 + ri) < tn;

    for (int cg = 0; cg < nrhs; cg += TS_CG) {
        const int nc = min(TS_CG, nrhs - cg);
        for (int q = tid; q < TS_BS * TS_CG; q += 256) {
            int i = q / TS_CG, c = q % TS_CG;
            xs[i][c] = (c < nc && i < bs) ? B[(long)(cg + c) * ldb + kb + i] : 0.0;
        }
        __syncthreads();
        // X_k = inv(T_kk) B_k
        {
            double acc[4] = {0, 0, 0, 0};
#pragma unroll 8
            for (int q = 0; q < TS_BS; ++q) {
                const double tv = is[ri][q];
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = fma(tv, xs[q][4 * cq + c], acc[c]);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) ys[ri][4 * cq + c] = acc[c];
        }
        __syncthreads();
        if (blockIdx.x == 0)
            for (int q = tid; q < bs * TS_CG; q += 256) {
                int i = q / TS_CG, c = q % TS_CG;
                if (c < nc) X[(long)(cg + c) * ldx + kb + i] = ys[i][c];
            }
        if (has_row) {
            double acc[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc[c] = (4 * cq + c < nc) ? B[(long)(cg + 4 * cq + c) * ldb + r] : 0.0;
            if (bs == TS_BS) {
#pragma unroll 16
                for (int q = 0; q < TS_BS; ++q) {
                    const double t = T(r, kb + q);
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[c] = fma(-t, ys[q][4 * cq + c], acc[c]);
                }
            } else {
                for (int q = 0; q < bs; ++q) {
                    const double t = T(r, kb + q);
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[c] = fma(-t, ys[q][4 * cq + c], acc[c]);
                }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (4 * cq + c < nc) B[(long)(cg + 4 * cq + c) * ldb + r] = acc[c];
        }
        __syncthreads();
    }
}

size_t trs_inv_doubles(int n) { return (size_t)((n + TS_BS - 1) / TS_BS) * TS_BS * TS_BS; }

hipError_t trsm_cm_f64(hipStream_t s, int upper, int trans, int unit, int n, int nrhs,
                       const double *LU, long lda, double *B, long ldb, double *X,
                       long ldx, double *inv)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    // forward order when the operand is lower triangular
    const int fwd = (!upper && !trans) || (upper && trans);
    const int nblk = (n + TS_BS - 1) / TS_BS;
    {
        ProfScope ps(tl_prof, s, PROF_TRS, (double)nblk * TS_BS * TS_BS * TS_BS / 3.0,
                     8.0 * nblk * TS_BS * TS_BS * 1.5);
#define FLX_TI(tr_, un_, fw_)                                                                \
    hipLaunchKernelGGL((tri_inv_kernel<tr_, un_, fw_>), dim3(nblk), dim3(64), 0, s, LU, lda, n, inv)
        if (!trans && unit && fwd) FLX_TI(0, 1, 1);
        else if (!trans && !unit && !fwd) FLX_TI(0, 0, 0);
        else if (trans && !unit && fwd) FLX_TI(1, 0, 1);
        else if (trans && unit && !fwd) FLX_TI(1, 1, 0);
        else if (!trans && unit && !fwd) FLX_TI(0, 1, 0);
        else if (!trans && !unit && fwd) FLX_TI(0, 0, 1);
        else if (trans && unit && fwd) FLX_TI(1, 1, 1);
        else FLX_TI(1, 0, 0);
#undef FLX_TI
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int b = 0; b < nblk; ++b) {
        const int blk = fwd ? b : nblk - 1 - b;
        const int kb = blk * TS_BS;
        const int bs = min(TS_BS, n - kb);
        const int tn = fwd ? n - (kb + bs) : kb;
        const int grid = tn > 0 ? (tn + 63) / 64 : 1;
        // block product + fold into tn rows: factor block and rows read once
        ProfScope ps(tl_prof, s, PROF_TRS, (double)nrhs * (2.0 * bs * bs + 2.0 * bs * tn),
                     8.0 * ((double)bs * bs + (double)bs * tn + 2.0 * nrhs * (bs + tn)));
        if (!trans && fwd) hipLaunchKernelGGL((trs_block_kernel<0, 1>), dim3(grid), dim3(256), 0, s, LU, lda, inv, B, ldb, X, ldx, n, nrhs, kb, bs);
        else if (!trans) hipLaunchKernelGGL((trs_block_kernel<0, 0>), dim3(grid), dim3(256), 0, s, LU, lda, inv, B, ldb, X, ldx, n, nrhs, kb, bs);
        else if (fwd) hipLaunchKernelGGL((trs_block_kernel<1, 1>), dim3(grid), dim3(256), 0, s, LU, lda, inv, B, ldb, X, ldx, n, nrhs, kb, bs);
        else hipLaunchKernelGGL((trs_block_kernel<1, 0>), dim3(grid), dim3(256), 0, s, LU, lda, inv, B, ldb, X, ldx, n, nrhs, kb, bs);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Single-launch triangular solve (NoTrans L then U), column-major FP64
// factor, up to 16 right-hand sides per launch.  One 256-thread workgroup
// per 64-row block; workgroup g owns block g (L) or nbk-1-g (U), so every
// block it depends on belongs to a workgroup dispatched before it.
//   * A finished block publishes its 64 x 16 solution values as pairs of
//     64-bit words (epoch << 32 | half of the double), so one agent-scope
//     load tests readiness and delivers the value; waits are bounded.
//   * Dependency phase: lane = row, wave w = columns w*16..+15 of each
//     finished block, all 16 right-hand sides; the factor entries of a pair
//     of blocks are requested before waiting for them.
//   * Diagonal block: wave w substitutes right-hand sides 4w..4w+3, one row
//     per lane, the solved entry broadcast by readlane.
// ---------------------------------------------------------------------------
constexpr int TP_BS = 64, TP_RHS = 16, TP_GRP = 2;
constexpr int TP_SPIN_LIMIT = 1 << 22;

FLX_DEV double tp_readlane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// FWD: blocks in ascending order (a lower triangle); UNIT: unit diagonal;
// TRANS: the triangle is read transposed (op = T^T: row i of the operand is
// column i of the stored factor).  R right-hand sides per launch (1, 4 or
// 16), GRP dependency blocks per wait.
template <int FWD, int UNIT, int TRANS, int R, int GRP>
__global__ __launch_bounds__(256) void trs_persist_kernel(const double *__restrict__ LU, long lda,
                                                          int n, int nrhs,
                                                          const double *__restrict__ Bin, long ldb,
                                                          double *__restrict__ Xout, long ldx,
                                                          uint64_t *__restrict__ xt,
                                                          unsigned epoch, int *__restrict__ err)
{
    __shared__ double s_x[2][GRP][TP_BS][R + 1];
    __shared__ double s_red[4][TP_BS][R + 1];
    __shared__ double s_d[TP_BS][TP_BS + 1];
    __shared__ int s_ok;
    const int nbk = (n + TP_BS - 1) / TP_BS;
    const int g = blockIdx.x;
    const int bi = FWD ? g : nbk - 1 - g;
    const int r0 = bi * TP_BS, rows = min(TP_BS, n - r0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ndep = FWD ? bi : nbk - 1 - bi;
    const bool row_ok = lane < rows;
    const int row = r0 + (row_ok ? lane : 0);
    // the diagonal block, staged in LDS at the start (identity outside the matrix)
    for (int q = tid; q < TP_BS * TP_BS; q += 256) {
        const int c = q / TP_BS, r = q % TP_BS;
        s_d[r][c] = (r < rows && c < rows)
                        ? (TRANS ? LU[(long)(r0 + r) * lda + r0 + c] : LU[(long)(r0 + c) * lda + r0 + r])
                        : (r == c ? 1.0 : 0.0);
    }
    double acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = 0.0;
    int par = 0, spins = 0;
    for (int d0 = 0; d0 < ndep; d0 += GRP, par ^= 1) {
        const int cnt = min(GRP, ndep - d0);
        double lv[GRP][16];
#pragma unroll
        for (int b = 0; b < GRP; ++b) {
            const int d = d0 + b;
            const int jb = FWD ? d : nbk - 1 - d;
            const int c0 = jb * TP_BS + wave * 16;
#pragma unroll
            for (int c = 0; c < 16; ++c)
                lv[b][c] = (b < cnt && row_ok && c0 + c < n)
                               ? (TRANS ? LU[(long)row * lda + c0 + c] : LU[(long)(c0 + c) * lda + row])
                               : 0.0;
        }
        // one thread waits for each block's sentinel word (written last by
        // its producer's thread), then every thread fetches 8 words per
        // block and all verify their tags (a late word repeats the wait)
        for (;;) {
            if (tid == 0) {
                for (;;) {
                    bool ready = true;
                    for (int b = 0; b < cnt; ++b) {
                        const int d = d0 + b;
                        const int jb = FWD ? d : nbk - 1 - d;
                        const uint64_t *w = xt + ((long)(jb * TP_BS + TP_BS - 1) * R + R - 1) * 2;
                        ready = ready && (unsigned)(__hip_atomic_load(w + 1, __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_AGENT) >> 32) == epoch;
                    }
                    // every poll counts, also those after a failed verification
                    if (++spins > TP_SPIN_LIMIT || ready) break;
                    __builtin_amdgcn_s_sleep(2);
                }
                s_ok = spins > TP_SPIN_LIMIT;
            }
            __syncthreads();
            const bool gave_up = s_ok;
            bool ok = true;
#pragma unroll
            for (int b = 0; b < GRP; ++b) {
                if (b < cnt) {
                    const int d = d0 + b;
                    const int jb = FWD ? d : nbk - 1 - d;
#pragma unroll
                    for (int idx = tid; idx < TP_BS * R; idx += 256) {   // row * R + rhs
                        const int rr = idx / R, k = idx % R;
                        const uint64_t *w = xt + ((long)(jb * TP_BS + rr) * R + k) * 2;
                        const uint64_t hi = __hip_atomic_load(w, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                        const uint64_t lo = __hip_atomic_load(w + 1, __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                        ok = ok && (unsigned)(hi >> 32) == epoch && (unsigned)(lo >> 32) == epoch;
                        s_x[par][b][rr][k] = __longlong_as_double(
                            (long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
                    }
                }
            }
            const int all = __syncthreads_and(ok);
            if (all) break;
            if (gave_up) {                 // go on with what is there
                if (tid == 0) atomicOr(err, 2);
                break;
            }
        }
        // only the blocks of this group: the slots of a short last group hold
        // another launch's LDS contents, and 0 * Inf / NaN is not 0
#pragma unroll
        for (int b = 0; b < GRP; ++b)
            if (b < cnt) {
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const double l = lv[b][c];
#pragma unroll
                    for (int k = 0; k < R; ++k) acc[k] = fma(l, s_x[par][b][wave * 16 + c][k], acc[k]);
                }
            }
    }
#pragma unroll
    for (int k = 0; k < R; ++k) s_red[wave][lane][k] = acc[k];
    __syncthreads();
    // right-hand sides of this block, wave w: KK of them from w * KK
    constexpr int KK = (R + 3) / 4;
    double v[KK];
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
        const int k = wave * KK + kk;
        double t = 0.0;
        if (row_ok && k < nrhs && k < R) {
            t = Bin[(long)k * ldb + r0 + lane];
#pragma unroll
            for (int w = 0; w < 4; ++w) t -= s_red[w][lane][k];
        }
        v[kk] = t;
    }
    double drow[TP_BS];
#pragma unroll
    for (int c = 0; c < TP_BS; ++c) drow[c] = s_d[lane][c];
    // the chain multiplies by the reciprocal of the row's diagonal entry
    // (a division per step would sit on the critical path)
    const double dinv = UNIT ? 1.0 : 1.0 / s_d[lane][lane];
    if (FWD) {
#pragma unroll
        for (int c = 0; c < TP_BS; ++c) {
#pragma unroll
            for (int kk = 0; kk < KK; ++kk) {
                if (!UNIT && lane == c) v[kk] = v[kk] * dinv;
                const double xc = tp_readlane(v[kk], c);
                if (lane > c) v[kk] = fma(-drow[c], xc, v[kk]);
            }
        }
    } else {
#pragma unroll
        for (int c = TP_BS - 1; c >= 0; --c) {
#pragma unroll
            for (int kk = 0; kk < KK; ++kk) {
                if (!UNIT && lane == c) v[kk] = v[kk] * dinv;
                const double xc = tp_readlane(v[kk], c);
                if (lane < c) v[kk] = fma(-drow[c], xc, v[kk]);
            }
        }
    }
    uint64_t wl = 0;
    uint64_t *wsent = nullptr;
#pragma unroll
    for (int kk = 0; kk < KK; ++kk) {
        const int k = wave * KK + kk;
        if (k >= R) continue;                 // waves beyond the last right-hand side
        const double x = (row_ok && k < nrhs) ? v[kk] : 0.0;
        if (row_ok && k < nrhs) Xout[(long)k * ldx + r0 + lane] = x;
        const uint64_t b = (uint64_t)__double_as_longlong(x);
        uint64_t *w = xt + ((long)(r0 + lane) * R + k) * 2;
        __hip_atomic_store(w, ((uint64_t)epoch << 32) | (b >> 32), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t lo = ((uint64_t)epoch << 32) | (b & 0xffffffffull);
        if (lane == TP_BS - 1 && k == R - 1) { wsent = w + 1; wl = lo; }
        else __hip_atomic_store(w + 1, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the sentinel goes last, after every other word of the block is out
    __threadfence();
    __syncthreads();
    if (wsent) __hip_atomic_store(wsent, wl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

size_t trs_persist_words(int n) { return (size_t)((n + TP_BS - 1) / TP_BS) * TP_BS * TP_RHS * 2; }

__global__ void zero_words_kernel(uint64_t *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// trans = 0: X = U^{-1} L^{-1} Bp (Bp already row-permuted);
// trans = 1: X = L^{-T} U^{-T} Bp (the caller applies P^T afterwards).
// T is n x 16 scratch.
hipError_t trs_persist_f64(hipStream_t s, int n, int nrhs, const double *LU, long lda,
                           const double *Bp, long ldbp, double *T, double *X, long ldx,
                           uint64_t *xt, unsigned *epoch, int *err, int trans)
{
    const int nbk = (n + TP_BS - 1) / TP_BS;
    if (nbk <= 0 || nrhs <= 0) return hipSuccess;
    auto next = [&]() -> unsigned {
        if (++*epoch == 0u) {
            const size_t nw = trs_persist_words(n);
            hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0,
                               s, xt, nw);
            *epoch = 1u;
        }
        return *epoch;
    };
    // 16 right-hand sides per pass; one or a few get a lighter variant (fewer
    // words per finished block, more blocks per wait).  First triangle: L
    // (forward, unit) or U^T (forward); second: U (backward) or L^T
    // (backward, unit).
    for (int k0 = 0; k0 < nrhs; k0 += TP_RHS) {
        const int kc = min(TP_RHS, nrhs - k0);
        ProfScope ps(tl_prof, s, PROF_TRS, 2.0 * (double)n * n * kc,
                     8.0 * ((double)n * n + 4.0 * n * kc));
#define FLX_TP(U1_, U2_, TR_, R_, G_)                                                            \
        do {                                                                                     \
            hipLaunchKernelGGL((trs_persist_kernel<1, U1_, TR_, R_, G_>), dim3(nbk), dim3(256),  \
                               0, s, LU, lda, n, kc, Bp + (long)k0 * ldbp, ldbp, T, (long)n, xt, \
                               next(), err);                                                     \
            hipLaunchKernelGGL((trs_persist_kernel<0, U2_, TR_, R_, G_>), dim3(nbk), dim3(256),  \
                               0, s, LU, lda, n, kc, (const double *)T, (long)n,                 \
                               X + (long)k0 * ldx, ldx, xt, next(), err);                        \
        } while (0)
        if (!trans) {
            if (kc == 1) FLX_TP(1, 0, 0, 1, 4);
            else if (kc <= 4) FLX_TP(1, 0, 0, 4, 4);
            else FLX_TP(1, 0, 0, 16, TP_GRP);
        } else {
            if (kc == 1) FLX_TP(0, 1, 1, 1, 4);
            else if (kc <= 4) FLX_TP(0, 1, 1, 4, 4);
            else FLX_TP(0, 1, 1, 16, TP_GRP);
        }
#undef FLX_TP
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Chain-and-helpers triangular solve (NoTrans: L unit forward, then U
// backward), column-major FP64 factor, up to 16 right-hand sides per launch,
// 64-row blocks, one launch per triangle after tri_inv_kernel has inverted
// every diagonal block.
//   * Workgroup 0 (the chain) walks the blocks in solve order g = 0, 1, ...:
//       v  = B_g - P_g - T(g, g-1) x_{g-1}      (x_{g-1} kept in LDS)
//       x_g = inv(T_gg) v                       (a 64 x 64 product, no
//                                                sequential substitution)
//     and publishes x_g as tagged words.  The factor block T(g+1, g), the
//     next inverse and the next right-hand side rows are requested one step
//     ahead, so the chain's step is two block products and two barriers.
//   * Workgroups 1 .. H (helpers) own blocks g = 2, 3, ... round robin and
//     fold every finished x_j, j <= g - 2, into P_g = sum_j T(g, j) x_j as
//     soon as it is published, then publish P_g -- the bulk of the work,
//     off the chain's critical path.
//   * Every published value is a pair of 64-bit words (epoch << 32 | half
//     of the double): a reader checks every word's tag, so producers need
//     no fence, and all waits are bounded (FLX flag bit 2 on a give-up).
// The grid is bounded by the CU count, so every workgroup is resident at
// once for any n.
// ---------------------------------------------------------------------------
constexpr int TC_BS = 64;
constexpr int TC_GRP = 2;        // helper: finished blocks per wait

FLX_DEV uint64_t tc_word_hi(unsigned epoch, double x)
{
    return ((uint64_t)epoch << 32) | ((uint64_t)__double_as_longlong(x) >> 32);
}
FLX_DEV uint64_t tc_word_lo(unsigned epoch, double x)
{
    return ((uint64_t)epoch << 32) | ((uint64_t)__double_as_longlong(x) & 0xffffffffull);
}
FLX_DEV double tc_join(uint64_t hi, uint64_t lo)
{
    return __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
}
FLX_DEV uint64_t tc_ld(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV void tc_st(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the pair of words of a tagged value (row r of block bi, right-hand side k)
template <int R>
FLX_DEV const uint64_t *tc_addr(const uint64_t *w, int bi, int r, int k)
{
    return w + ((long)(bi * TC_BS + r) * R + k) * 2;
}
// a value whose words were loaded ahead (hi, lo): re-polled only while a
// word still carries an older epoch; false on a give-up
FLX_DEV bool tc_settle(const uint64_t *p, uint64_t &hi, uint64_t &lo, unsigned epoch, int &spins)
{
    while ((unsigned)(hi >> 32) != epoch || (unsigned)(lo >> 32) != epoch) {
        if (++spins > TP_SPIN_LIMIT) return false;
        __builtin_amdgcn_s_sleep(1);
        hi = tc_ld(p);
        lo = tc_ld(p + 1);
    }
    return true;
}

// the chain's operands of step g into registers: the inverse of its
// diagonal block and the factor block left of it (columns of block g - 1),
// IPT entries of each per thread, coalesced; its right-hand side rows
template <int FWD, int R, int TR>
FLX_DEV void tc_fetch(int g, int nbk, int n, const double *__restrict__ LU, long lda,
                      const double *__restrict__ inv, const double *__restrict__ Bin, long ldb,
                      int nrhs, double (&inext)[TC_BS * TC_BS / 256],
                      double (&tnext)[TC_BS * TC_BS / 256], double (&bnext)[4])
{
    constexpr int IPT = TC_BS * TC_BS / 256;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int bi = FWD ? g : nbk - 1 - g, r0 = bi * TC_BS, rows = min(TC_BS, n - r0);
    const double *ib = inv + (size_t)bi * TC_BS * TC_BS;
#pragma unroll
    for (int u = 0; u < IPT; ++u) inext[u] = ib[tid + 256 * u];
    const int c0 = g > 0 ? (FWD ? g - 1 : nbk - g) * TC_BS : 0;
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
        // consecutive threads read consecutive addresses: down a column of a
        // column-major factor, along a row of a row-major one (TR)
        const int q = tid + 256 * u;
        const int c = TR ? q % TC_BS : q / TC_BS, r = TR ? q / TC_BS : q % TC_BS;
        tnext[u] = (g > 0 && r < rows && c0 + c < n)
                       ? (TR ? LU[(long)(r0 + r) * lda + c0 + c] : LU[(long)(c0 + c) * lda + r0 + r])
                       : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rr = 16 * wave + lk + 4 * q;
        bnext[q] = (rr < rows && lr < R && lr < nrhs)
                       ? (TR ? Bin[(long)(r0 + rr) * ldb + lr] : Bin[(long)lr * ldb + r0 + rr])
                       : 0.0;
    }
}

// a barrier for LDS traffic only: outstanding global loads (prefetches) and
// stores (published words) stay in flight across it
FLX_DEV void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// TR = 1: the factor and the right-hand sides are row-major (element (i, j)
// at LU[i * lda + j], Bin[i * ldb + j]); the solution is column-major.
template <int FWD, int R, int TR>
__global__ __launch_bounds__(256) void trs_chain_kernel(const double *__restrict__ LU, long lda,
                                                        int n, int nrhs,
                                                        const double *__restrict__ inv,
                                                        const double *__restrict__ Bin, long ldb,
                                                        double *__restrict__ Xout, long ldx,
                                                        uint64_t *__restrict__ xw,
                                                        uint64_t *__restrict__ pw,
                                                        unsigned epoch, int *__restrict__ err)
{
    constexpr int IPT = TC_BS * TC_BS / 256; // inverse entries per thread
    constexpr int XPT = (TC_BS * R + 255) / 256;   // block values per thread (helpers)
    __shared__ double s_inv[TC_BS][TC_BS + 1];
    __shared__ double s_t[TC_BS][TC_BS + 1];      // T(g, g-1), row-major
    __shared__ double s_v[TC_BS][17];              // chain: 16 right-hand side slots
    __shared__ double s_xc[TC_BS][17];
    __shared__ double s_x[2][TC_GRP][TC_BS][R + 1];
    __shared__ double s_red[4][TC_BS][R + 1];
    __shared__ int s_bad;
    const int nbk = (n + TC_BS - 1) / TC_BS;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto blk = [&](int g) { return FWD ? g : nbk - 1 - g; };
    if (tid == 0) s_bad = 0;
    __syncthreads();
    if (blockIdx.x == 0) {
        // ------------------------------------------------------------ chain
        // Wave w owns rows 16w .. 16w + 15 of the block and every right-hand
        // side: both block products are 16 x 16 x 64 MFMA tiles (sixteen
        // v_mfma_f64_16x16x4 steps per wave).  Lane (lr, lk) holds rows
        // 16w + lk + 4q (q = 0..3) of right-hand side lr.  Step g's operands
        // -- the inverse of its diagonal block, the factor block left of it
        // (columns of block g-1), its right-hand side rows -- are requested
        // during step g-1; the chain's barriers wait for LDS traffic only,
        // so those loads and the published stores stay in flight.
        const int lr = lane & 15, lk = lane >> 4;
        int spins = 0;
        bool bad = false;
        double inext[IPT], tnext[IPT], bnext[4];
        tc_fetch<FWD, R, TR>(0, nbk, n, LU, lda, inv, Bin, ldb, nrhs, inext, tnext, bnext);
        for (int g = 0; g < nbk; ++g) {
            const int bi = blk(g), r0 = bi * TC_BS, rows = min(TC_BS, n - r0);
#pragma unroll
            for (int u = 0; u < IPT; ++u) {
                const int q = tid + 256 * u;
                s_inv[q / TC_BS][q % TC_BS] = inext[u];
                if (TR) s_t[q / TC_BS][q % TC_BS] = tnext[u];
                else s_t[q % TC_BS][q / TC_BS] = tnext[u];
            }
            double v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = bnext[q];
            lds_barrier();
            // P_g, published by a helper (g >= 2): every word requested first
            if (g >= 2) {
                uint64_t ph[4], pl[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ph[q] = pl[q] = (uint64_t)epoch << 32;
                    if (lr < R) {
                        const uint64_t *p = tc_addr<R>(pw, bi, 16 * wave + lk + 4 * q, lr);
                        ph[q] = tc_ld(p);
                        pl[q] = tc_ld(p + 1);
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (lr < R) {
                        if (!bad)
                            bad = !tc_settle(tc_addr<R>(pw, bi, 16 * wave + lk + 4 * q, lr), ph[q],
                                             pl[q], epoch, spins);
                        v[q] -= tc_join(ph[q], pl[q]);
                    }
                }
            }
            // v -= T(g, g-1) x_{g-1}
            if (g > 0) {
                // four accumulators: four independent MFMA chains of four steps
                double4_t acc[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) acc[a] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int q = 0; q < TC_BS / 4; ++q)
                    acc[q & 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        s_t[16 * wave + lr][4 * q + lk], s_xc[4 * q + lk][lr], acc[q & 3], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; ++q) v[q] -= (acc[0][q] + acc[1][q]) + (acc[2][q] + acc[3][q]);
            }
            // the next step's operands are requested now and arrive during
            // this step's second product and stores
            if (g + 1 < nbk)
                tc_fetch<FWD, R, TR>(g + 1, nbk, n, LU, lda, inv, Bin, ldb, nrhs, inext, tnext, bnext);
#pragma unroll
            for (int q = 0; q < 4; ++q) s_v[16 * wave + lk + 4 * q][lr] = v[q];
            lds_barrier();
            // x_g = inv(T_gg) v
            double4_t xa[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) xa[a] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < TC_BS / 4; ++q)
                xa[q & 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(
                    s_inv[16 * wave + lr][4 * q + lk], s_v[4 * q + lk][lr], xa[q & 3], 0, 0, 0);
            double x[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) x[q] = (xa[0][q] + xa[1][q]) + (xa[2][q] + xa[3][q]);
            lds_barrier();            // every read of s_xc and s_v is done
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int rr = 16 * wave + lk + 4 * q;
                const double xv = rr < rows ? x[q] : 0.0;
                s_xc[rr][lr] = xv;
                if (lr < R) {
                    if (rr < rows && lr < nrhs) Xout[(long)lr * ldx + r0 + rr] = xv;
                    uint64_t *p = xw + ((long)(bi * TC_BS + rr) * R + lr) * 2;
                    tc_st(p, tc_word_hi(epoch, xv));
                    tc_st(p + 1, tc_word_lo(epoch, xv));
                }
            }
        }
        if (bad) atomicOr(err, 2);
        return;
    }
    // ------------------------------------------------------------- helpers
    const int H = gridDim.x - 1;
    int spins = 0;
    bool gave_up = false;
    for (int g = 2 + (blockIdx.x - 1); g < nbk; g += H) {
        const int bi = blk(g), r0 = bi * TC_BS, rows = min(TC_BS, n - r0);
        const bool row_ok = lane < rows;
        const int row = r0 + (row_ok ? lane : 0);
        const int ndep = g - 1;                  // x_0 .. x_{g-2}
        double acc[R];
#pragma unroll
        for (int k = 0; k < R; ++k) acc[k] = 0.0;
        int par = 0;
        for (int d0 = 0; d0 < ndep; d0 += TC_GRP, par ^= 1) {
            const int cnt = min(TC_GRP, ndep - d0);
            double lv[TC_GRP][16];
#pragma unroll
            for (int b = 0; b < TC_GRP; ++b) {
                const int c0 = blk(d0 + b) * TC_BS + wave * 16;
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    lv[b][c] = (b < cnt && row_ok && c0 + c < n)
                                   ? (TR ? LU[(long)row * lda + c0 + c] : LU[(long)(c0 + c) * lda + row])
                                   : 0.0;
            }
            // the group's x values into LDS: every word requested, then
            // each checked against the epoch (a stale one is re-polled)
            uint64_t wh[TC_GRP][XPT], wl[TC_GRP][XPT];
#pragma unroll
            for (int b = 0; b < TC_GRP; ++b)
#pragma unroll
                for (int u = 0; u < XPT; ++u) {
                    const int idx = tid + 256 * u;
                    wh[b][u] = wl[b][u] = (uint64_t)epoch << 32;
                    if (b < cnt && idx < TC_BS * R) {
                        const uint64_t *p = tc_addr<R>(xw, blk(d0 + b), idx / R, idx % R);
                        wh[b][u] = tc_ld(p);
                        wl[b][u] = tc_ld(p + 1);
                    }
                }
#pragma unroll
            for (int b = 0; b < TC_GRP; ++b)
#pragma unroll
                for (int u = 0; u < XPT; ++u) {
                    const int idx = tid + 256 * u;
                    if (b < cnt && idx < TC_BS * R) {
                        if (!gave_up)
                            gave_up = !tc_settle(tc_addr<R>(xw, blk(d0 + b), idx / R, idx % R),
                                                 wh[b][u], wl[b][u], epoch, spins);
                        s_x[par][b][idx / R][idx % R] = tc_join(wh[b][u], wl[b][u]);
                    }
                }
            __syncthreads();
#pragma unroll
            for (int b = 0; b < TC_GRP; ++b)
                if (b < cnt) {
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const double l = lv[b][c];
#pragma unroll
                        for (int k = 0; k < R; ++k) acc[k] = fma(l, s_x[par][b][wave * 16 + c][k], acc[k]);
                    }
                }
        }
#pragma unroll
        for (int k = 0; k < R; ++k) s_red[wave][lane][k] = acc[k];
        __syncthreads();
        for (int idx = tid; idx < TC_BS * R; idx += 256) {
            const int rr = idx / R, k = idx % R;
            const double pv = s_red[0][rr][k] + s_red[1][rr][k] + s_red[2][rr][k] + s_red[3][rr][k];
            uint64_t *p = pw + ((long)(bi * TC_BS + rr) * R + k) * 2;
            tc_st(p, tc_word_hi(epoch, pv));
            tc_st(p + 1, tc_word_lo(epoch, pv));
        }
        __syncthreads();           // s_red and s_x are reused by the next block
    }
    if (gave_up) atomicOr(err, 2);
}

// words of the chain solve's two tagged arrays (x and partial sums)
size_t trs_chain_words(int n) { return 2 * trs_persist_words(n); }

// NoTrans: X = U^{-1} L^{-1} Bp (Bp already row-permuted), T n x nrhs scratch,
// inv trs_inv_doubles(n) scratch, w trs_chain_words(n) tagged words.  With
// lower = 0 Bp already holds L^{-1} P B: X = U^{-1} Bp (T unused).
hipError_t trs_chain_f64(hipStream_t s, int n, int nrhs, const double *LU, long lda,
                         const double *Bp, long ldbp, double *T, double *X, long ldx,
                         double *inv, uint64_t *w, unsigned *epoch, int *err, int cus,
                         int lower, int rowmajor, hipEvent_t inv_done)
{
    if (rowmajor && lower) return hipErrorInvalidValue;   // the U solve only
    const int nbk = (n + TC_BS - 1) / TC_BS;
    if (nbk <= 0 || nrhs <= 0) return hipSuccess;
    const size_t half = trs_persist_words(n);
    uint64_t *xw = w, *pw = w + half;
    auto next = [&]() -> unsigned {
        if (++*epoch == 0u) {
            hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((2 * half + 255) / 256)), dim3(256),
                               0, s, w, 2 * half);
            *epoch = 1u;
        }
        return *epoch;
    };
    // chain + helpers, all resident: at most one workgroup per CU
    const int grid = 1 + std::max(1, std::min(nbk - 2, cus - 1));
    if (lower) hipLaunchKernelGGL((tri_inv_kernel<0, 1, 1>), dim3(nbk), dim3(64), 0, s, LU, lda, n, inv);
    for (int k0 = 0; k0 < (lower ? nrhs : 0); k0 += TP_RHS) {
        const int kc = min(TP_RHS, nrhs - k0);
        ProfScope ps(tl_prof, s, PROF_TRS, 2.0 * (double)n * n * kc,
                     8.0 * ((double)n * n + 4.0 * n * kc));
        const double *bsrc = Bp + (long)k0 * ldbp;
        double *tdst = T + (long)k0 * n;
#define FLX_TC(FW_, R_, IN_, LDI_, OUT_, LDO_)                                                     \
        hipLaunchKernelGGL((trs_chain_kernel<FW_, R_, 0>), dim3(grid), dim3(256), 0, s, LU, lda, n, \
                           kc, inv, IN_, LDI_, OUT_, LDO_, xw, pw, next(), err)
#define FLX_TCR(R_, IN_, LDI_, OUT_, LDO_)                                                        \
        hipLaunchKernelGGL((trs_chain_kernel<0, R_, 1>), dim3(grid), dim3(256), 0, s, LU, lda, n, \
                           kc, inv, IN_, LDI_, OUT_, LDO_, xw, pw, next(), err)
        if (kc == 1) FLX_TC(1, 1, bsrc, ldbp, tdst, (long)n);
        else if (kc <= 4) FLX_TC(1, 4, bsrc, ldbp, tdst, (long)n);
        else FLX_TC(1, 16, bsrc, ldbp, tdst, (long)n);
    }
    if (rowmajor)
        hipLaunchKernelGGL((tri_inv_kernel<1, 0, 0>), dim3(nbk), dim3(64), 0, s, LU, lda, n, inv);
    else
        hipLaunchKernelGGL((tri_inv_kernel<0, 0, 0>), dim3(nbk), dim3(64), 0, s, LU, lda, n, inv);
    if (inv_done) {
        hipError_t e = hipEventRecord(inv_done, s);
        if (e != hipSuccess) return e;
    }
    for (int k0 = 0; k0 < nrhs; k0 += TP_RHS) {
        const int kc = min(TP_RHS, nrhs - k0);
        ProfScope ps(tl_prof, s, PROF_TRS, lower ? 0.0 : (double)n * n * kc,
                     lower ? 0.0 : 8.0 * ((double)n * n / 2.0 + 2.0 * n * kc));
        const double *tsrc = lower ? T + (long)k0 * n : rowmajor ? Bp + k0 : Bp + (long)k0 * ldbp;
        const long ldt = lower ? (long)n : ldbp;
        double *xdst = X + (long)k0 * ldx;
        if (rowmajor) {
            if (kc == 1) FLX_TCR(1, tsrc, ldt, xdst, ldx);
            else if (kc <= 4) FLX_TCR(4, tsrc, ldt, xdst, ldx);
            else FLX_TCR(16, tsrc, ldt, xdst, ldx);
        } else {
            if (kc == 1) FLX_TC(0, 1, tsrc, ldt, xdst, ldx);
            else if (kc <= 4) FLX_TC(0, 4, tsrc, ldt, xdst, ldx);
            else FLX_TC(0, 16, tsrc, ldt, xdst, ldx);
        }
#undef FLX_TC
#undef FLX_TCR
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// dlaswp on a column-major matrix: rows k1..k2 (1-based), incx = +1 forward
// / -1 backward (flens/lapack/la/laswp.tcc:54-93).  One thread per column.
// ---------------------------------------------------------------------------
__global__ void laswp_cm_kernel(double *__restrict__ A, long lda, int ncols, int k1, int k2,
                                const int *__restrict__ ipiv, int incx)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    double *col = A + (long)c * lda;
    if (incx > 0) {
        for (int i = k1 - 1; i < k2; ++i) {
            int q = ipiv[i] - 1;
            if (q != i) { double t = col[i]; col[i] = col[q]; col[q] = t; }
        }
    } else {
        for (int i = k2 - 1; i >= k1 - 1; --i) {
            int q = ipiv[i] - 1;
            if (q != i) { double t = col[i]; col[i] = col[q]; col[q] = t; }
        }
    }
}

hipError_t laswp_cm_f64(hipStream_t s, double *A, long lda, int ncols, int k1, int k2,
                        const int *ipiv, int incx)
{
    if (ncols <= 0 || k2 < k1) return hipSuccess;
    hipLaunchKernelGGL(laswp_cm_kernel, dim3((ncols + 63) / 64), dim3(64), 0, s, A, lda, ncols,
                       k1, k2, ipiv, incx);
    return hipGetLastError();
}

} // namespace flx
