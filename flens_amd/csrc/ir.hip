// ir.hip -- FP64 iterative refinement around a low-precision LU factor.
//
// Kernels for the refinement loop of the mixed-precision solver (mixed.hip
// computes the factor):
//   * FP64 column-major GEMV in two deterministic phases (row blocks x
//     column chunks, then a per-row sum): r = b - A x, |A| row sums for
//     ||A||_inf, row maxima for the equilibration;
//   * the preconditioner solve z = M^{-1} v with the FP32 factor,
//     M = D_r^{-1} P^T L U D_c^{-1}: gather + scale into FP32, forward and
//     backward substitution, scale back into FP64;
//   * the substitutions: one launch per triangle, one workgroup per 64-row
//     block, each consuming (with bounded waits) the blocks it depends on
//     as they are published, and solving its diagonal block by wave-level
//     substitution;
//   * vector kernels for GMRES (several dot products at once, w -= V h).
// The host-side loops (classic refinement with dsgesv's stopping rule,
// netlib dsgesv.f:199-296, and GMRES-IR) are in api.hip.
#include <cfloat>
#include <climits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {


// ---------------------------------------------------------------------------
// Row reductions of a column-major FP64 matrix, phase 1: block (bx, by)
// covers rows [bx*512, +512) (two per thread) and columns [by*RW, +RW).
//   mode 0: part = sum_j A(i,j) x_j      mode 1: part = sum_j |A(i,j)|
//   mode 2: part = max_j |A(i,j)| * (dc ? dc[j] : 1)
// ---------------------------------------------------------------------------
constexpr int RW = 256;

template <int MODE>
__global__ __launch_bounds__(256) void rowred_kernel(const double *__restrict__ A, long lda, int m,
                                                     int n, const double *__restrict__ x,
                                                     double *__restrict__ part)
{
    const int i = blockIdx.x * 512 + 2 * threadIdx.x;
    const int j0 = blockIdx.y * RW, j1 = min(n, j0 + RW);
    double a0 = 0.0, a1 = 0.0;
    if (i + 1 < m && (lda & 1) == 0 && ((uintptr_t)A & 15) == 0) {
        const double *p = A + (long)j0 * lda + i;
        for (int j = j0; j < j1; ++j, p += lda) {
            const double2 v = *reinterpret_cast<const double2 *>(p);
            if (MODE == 0) { const double xj = x[j]; a0 = fma(v.x, xj, a0); a1 = fma(v.y, xj, a1); }
            if (MODE == 1) { a0 += fabs(v.x); a1 += fabs(v.y); }
            if (MODE == 2) {
                const double s = x ? x[j] : 1.0;
                a0 = fmax(a0, fabs(v.x) * s); a1 = fmax(a1, fabs(v.y) * s);
            }
        }
    } else {
        for (int j = j0; j < j1; ++j) {
            const double v0 = (i < m) ? A[(long)j * lda + i] : 0.0;
            const double v1 = (i + 1 < m) ? A[(long)j * lda + i + 1] : 0.0;
            if (MODE == 0) { const double xj = x[j]; a0 = fma(v0, xj, a0); a1 = fma(v1, xj, a1); }
            if (MODE == 1) { a0 += fabs(v0); a1 += fabs(v1); }
            if (MODE == 2) {
                const double s = x ? x[j] : 1.0;
                a0 = fmax(a0, fabs(v0) * s); a1 = fmax(a1, fabs(v1) * s);
            }
        }
    }
    double *o = part + (long)blockIdx.y * m;
    if (i < m) o[i] = a0;
    if (i + 1 < m) o[i + 1] = a1;
}

// phase 2 for mode 0: r_i = b_i - sum_s part[s][i]; ||r||_inf -> nrm
__global__ void resid_fin_kernel(const double *__restrict__ part, int S, int m,
                                 const double *__restrict__ b, double *__restrict__ r,
                                 unsigned long long *nrm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < m) {
        double acc = 0.0;
        for (int s = 0; s < S; ++s) acc += part[(long)s * m + i];
        v = b[i] - acc;
        r[i] = v;
    }
    double a = fabs(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(nrm, a);
}

// phase 2 for mode 1: max_i sum_s part -> nrm (||A||_inf)
__global__ void rowsum_fin_kernel(const double *__restrict__ part, int S, int m,
                                  unsigned long long *nrm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc = 0.0;
    if (i < m)
        for (int s = 0; s < S; ++s) acc += part[(long)s * m + i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc = fmax(acc, __shfl_xor(acc, o, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(nrm, acc);
}


// phase 2 for mode 2: dr_i = 2^-e of the row maximum
__global__ void rowmax_fin_kernel(const double *__restrict__ part, int S, int m,
                                  double *__restrict__ dr)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double mx = 0.0;
    for (int s = 0; s < S; ++s) mx = fmax(mx, part[(long)s * m + i]);
    dr[i] = pow2_recip(mx);
}

// dc_j = 2^-e of max_i |A(i,j)| dr_i ; one workgroup per column
__global__ __launch_bounds__(256) void colmax_kernel(const double *__restrict__ A, long lda, int m,
                                                     const double *__restrict__ dr,
                                                     double *__restrict__ dc)
{
    const int j = blockIdx.x;
    double mx = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) mx = fmax(mx, fabs(A[(long)j * lda + i]) * dr[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    __shared__ double s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) dc[j] = pow2_recip(fmax(fmax(s[0], s[1]), fmax(s[2], s[3])));
}

static int rowred_chunks(int n) { return (n + RW - 1) / RW; }

size_t ir_part_doubles(int n) { return (size_t)rowred_chunks(n) * (size_t)n; }

hipError_t gemv_resid(hipStream_t s, const double *A, long lda, int n, const double *x,
                      const double *b, double *r, double *part, unsigned long long *nrm)
{
    if (n <= 0) return hipSuccess;
    const int S = rowred_chunks(n);
    ProfScope ps(tl_prof, s, PROF_OTHER, 2.0 * n * n, 8.0 * n * (double)n);
    hipLaunchKernelGGL(rowred_kernel<0>, dim3((n + 511) / 512, S), dim3(256), 0, s, A, lda, n, n,
                       x, part);
    hipLaunchKernelGGL(resid_fin_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const double *)part, S, n, b, r, nrm);
    return hipGetLastError();
}

hipError_t norm_inf(hipStream_t s, const double *A, long lda, int n, double *part,
                    unsigned long long *nrm)
{
    if (n <= 0) return hipSuccess;
    const int S = rowred_chunks(n);
    hipLaunchKernelGGL(rowred_kernel<1>, dim3((n + 511) / 512, S), dim3(256), 0, s, A, lda, n, n,
                       (const double *)nullptr, part);
    hipLaunchKernelGGL(rowsum_fin_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const double *)part, S, n, nrm);
    return hipGetLastError();
}

hipError_t equilibrate(hipStream_t s, const double *A, long lda, int n, double *part,
                       double *dr, double *dc)
{
    if (n <= 0) return hipSuccess;
    const int S = rowred_chunks(n);
    hipLaunchKernelGGL(rowred_kernel<2>, dim3((n + 511) / 512, S), dim3(256), 0, s, A, lda, n, n,
                       (const double *)nullptr, part);
    hipLaunchKernelGGL(rowmax_fin_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const double *)part, S, n, dr);
    hipLaunchKernelGGL(colmax_kernel, dim3(n), dim3(256), 0, s, A, lda, n, (const double *)dr, dc);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// y <- T^{-1} y for T the unit-lower (UPPER = 0) or upper triangle of the
// row-major FP32 factor W.  One launch per triangle; workgroup g owns the
// 64-row block bi = g (lower) or nbk-1-g (upper), so every block it waits
// for belongs to a workgroup dispatched before it.
//   * Finished blocks are published as 64-bit words (epoch << 32 | x bits)
//     in xt: a word is its own ready flag, so one agent-scope load both
//     tests readiness and delivers the value.  Waits are bounded (err |= 2).
//   * Thread (rr, q) = (tid / 4, tid % 4) accumulates row rr of its block
//     over columns q*16..+15 of every finished block: the older blocks in
//     groups of 8 (wave 0 waits until all 64 words of each carry this
//     launch's epoch, then the group goes through LDS), the block finished
//     last on its own -- its factor entries are loaded at kernel start and
//     every wave polls its words itself, so no barrier sits between its
//     arrival and the product.
//   * The diagonal block is applied as its inverse (computed once per
//     factor by trsv_inv_kernel, in FP64 from the FP32 entries): a 64 x 64
//     product whose rows are loaded at kernel start, one barrier after the
//     right-hand side is known, instead of a 64-step substitution chain.
// ---------------------------------------------------------------------------
constexpr int TV = 64;
constexpr int TV_FETCH = 8;
constexpr int TV_SPIN_LIMIT = 1 << 22;

// Published solution words: a float travels as one tagged word, a double
// as two (high and low halves), each (epoch << 32 | 32 payload bits).
template <typename TY> struct TvWords;
template <> struct TvWords<float> {
    static constexpr int N = 1;
    FLX_DEV void put(uint64_t *p, unsigned e, float v)
    {
        __hip_atomic_store(p, ((uint64_t)e << 32) | __float_as_uint(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    FLX_DEV bool get(const uint64_t *p, unsigned e, float &v)
    {
        const uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = __uint_as_float((unsigned)(w & 0xffffffffull));
        return (unsigned)(w >> 32) == e;
    }
};
template <> struct TvWords<double> {
    static constexpr int N = 2;
    FLX_DEV void put(uint64_t *p, unsigned e, double v)
    {
        const uint64_t b = (uint64_t)__double_as_longlong(v);
        __hip_atomic_store(p, ((uint64_t)e << 32) | (b >> 32), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(p + 1, ((uint64_t)e << 32) | (b & 0xffffffffull), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    FLX_DEV bool get(const uint64_t *p, unsigned e, double &v)
    {
        const uint64_t hi = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t lo = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
        return (unsigned)(hi >> 32) == e && (unsigned)(lo >> 32) == e;
    }
};

// Inverses of the 64 x 64 diagonal blocks of W's two triangles, FP64,
// row-major: dinv[(t * nbk + bi) * 4096 + 64 i + j], t = 0 for the unit
// lower triangle, 1 for the upper.  Rows and columns beyond n are the
// identity's.  One wave per block; lane j forms column j of the inverse by
// substitution against the block held in LDS (eight terms of each sum
// requested at once).
__global__ __launch_bounds__(64) void trsv_inv_kernel(const float *__restrict__ W, long ldw,
                                                      int n, double *__restrict__ dinv)
{
    __shared__ double s_t[TV][TV + 1];
    __shared__ double s_x[TV][TV + 1];
    const int nbk = (n + TV - 1) / TV;
    const int bi = blockIdx.x, upper = blockIdx.y;
    const int r0 = bi * TV, rows = min(TV, n - r0);
    const int j = threadIdx.x;
    {
        float v[TV];
#pragma unroll
        for (int i = 0; i < TV; ++i)
            v[i] = (i < rows && j < rows) ? W[(long)(r0 + i) * ldw + r0 + j]
                                          : (i == j ? 1.0f : 0.0f);
#pragma unroll
        for (int i = 0; i < TV; ++i) s_t[i][j] = (double)v[i];
    }
    __syncthreads();
    if (!upper) {
        for (int i = 0; i < TV; ++i) {
            double a = (i == j) ? 1.0 : 0.0;
#pragma unroll 8
            for (int k = j; k < i; ++k) a = fma(-s_t[i][k], s_x[k][j], a);
            s_x[i][j] = (i < j) ? 0.0 : a;
        }
    } else {
        for (int i = TV - 1; i >= 0; --i) {
            double a = (i == j) ? 1.0 : 0.0;
#pragma unroll 8
            for (int k = i + 1; k <= j; ++k) a = fma(-s_t[i][k], s_x[k][j], a);
            s_x[i][j] = (i > j) ? 0.0 : a / s_t[i][i];
        }
    }
    double *out = dinv + ((size_t)upper * nbk + bi) * TV * TV;
#pragma unroll 8
    for (int i = 0; i < TV; ++i) out[i * TV + j] = s_x[i][j];
}

// TY = float: FP32 solve (dsgesv's single-precision solve, the inverse
// blocks rounded to FP32); TY = double: the FP32 factor applied in FP64
// arithmetic (GMRES-IR).
template <typename TY, int UPPER>
__global__ __launch_bounds__(256) void trsv_rm_kernel(const float *__restrict__ W, long ldw, int n,
                                                      const double *__restrict__ dinv,
                                                      TY *__restrict__ y,
                                                      uint64_t *__restrict__ xt,
                                                      unsigned epoch, int *__restrict__ err)
{
    typedef TvWords<TY> TW;
    __shared__ TY s_x[2][TV_FETCH][TV];
    __shared__ TY s_rhs[TV];
    const int nbk = (n + TV - 1) / TV;
    const int g = blockIdx.x;
    const int bi = UPPER ? nbk - 1 - g : g;
    const int r0 = bi * TV, rows = min(TV, n - r0);
    const int tid = threadIdx.x, lane = tid & 63;
    const int rr = tid >> 2, q = tid & 3;
    const int ndep = UPPER ? nbk - 1 - bi : bi;
    const bool row_ok = rr < rows;
    const float *wrow = W + (long)(r0 + (row_ok ? rr : 0)) * ldw;
    const TY rhs0 = row_ok ? y[r0 + rr] : TY(0);
    // ahead of every wait: the inverse's row rr (columns q*16..+15) and the
    // factor entries of the block finished last
    TY iv[16];
    {
        const double *ib = dinv + ((size_t)UPPER * nbk + bi) * TV * TV + rr * TV + q * 16;
#pragma unroll
        for (int c = 0; c < 16; ++c) iv[c] = (TY)ib[c];
    }
    const int jl = UPPER ? bi + 1 : bi - 1;   // the block finished last (ndep >= 1)
    float wl[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        const int col = jl * TV + q * 16 + c;
        wl[c] = (ndep >= 1 && row_ok && col < n) ? wrow[col] : 0.0f;
    }
    TY acc = TY(0);
    int par = 0, spins = 0;
    const int nold = ndep - 1;   // blocks finished before the last one
    for (int d0 = 0; d0 < nold; d0 += TV_FETCH, par ^= 1) {
        const int cnt = min(TV_FETCH, nold - d0);
        float wv[TV_FETCH][16];
#pragma unroll
        for (int b = 0; b < TV_FETCH; ++b) {
            const int d = d0 + b;
            const int jb = UPPER ? nbk - 1 - d : d;
            const int c0 = jb * TV + q * 16;
#pragma unroll
            for (int c = 0; c < 16; ++c)
                wv[b][c] = (b < cnt && row_ok && c0 + c < n) ? wrow[c0 + c] : 0.0f;
        }
        if (tid < 64) {
            for (;;) {
                TY v[TV_FETCH];
                bool all = true;
#pragma unroll
                for (int b = 0; b < TV_FETCH; ++b) {
                    const int d = d0 + b;
                    const int jb = UPPER ? nbk - 1 - d : d;
                    bool ok = true;
                    v[b] = TY(0);
                    if (b < cnt) ok = TW::get(&xt[((long)jb * TV + lane) * TW::N], epoch, v[b]);
                    all = all && (__ballot(ok) == ~0ull);
                }
                if (all || spins > TV_SPIN_LIMIT) {
                    if (!all && lane == 0) atomicOr(err, 2);   // gave up: go on with what is there
#pragma unroll
                    for (int b = 0; b < TV_FETCH; ++b) s_x[par][b][lane] = v[b];
                    break;
                }
                ++spins;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < TV_FETCH; ++b)
#pragma unroll
            for (int c = 0; c < 16; ++c) acc = acc + (TY)wv[b][c] * s_x[par][b][q * 16 + c];
    }
    if (ndep >= 1) {
        // every wave polls the last block's 64 words (lane l: row l)
        TY xl = TY(0);
        for (;;) {
            const bool ok = TW::get(&xt[((long)jl * TV + lane) * TW::N], epoch, xl);
            const bool all = __ballot(ok) == ~0ull;
            if (all || spins > TV_SPIN_LIMIT) {
                if (!all && lane == 0) atomicOr(err, 2);
                break;
            }
            ++spins;
            __builtin_amdgcn_s_sleep(1);
        }
#pragma unroll
        for (int c = 0; c < 16; ++c) acc = acc + (TY)wl[c] * __shfl(xl, q * 16 + c, 64);
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (q == 0) s_rhs[rr] = row_ok ? rhs0 - acc : TY(0);
    __syncthreads();
    TY x = TY(0);
#pragma unroll
    for (int c = 0; c < 16; ++c) x = x + iv[c] * s_rhs[q * 16 + c];
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    if (q == 0) {
        if (row_ok) y[r0 + rr] = x;
        TW::put(&xt[((long)bi * TV + rr) * TW::N], epoch, row_ok ? x : TY(0));
    }
}

size_t trsv_inv_doubles(int n) { return 2 * (size_t)((n + TV - 1) / TV) * TV * TV; }

hipError_t trsv_diag_inv(hipStream_t s, const float *W, long ldw, int n, double *dinv)
{
    const int nbk = (n + TV - 1) / TV;
    if (nbk <= 0) return hipSuccess;
    hipLaunchKernelGGL(trsv_inv_kernel, dim3(nbk, 2), dim3(64), 0, s, W, ldw, n, dinv);
    return hipGetLastError();
}

template <typename TY>
hipError_t trsv_rm(hipStream_t s, int upper, const float *W, long ldw, int n, const double *dinv,
                   TY *y, uint64_t *xt, unsigned epoch, int *err)
{
    const int nbk = (n + TV - 1) / TV;
    if (nbk <= 0) return hipSuccess;
    ProfScope ps(tl_prof, s, PROF_TRS, (double)n * n, 2.0 * n * (double)n);
    if (upper)
        hipLaunchKernelGGL((trsv_rm_kernel<TY, 1>), dim3(nbk), dim3(256), 0, s, W, ldw, n, dinv,
                           y, xt, epoch, err);
    else
        hipLaunchKernelGGL((trsv_rm_kernel<TY, 0>), dim3(nbk), dim3(256), 0, s, W, ldw, n, dinv,
                           y, xt, epoch, err);
    return hipGetLastError();
}
template hipError_t trsv_rm<float>(hipStream_t, int, const float *, long, int, const double *,
                                   float *, uint64_t *, unsigned, int *);
template hipError_t trsv_rm<double>(hipStream_t, int, const float *, long, int, const double *,
                                    double *, uint64_t *, unsigned, int *);

size_t trsv_words(int n) { return 2 * (size_t)((n + TV - 1) / TV) * TV; }   // room for doubles

// ---------------------------------------------------------------------------
// Preconditioner glue and vector kernels (FP64 vectors of length n)
// ---------------------------------------------------------------------------
FLX_DEV double nrm_scale(const unsigned long long *nrm)
{
    return pow2_recip(__longlong_as_double((long long)*nrm));
}

// y32[i] = v[perm[i]] * dr[perm[i]] * s   with s = 2^-e of ||v||_inf
__global__ void prec_in_kernel(const double *__restrict__ v, const int *__restrict__ perm,
                               const double *__restrict__ dr, const unsigned long long *nrm,
                               float *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    double t = v[p] * nrm_scale(nrm);
    if (dr) t *= dr[p];
    y[i] = (float)t;
}

// out[j] = (add ? out[j] : 0) + y32[j] * dc[j] / s ; ||out||_inf -> onrm
__global__ void prec_out_kernel(const float *__restrict__ y, const double *__restrict__ dc,
                                const unsigned long long *nrm, double *__restrict__ out, int add,
                                int n, unsigned long long *onrm)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (j < n) {
        double t = (double)y[j] / nrm_scale(nrm);
        if (dc) t *= dc[j];
        v = add ? out[j] + t : t;
        out[j] = v;
    }
    if (onrm) {
        double a = fabs(v);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
        if ((threadIdx.x & 63) == 0) atomic_max_nonneg(onrm, a);
    }
}

__global__ void vmax_kernel(const double *__restrict__ v, int n, unsigned long long *nrm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = (i < n) ? fabs(v[i]) : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(nrm, a);
}

__global__ void zero_u64_slots_kernel(unsigned long long *p, int k)
{
    if ((int)threadIdx.x < k) p[threadIdx.x] = 0ull;
}

hipError_t zero_slots(hipStream_t s, unsigned long long *p, int k)
{
    hipLaunchKernelGGL(zero_u64_slots_kernel, dim3(1), dim3(64), 0, s, p, k);
    return hipGetLastError();
}

hipError_t vmax(hipStream_t s, const double *v, int n, unsigned long long *nrm)
{
    hipLaunchKernelGGL(vmax_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v, n, nrm);
    return hipGetLastError();
}

__global__ void zero_u64_n_kernel(uint64_t *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// a fresh tag per substitution launch; the words are cleared on wrap
static unsigned next_epoch(hipStream_t s, const PrecFactor &f)
{
    if (++*f.epoch == 0u) {
        const int nw = (int)trsv_words(f.n);
        hipLaunchKernelGGL(zero_u64_n_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, f.xt, nw);
        *f.epoch = 1u;
    }
    return *f.epoch;
}

// z = M^{-1} v   (vnrm: a zeroed slot that receives ||v||_inf first)
hipError_t prec_apply(hipStream_t s, const PrecFactor &f, const double *v, double *z, int add,
                      unsigned long long *vnrm, unsigned long long *znrm)
{
    const int n = f.n;
    FLX_CHECK(vmax(s, v, n, vnrm));
    if (f.rbt)
        FLX_CHECK(rbt_in(s, v, f.dr, vnrm, f.rb, f.y32, nullptr, n));
    else
        hipLaunchKernelGGL(prec_in_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v,
                           (const int *)f.perm, f.dr, (const unsigned long long *)vnrm, f.y32, n);
    FLX_CHECK(trsv_rm<float>(s, 0, f.W, f.ldw, n, f.dinv, f.y32, f.xt, next_epoch(s, f), f.err));
    FLX_CHECK(trsv_rm<float>(s, 1, f.W, f.ldw, n, f.dinv, f.y32, f.xt, next_epoch(s, f), f.err));
    if (f.rbt)
        return rbt_out(s, f.y32, nullptr, f.dc, vnrm, f.rb, z, add, n, znrm);
    hipLaunchKernelGGL(prec_out_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const float *)f.y32, f.dc, (const unsigned long long *)vnrm, z, add, n,
                       znrm);
    return hipGetLastError();
}

// FP64 application of the FP32 factor: z = M^{-1} v without rounding the
// operand or the solution to FP32 (the factor entries are exact in FP64)
__global__ void prec_in64_kernel(const double *__restrict__ v, const int *__restrict__ perm,
                                 const double *__restrict__ dr, double *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = perm[i];
    y[i] = dr ? v[p] * dr[p] : v[p];
}

__global__ void prec_out64_kernel(const double *__restrict__ y, const double *__restrict__ dc,
                                  double *__restrict__ out, int add, int n)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double t = dc ? y[j] * dc[j] : y[j];
    out[j] = add ? out[j] + t : t;
}

hipError_t prec_apply64(hipStream_t s, const PrecFactor &f, const double *v, double *z, int add)
{
    const int n = f.n;
    if (f.rbt)
        FLX_CHECK(rbt_in(s, v, f.dr, nullptr, f.rb, nullptr, f.y64, n));
    else
        hipLaunchKernelGGL(prec_in64_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v,
                           (const int *)f.perm, f.dr, f.y64, n);
    FLX_CHECK(trsv_rm<double>(s, 0, f.W, f.ldw, n, f.dinv, f.y64, f.xt, next_epoch(s, f), f.err));
    FLX_CHECK(trsv_rm<double>(s, 1, f.W, f.ldw, n, f.dinv, f.y64, f.xt, next_epoch(s, f), f.err));
    if (f.rbt)
        return rbt_out(s, nullptr, f.y64, f.dc, nullptr, f.rb, z, add, n, nullptr);
    hipLaunchKernelGGL(prec_out64_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const double *)f.y64, f.dc, z, add, n);
    return hipGetLastError();
}

// ---- GMRES helpers -----------------------------------------------------------
// part[blk][c] = sum over the block's rows of V[:, c] * w   (c < k)
constexpr int MD_ROWS = 2048;

__global__ __launch_bounds__(256) void mdot_part_kernel(const double *__restrict__ V, long ldv, int n,
                                                        int k, const double *__restrict__ w,
                                                        double *__restrict__ part)
{
    __shared__ double s[4];
    const int i0 = blockIdx.x * MD_ROWS;
    const int i1 = min(n, i0 + MD_ROWS);
    for (int c = 0; c < k; ++c) {
        double a = 0.0;
        for (int i = i0 + threadIdx.x; i < i1; i += 256) a = fma(V[(long)c * ldv + i], w[i], a);
        a = wave_sum(a);
        if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) part[(long)blockIdx.x * k + c] = (s[0] + s[1]) + (s[2] + s[3]);
        __syncthreads();
    }
}

__global__ void mdot_fin_kernel(const double *__restrict__ part, int nblk, int k,
                                double *__restrict__ out)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double a = 0.0;
    for (int b = 0; b < nblk; ++b) a += part[(long)b * k + c];
    out[c] = a;
}

// w -= V[:, 0..k) h ; then optionally w *= scale (scale read from *sc if given)
__global__ void maxpy_kernel(const double *__restrict__ V, long ldv, int n, int k,
                             const double *__restrict__ h, double *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = w[i];
    for (int c = 0; c < k; ++c) a = fma(-V[(long)c * ldv + i], h[c], a);
    w[i] = a;
}

__global__ void vscale_copy_kernel(const double *__restrict__ src, double alpha,
                                   double *__restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = alpha * src[i];
}

__global__ void fill_const_kernel(double *v, int n, double c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = c;
}
__global__ void sign_of_kernel(const double *__restrict__ x, double *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] >= 0.0 ? 1.0 : -1.0;
}
hipError_t fill_const(hipStream_t s, double *v, int n, double c)
{
    hipLaunchKernelGGL(fill_const_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v, n, c);
    return hipGetLastError();
}
hipError_t sign_of(hipStream_t s, const double *x, double *y, int n)
{
    hipLaunchKernelGGL(sign_of_kernel, dim3((n + 255) / 256), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

size_t mdot_part_doubles(int n, int kmax) { return (size_t)((n + MD_ROWS - 1) / MD_ROWS) * kmax; }

hipError_t mdot(hipStream_t s, const double *V, long ldv, int n, int k, const double *w,
                double *part, double *out)
{
    if (k <= 0) return hipSuccess;
    const int nblk = (n + MD_ROWS - 1) / MD_ROWS;
    hipLaunchKernelGGL(mdot_part_kernel, dim3(nblk), dim3(256), 0, s, V, ldv, n, k, w, part);
    hipLaunchKernelGGL(mdot_fin_kernel, dim3(1), dim3(256), 0, s, (const double *)part, nblk, k, out);
    return hipGetLastError();
}

hipError_t maxpy(hipStream_t s, const double *V, long ldv, int n, int k, const double *h, double *w)
{
    if (k <= 0) return hipSuccess;
    hipLaunchKernelGGL(maxpy_kernel, dim3((n + 255) / 256), dim3(256), 0, s, V, ldv, n, k, h, w);
    return hipGetLastError();
}

hipError_t vscale_copy(hipStream_t s, const double *src, double alpha, double *dst, int n)
{
    hipLaunchKernelGGL(vscale_copy_kernel, dim3((n + 255) / 256), dim3(256), 0, s, src, alpha, dst, n);
    return hipGetLastError();
}

} // namespace flx
