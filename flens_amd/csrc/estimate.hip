// estimate.hip -- norms, equilibration, condition estimate, refinement with
// error bounds and the expert driver, on the device.
//
// Replaces the FLENS-LAPACK generic implementations around the LU:
//   lan (GeMatrix)  flens/lapack/la/lan.tcc:55-130
//   equ             flens/lapack/ge/equ.tcc:62-194
//   laq             flens/lapack/la/laq.tcc:57-118
//   con (+ lacn2)   flens/lapack/ge/cond.tcc:62-165, flens/lapack/la/lacn2.tcc:55-185
//   rfs             flens/lapack/ge/rfs.tcc:63-261
//   svx             flens/lapack/ge/svx.tcc:60-262
// and the Fortran seams dlange_, dgeequ_, dgecon_, dgerfs_, dgesvx_
// (cxxlapack/netlib/interface/lapack.in.h:6143-6166, 6535-6553, 6676-6699,
// 8145-8151).
//
// Matrix passes are one kernel each.  The 1- and infinity-norm sums and the
// equilibration maxima visit every row (column) in the reference's order,
// one workgroup per 64-row (64-column) slab staged through LDS, so those
// results are bit-identical to the reference; the Frobenius norm and the
// residuals use parallel partial sums (rounding differs within FP64
// tolerance).  lacn2's reverse-communication state machine runs on the host
// exactly as the reference's; its vector steps (sums, sign vectors, argmax)
// and every solve run on the device, only scalars cross to the host.
// The triangular solves of con are the LU's own solves (latrs without its
// scaling branch): when a solution approaches the overflow threshold at
// which latrs would rescale, the estimate is reported as 0.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>
#include "flx_common.h"
#include "context.h"

namespace flx {

constexpr double EPS64 = DBL_EPSILON * 0.5;   // lamch Eps
constexpr double PREC64 = DBL_EPSILON;        // lamch Precision
constexpr double SAFEMIN = DBL_MIN;           // lamch SafeMin

// slots of AuxBufs::slot (unsigned long long each)
enum { SL_MAX = 0, SL_MIN, SL_ZERO, SL_FLAG, SL_SUM, SL_IDX, SL_CNT, SL_N };

FLX_DEV void amax_nonneg(unsigned long long *slot, double v)
{
    // v >= 0 (or NaN, never passed): the bit pattern orders like the value
    if (v > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}
FLX_DEV double std_max(double a, double b) { return (a < b) ? b : a; }   // std::max
FLX_DEV double std_min(double a, double b) { return (b < a) ? b : a; }   // std::min

// ---------------------------------------------------------------------------
// Slab reductions in the reference's order.  Workgroup b owns lines
// [32b, 32b+32) -- rows (BYROW) or columns -- and walks the other dimension
// in 64-wide tiles staged through LDS; thread t < 32 folds its line's tile
// entries in ascending order.  Loads run SLAB_AHEAD tiles ahead of the fold
// (registers), so several tiles per workgroup are in flight.
// MODE 0: sum |a|; 1: std::max(acc, |a|); 2: std::max(acc, |a| * s[row]).
// ---------------------------------------------------------------------------
constexpr int SLAB = 32;          // lines per workgroup
constexpr int SLAB_K = 64;        // tile length along the other dimension
constexpr int SLAB_AHEAD = 4;     // tiles in flight
constexpr int SLAB_PER = SLAB * SLAB_K / 256;   // elements per thread per tile (8)

template <bool BYROW, int MODE>
__global__ __launch_bounds__(256) void slab_reduce_kernel(const double *__restrict__ A, long lda,
                                                          int m, int n,
                                                          const double *__restrict__ s,
                                                          double *__restrict__ out)
{
    __shared__ double tile[SLAB][SLAB_K + 1];   // [line][k]
    const int t = threadIdx.x;
    const int l0 = blockIdx.x * SLAB;
    const int lines = BYROW ? m : n, len = BYROW ? n : m;
    const int ntiles = (len + SLAB_K - 1) / SLAB_K;
    double acc = 0.0;
    // element e of a tile: rows -> (line e % 32, k e / 32); columns -> (k e % 64, line e / 64),
    // both contiguous in memory along the fast index
    auto fetch = [&](int tk, double (&v)[SLAB_PER]) {
        const int k0 = tk * SLAB_K;
#pragma unroll
        for (int u = 0; u < SLAB_PER; ++u) {
            const int e = t + u * 256;
            int q, k;
            if (BYROW) { q = e % SLAB; k = e / SLAB; }
            else { k = e % SLAB_K; q = e / SLAB_K; }
            const int i = BYROW ? l0 + q : k0 + k, j = BYROW ? k0 + k : l0 + q;
            double x = 0.0;
            if (tk < ntiles && i < m && j < n) {
                x = fabs(A[(long)j * lda + i]);
                if (MODE == 2) x = x * s[i];
            }
            v[u] = x;
        }
    };
    double buf[SLAB_AHEAD][SLAB_PER];
#pragma unroll
    for (int a = 0; a < SLAB_AHEAD; ++a) fetch(a, buf[a]);
    for (int tk0 = 0; tk0 < ntiles; tk0 += SLAB_AHEAD) {
#pragma unroll
        for (int a = 0; a < SLAB_AHEAD; ++a) {
            const int tk = tk0 + a;
#pragma unroll
            for (int u = 0; u < SLAB_PER; ++u) {
                const int e = t + u * 256;
                int q, k;
                if (BYROW) { q = e % SLAB; k = e / SLAB; }
                else { k = e % SLAB_K; q = e / SLAB_K; }
                tile[q][k] = buf[a][u];
            }
            fetch(tk + SLAB_AHEAD, buf[a]);   // the next group's tile, in flight during the fold
            __syncthreads();
            if (t < SLAB && tk < ntiles) {
                const int kk = min(SLAB_K, len - tk * SLAB_K);
                for (int k = 0; k < kk; ++k) {
                    const double v = tile[t][k];
                    if (MODE == 0) acc += v;
                    else acc = std_max(acc, v);
                }
            }
            __syncthreads();
        }
    }
    if (t < SLAB && l0 + t < lines) out[l0 + t] = acc;
}

// column-strided matrix pass: workgroup g takes columns g, g + grid, ...;
// the 256 threads stride down each column (coalesced)
#define FOR_EACH_ELEM(m, n)                                                   \
    for (int j = blockIdx.x; j < (n); j += gridDim.x)                         \
        for (int i = threadIdx.x; i < (m); i += blockDim.x)

// max (std::max fold from 0), min (std::min fold from `big`) and the first
// index holding exactly zero, over v[0, L)
__global__ void vec_stats_kernel(const double *__restrict__ v, int L, double big,
                                 unsigned long long *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0, mn = big;
    if (i < L) {
        const double x = v[i];
        mx = std_max(0.0, x);
        mn = std_min(big, x);
        if (x == 0.0) atomicMin(&slot[SL_ZERO], (unsigned long long)i);
        if (!(x >= 0.0)) atomicOr(&slot[SL_FLAG], 1ull);   // negative or NaN
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o, 64));
        mn = fmin(mn, __shfl_xor(mn, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        amax_nonneg(&slot[SL_MAX], mx);
        if (mn >= 0.0) atomicMin(&slot[SL_MIN], (unsigned long long)__double_as_longlong(mn));
    }
}

// max |a| over the matrix, NaN skipped (std::max fold from 0)
__global__ void mat_absmax_kernel(const double *__restrict__ A, long lda, int m, int n,
                                  unsigned long long *slot)
{
    double mx = 0.0;
    FOR_EACH_ELEM(m, n) mx = std_max(mx, fabs(A[(long)j * lda + i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) amax_nonneg(&slot[SL_MAX], mx);
}

// max |a_ij| over the upper triangle (i <= j) of the leading k x k block
__global__ void upper_absmax_kernel(const double *__restrict__ A, long lda, int k,
                                    unsigned long long *slot)
{
    double mx = 0.0;
    for (int j = blockIdx.x; j < k; j += gridDim.x)
        for (int i = threadIdx.x; i <= j; i += blockDim.x)
            mx = std_max(mx, fabs(A[(long)j * lda + i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) amax_nonneg(&slot[SL_MAX], mx);
}

// Frobenius pass 1: max finite |a|, number of infinities, any NaN
__global__ void fro_scan_kernel(const double *__restrict__ A, long lda, int m, int n,
                                unsigned long long *slot)
{
    double mx = 0.0;
    unsigned ninf = 0, nan = 0;
    FOR_EACH_ELEM(m, n) {
        const double a = fabs(A[(long)j * lda + i]);
        if (a != a) nan = 1;
        else if (a > DBL_MAX) ++ninf;
        else mx = fmax(mx, a);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o, 64));
        ninf += __shfl_xor(ninf, o, 64);
        nan |= __shfl_xor(nan, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        amax_nonneg(&slot[SL_MAX], mx);
        if (ninf) atomicAdd(&slot[SL_CNT], (unsigned long long)ninf);
        if (nan) atomicOr(&slot[SL_FLAG], 1ull);
    }
}

// Frobenius pass 2: per-block partial sums of (|a| / scale)^2
__global__ __launch_bounds__(256) void fro_sum_kernel(const double *__restrict__ A, long lda,
                                                      int m, int n, double scale,
                                                      double *__restrict__ part)
{
    __shared__ double red[4];
    double s = 0.0;
    FOR_EACH_ELEM(m, n) {
        const double q = fabs(A[(long)j * lda + i]) / scale;
        s += q * q;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// fixed-order sum of part[0, P) into the SL_SUM slot (as a double)
__global__ void fixed_sum_kernel(const double *__restrict__ part, int P, unsigned long long *slot)
{
    __shared__ double red[4];
    double s = 0.0;
    for (int i = threadIdx.x; i < P; i += 256) s += part[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        slot[SL_SUM] = (unsigned long long)__double_as_longlong((red[0] + red[1]) + (red[2] + red[3]));
}

// per-block partial sums of |x| (lacn2's asum), then fixed_sum_kernel
__global__ __launch_bounds__(256) void asum_part_kernel(const double *__restrict__ x, int n,
                                                        double *__restrict__ part)
{
    __shared__ double red[4];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += fabs(x[i]);
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// r_i <- 1 / min(max(r_i, small), big)
__global__ void inv_clamp_kernel(double *v, int L, double small, double big)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L) v[i] = 1.0 / std_min(std_max(v[i], small), big);
}

// laq: A(i,j) *= c_j (mode 'C'), r_i ('R'), (c_j r_i) ('B')
__global__ void laqge_kernel(double *__restrict__ A, long lda, int m, int n,
                             const double *__restrict__ r, const double *__restrict__ c, int mode)
{
    FOR_EACH_ELEM(m, n) {
        double &a = A[(long)j * lda + i];
        if (mode == 'C') a *= c[j];
        else if (mode == 'R') a *= r[i];
        else a *= c[j] * r[i];
    }
}

// B(i,j) *= s_i over an n x nrhs block
__global__ void scale_rows_kernel(double *__restrict__ B, long ldb, int n, int nrhs,
                                  const double *__restrict__ s)
{
    const long total = (long)n * nrhs;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        B[(long)j * ldb + i] *= s[i];
    }
}

// ---------------------------------------------------------------------------
// rfs residual: r = b - op(A) x, w = |b| + |op(A)| |x|, and the
// componentwise backward error max_i |r_i| / w_i (safe1-guarded).
// NoTrans: rows in 512-row blocks x 256-column chunks -> partials, then a
// fixed-order finish.  Trans: one workgroup per column.
// ---------------------------------------------------------------------------
constexpr int RR = 256;   // column chunk of the NoTrans residual

__global__ __launch_bounds__(256) void resid_nt_kernel(const double *__restrict__ A, long lda,
                                                       int n, const double *__restrict__ x,
                                                       double *__restrict__ pr,
                                                       double *__restrict__ pw)
{
    const int i = blockIdx.x * 512 + 2 * threadIdx.x;
    const int j0 = blockIdx.y * RR, j1 = min(n, j0 + RR);
    double r0 = 0.0, r1 = 0.0, w0 = 0.0, w1 = 0.0;
    for (int j = j0; j < j1; ++j) {
        const double xj = x[j], ax = fabs(xj);
        const double v0 = (i < n) ? A[(long)j * lda + i] : 0.0;
        const double v1 = (i + 1 < n) ? A[(long)j * lda + i + 1] : 0.0;
        r0 = fma(v0, xj, r0); r1 = fma(v1, xj, r1);
        w0 = fma(fabs(v0), ax, w0); w1 = fma(fabs(v1), ax, w1);
    }
    double *o = pr + (long)blockIdx.y * n, *q = pw + (long)blockIdx.y * n;
    if (i < n) { o[i] = r0; q[i] = w0; }
    if (i + 1 < n) { o[i + 1] = r1; q[i + 1] = w1; }
}

FLX_DEV double berr_term(double r, double w, double safe1, double safe2)
{
    return (w > safe2) ? fabs(r) / w : (fabs(r) + safe1) / (w + safe1);
}

__global__ void resid_nt_fin_kernel(const double *__restrict__ pr, const double *__restrict__ pw,
                                    int S, int n, const double *__restrict__ b,
                                    double *__restrict__ r, double *__restrict__ w, double safe1,
                                    double safe2, unsigned long long *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0.0;
    if (i < n) {
        double ar = 0.0, aw = 0.0;
        for (int s = 0; s < S; ++s) { ar += pr[(long)s * n + i]; aw += pw[(long)s * n + i]; }
        const double ri = b[i] - ar, wi = fabs(b[i]) + aw;
        r[i] = ri;
        w[i] = wi;
        q = berr_term(ri, wi, safe1, safe2);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) q = fmax(q, __shfl_xor(q, o, 64));
    if ((threadIdx.x & 63) == 0) amax_nonneg(&slot[SL_MAX], q);
}

__global__ __launch_bounds__(256) void resid_t_kernel(const double *__restrict__ A, long lda,
                                                      int n, const double *__restrict__ x,
                                                      const double *__restrict__ b,
                                                      double *__restrict__ r,
                                                      double *__restrict__ w, double safe1,
                                                      double safe2, unsigned long long *slot)
{
    __shared__ double red[2][4];
    const int k = blockIdx.x;
    const double *col = A + (long)k * lda;
    double sr = 0.0, sw = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = col[i], xi = x[i];
        sr = fma(a, xi, sr);
        sw = fma(fabs(a), fabs(xi), sw);
    }
    sr = wave_sum(sr);
    sw = wave_sum(sw);
    if ((threadIdx.x & 63) == 0) { red[0][threadIdx.x >> 6] = sr; red[1][threadIdx.x >> 6] = sw; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double tr = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        const double tw = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        const double rk = b[k] - tr, wk = fabs(b[k]) + tw;
        r[k] = rk;
        w[k] = wk;
        amax_nonneg(&slot[SL_MAX], berr_term(rk, wk, safe1, safe2));
    }
}

// w_i <- |r_i| + nz eps w_i (+ safe1 where w_i <= safe2)
__global__ void ferr_weight_kernel(const double *__restrict__ r, double *__restrict__ w, int n,
                                   double nzeps, double safe1, double safe2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double wi = w[i];
        w[i] = (wi > safe2) ? fabs(r[i]) + nzeps * wi : fabs(r[i]) + nzeps * wi + safe1;
    }
}

__global__ void vadd_kernel(double *__restrict__ x, const double *__restrict__ d, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

__global__ void vmul_kernel(double *__restrict__ x, const double *__restrict__ w, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= w[i];
}

__global__ void vcopy_kernel(double *__restrict__ y, const double *__restrict__ x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i];
}

__global__ void vfill_kernel(double *__restrict__ x, int n, double v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// x = e_j
__global__ void unit_vec_kernel(double *__restrict__ x, int n, int j)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (i == j) ? 1.0 : 0.0;
}

// lacn2's final-stage vector: x_i = (-1)^i (1 + i / (n - 1))
__global__ void altsgn_kernel(double *__restrict__ x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = ((i & 1) ? -1.0 : 1.0) * (1.0 + (double)i / (double)(n - 1));
}

FLX_DEV double sign1(double y)   // flens/lapack/auxiliary/sign.tcc:48-62
{
    if (y > 0.0) return 1.0;
    if (y < 0.0) return -1.0;
    return signbit(y) ? -1.0 : 1.0;
}

// count of i with nint(sign(1, x_i)) != sgn_i into SL_CNT
__global__ void sign_cmp_kernel(const double *__restrict__ x, const int *__restrict__ sgn, int n,
                                unsigned long long *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool diff = i < n && (int)floor(sign1(x[i]) + 0.5) != sgn[i];
    const unsigned long long b = __ballot(diff);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&slot[SL_CNT], 1ull);
}

// x_i <- sign(1, x_i), sgn_i <- nint(x_i)
__global__ void sign_set_kernel(double *__restrict__ x, int *__restrict__ sgn, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double s = sign1(x[i]);
        x[i] = s;
        sgn[i] = (int)floor(s + 0.5);
    }
}

// iamax, phase 1: max |x_i| over non-NaN entries
__global__ void absmax_vec_kernel(const double *__restrict__ x, int n, unsigned long long *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (i < n) { const double v = fabs(x[i]); a = (v == v) ? v : 0.0; }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((threadIdx.x & 63) == 0) amax_nonneg(&slot[SL_MAX], a);
}

// iamax, phase 2: first index holding the maximum (0 when x_0 is NaN: the
// reference's first maximum never moves past a leading NaN)
__global__ void argmax_vec_kernel(const double *__restrict__ x, int n, unsigned long long *slot)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mx = __longlong_as_double((long long)slot[SL_MAX]);
    if ((i == 0 && x[0] != x[0]) || fabs(x[i]) == mx) atomicMin(&slot[SL_IDX], (unsigned long long)i);
}

__global__ void slots_init_kernel(unsigned long long *slot, double big)
{
    const int i = threadIdx.x;
    if (i < SL_N) {
        unsigned long long v = 0;
        if (i == SL_MIN) v = (unsigned long long)__double_as_longlong(big);
        if (i == SL_ZERO || i == SL_IDX) v = ~0ull;
        slot[i] = v;
    }
}

} // namespace flx

using namespace flx;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

// workgroups of a column-strided matrix pass (at most 1024, which also
// bounds the partial sums of the Frobenius norm)
inline int col_grid(int n) { return n < 1 ? 1 : (n < 1024 ? n : 1024); }

inline int grid1(long n, int per = 256)
{
    long g = (n + per - 1) / per;
    return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

struct Aux {
    flx_context *h;
    hipStream_t s;
    AuxBufs &b;
    explicit Aux(flx_context *h_) : h(h_), s(h_->stream), b(h_->aux) {}

    unsigned long long *slot() { return (unsigned long long *)b.slot.p; }

    int need(int n, int nrhs = 1)
    {
        const size_t nn = (size_t)(n > 0 ? n : 1);
        FLX_TRY(ensure(h, b.slot, sizeof(unsigned long long) * SL_N));
        FLX_TRY(ensure(h, b.v, sizeof(double) * nn));
        FLX_TRY(ensure(h, b.x, sizeof(double) * nn));
        FLX_TRY(ensure(h, b.r, sizeof(double) * nn));
        FLX_TRY(ensure(h, b.w, sizeof(double) * nn));
        FLX_TRY(ensure(h, b.lines, sizeof(double) * nn));
        FLX_TRY(ensure(h, b.isgn, sizeof(int) * nn));
        FLX_TRY(ensure(h, b.part, sizeof(double) * (2 * nn * ((nn + RR - 1) / RR) + 4096)));
        (void)nrhs;
        return FLX_SUCCESS;
    }

    int init_slots(double big = DBL_MAX)
    {
        hipLaunchKernelGGL(slots_init_kernel, dim3(1), dim3(64), 0, s, slot(), big);
        FLX_TRY(hipGetLastError());
        return FLX_SUCCESS;
    }

    int read_slots(unsigned long long *out)
    {
        FLX_TRY(hipMemcpyAsync(out, slot(), sizeof(unsigned long long) * SL_N,
                               hipMemcpyDeviceToHost, s));
        FLX_TRY(hipStreamSynchronize(s));
        return FLX_SUCCESS;
    }

    static double as_double(unsigned long long u)
    {
        double d;
        memcpy(&d, &u, sizeof d);
        return d;
    }

    // ---- lan ---------------------------------------------------------------
    int lange(char norm, int m, int n, const double *A, int lda, double *value)
    {
        if ((m < n ? m : n) == 0) { *value = 0.0; return FLX_SUCCESS; }
        const int L = m > n ? m : n;
        FLX_RET(need(L));
        unsigned long long hs[SL_N];
        FLX_RET(init_slots());
        double *lines = (double *)b.lines.p;
        if (norm == 'M') {
            hipLaunchKernelGGL(mat_absmax_kernel, dim3(col_grid(n)), dim3(256), 0, s, A,
                               (long)lda, m, n, slot());
            FLX_RET(read_slots(hs));
            *value = as_double(hs[SL_MAX]);
        } else if (norm == 'O' || norm == 'I') {
            const bool rows = norm == 'I';
            const int lines_n = rows ? m : n;
            if (rows)
                hipLaunchKernelGGL((slab_reduce_kernel<true, 0>), dim3((m + SLAB - 1) / SLAB),
                                   dim3(256), 0, s, A, (long)lda, m, n, (const double *)nullptr,
                                   lines);
            else
                hipLaunchKernelGGL((slab_reduce_kernel<false, 0>), dim3((n + SLAB - 1) / SLAB),
                                   dim3(256), 0, s, A, (long)lda, m, n, (const double *)nullptr,
                                   lines);
            hipLaunchKernelGGL(vec_stats_kernel, dim3(grid1(lines_n)), dim3(256), 0, s,
                               (const double *)lines, lines_n, DBL_MAX, slot());
            FLX_RET(read_slots(hs));
            *value = as_double(hs[SL_MAX]);
        } else {
            hipLaunchKernelGGL(fro_scan_kernel, dim3(col_grid(n)), dim3(256), 0, s, A,
                               (long)lda, m, n, slot());
            FLX_RET(read_slots(hs));
            const double mx = as_double(hs[SL_MAX]);
            if (hs[SL_FLAG] || hs[SL_CNT] >= 2) { *value = NAN; return FLX_SUCCESS; }
            if (hs[SL_CNT] == 1) { *value = INFINITY; return FLX_SUCCESS; }
            if (mx == 0.0) { *value = 0.0; return FLX_SUCCESS; }
            const int P = col_grid(n);
            hipLaunchKernelGGL(fro_sum_kernel, dim3(P), dim3(256), 0, s, A, (long)lda, m, n, mx,
                               (double *)b.part.p);
            hipLaunchKernelGGL(fixed_sum_kernel, dim3(1), dim3(256), 0, s,
                               (const double *)b.part.p, P, slot());
            FLX_RET(read_slots(hs));
            *value = mx * std::sqrt(as_double(hs[SL_SUM]));
        }
        FLX_TRY(hipGetLastError());
        return FLX_SUCCESS;
    }

    // ---- equ ---------------------------------------------------------------
    int geequ(int m, int n, const double *A, int lda, double *r, double *c, double *rowcnd,
              double *colcnd, double *amax, int *info)
    {
        *info = 0;
        if (m == 0 || n == 0) { *rowcnd = 1.0; *colcnd = 1.0; *amax = 0.0; return FLX_SUCCESS; }
        FLX_RET(need(m > n ? m : n));
        const double small = SAFEMIN, big = 1.0 / small;
        unsigned long long hs[SL_N];
        hipLaunchKernelGGL((slab_reduce_kernel<true, 1>), dim3((m + SLAB - 1) / SLAB), dim3(256),
                           0, s, A, (long)lda, m, n, (const double *)nullptr, r);
        FLX_RET(init_slots(big));
        hipLaunchKernelGGL(vec_stats_kernel, dim3(grid1(m)), dim3(256), 0, s, (const double *)r,
                           m, big, slot());
        FLX_RET(read_slots(hs));
        double rcmax = as_double(hs[SL_MAX]), rcmin = as_double(hs[SL_MIN]);
        *amax = rcmax;
        if (rcmin == 0.0) { *info = (int)hs[SL_ZERO] + 1; return FLX_SUCCESS; }
        hipLaunchKernelGGL(inv_clamp_kernel, dim3(grid1(m)), dim3(256), 0, s, r, m, small, big);
        *rowcnd = std::max(rcmin, small) / std::min(rcmax, big);
        hipLaunchKernelGGL((slab_reduce_kernel<false, 2>), dim3((n + SLAB - 1) / SLAB), dim3(256),
                           0, s, A, (long)lda, m, n, (const double *)r, c);
        FLX_RET(init_slots(big));
        hipLaunchKernelGGL(vec_stats_kernel, dim3(grid1(n)), dim3(256), 0, s, (const double *)c,
                           n, big, slot());
        FLX_RET(read_slots(hs));
        rcmax = as_double(hs[SL_MAX]);
        rcmin = as_double(hs[SL_MIN]);
        if (rcmin == 0.0) { *info = m + (int)hs[SL_ZERO] + 1; return FLX_SUCCESS; }
        hipLaunchKernelGGL(inv_clamp_kernel, dim3(grid1(n)), dim3(256), 0, s, c, n, small, big);
        *colcnd = std::max(rcmin, small) / std::min(rcmax, big);
        FLX_TRY(hipGetLastError());
        return FLX_SUCCESS;
    }

    // ---- laq ---------------------------------------------------------------
    int laqge(int m, int n, double *A, int lda, const double *r, const double *c, double rowcnd,
              double colcnd, double amax, char *equed)
    {
        *equed = 'N';
        if (m == 0 || n == 0) return FLX_SUCCESS;
        const double thresh = 0.1, small = SAFEMIN / PREC64, large = 1.0 / small;
        char mode;
        if (rowcnd >= thresh && amax >= small && amax <= large) {
            if (colcnd >= thresh) return FLX_SUCCESS;
            mode = 'C';
        } else if (colcnd >= thresh) {
            mode = 'R';
        } else {
            mode = 'B';
        }
        hipLaunchKernelGGL(laqge_kernel, dim3(col_grid(n)), dim3(256), 0, s, A, (long)lda, m, n,
                           r, c, (int)mode);
        FLX_TRY(hipGetLastError());
        *equed = mode;
        return FLX_SUCCESS;
    }

    // ---- lacn2 vector steps -------------------------------------------------
    int asum(const double *x, int n, double *out)
    {
        const int P = grid1(n) < 256 ? grid1(n) : 256;
        hipLaunchKernelGGL(asum_part_kernel, dim3(P), dim3(256), 0, s, x, n, (double *)b.part.p);
        hipLaunchKernelGGL(fixed_sum_kernel, dim3(1), dim3(256), 0, s, (const double *)b.part.p,
                           P, slot());
        unsigned long long hs[SL_N];
        FLX_RET(read_slots(hs));
        *out = as_double(hs[SL_SUM]);
        return FLX_SUCCESS;
    }

    int iamax(const double *x, int n, int *out)   // 0-based
    {
        FLX_RET(init_slots());
        hipLaunchKernelGGL(absmax_vec_kernel, dim3(grid1(n)), dim3(256), 0, s, x, n, slot());
        hipLaunchKernelGGL(argmax_vec_kernel, dim3(grid1(n)), dim3(256), 0, s, x, n, slot());
        unsigned long long hs[SL_N];
        FLX_RET(read_slots(hs));
        *out = hs[SL_IDX] == ~0ull ? 0 : (int)hs[SL_IDX];
        return FLX_SUCCESS;
    }

    int get(const double *x, int i, double *out)
    {
        FLX_TRY(hipMemcpyAsync(out, x + i, sizeof(double), hipMemcpyDeviceToHost, s));
        FLX_TRY(hipStreamSynchronize(s));
        return FLX_SUCCESS;
    }

    // flens/lapack/la/lacn2.tcc:55-185, vectors on the device
    struct Lacn2 {
        int isave[3] = {0, 0, 0};
        int kase = 0;
        double est = 0.0;
    };
    int lacn2(Lacn2 &st, int n, double *v, double *x, int *sgn)
    {
        const int itmax = 5;
        const int g = grid1(n);
        if (st.kase == 0) {
            hipLaunchKernelGGL(vfill_kernel, dim3(g), dim3(256), 0, s, x, n, 1.0 / (double)n);
            st.kase = 1;
            st.isave[0] = 1;
            return FLX_SUCCESS;
        }
        bool main_loop = false, complete = false;
        switch (st.isave[0]) {
        case 1:
            if (n == 1) {
                hipLaunchKernelGGL(vcopy_kernel, dim3(1), dim3(64), 0, s, v, (const double *)x, 1);
                double x0;
                FLX_RET(get(x, 0, &x0));
                st.est = std::fabs(x0);
                st.kase = 0;
                return FLX_SUCCESS;
            }
            FLX_RET(asum(x, n, &st.est));
            hipLaunchKernelGGL(sign_set_kernel, dim3(g), dim3(256), 0, s, x, sgn, n);
            st.kase = 2;
            st.isave[0] = 2;
            return FLX_SUCCESS;
        case 2: {
            int i;
            FLX_RET(iamax(x, n, &i));
            st.isave[1] = i + 1;
            st.isave[2] = 2;
            main_loop = true;
            break;
        }
        case 3: {
            hipLaunchKernelGGL(vcopy_kernel, dim3(g), dim3(256), 0, s, v, (const double *)x, n);
            const double estold = st.est;
            FLX_RET(asum(v, n, &st.est));
            FLX_RET(init_slots());
            hipLaunchKernelGGL(sign_cmp_kernel, dim3(g), dim3(256), 0, s, (const double *)x,
                               (const int *)sgn, n, slot());
            unsigned long long hs[SL_N];
            FLX_RET(read_slots(hs));
            if (hs[SL_CNT] == 0 || st.est <= estold) {
                complete = true;
                break;
            }
            hipLaunchKernelGGL(sign_set_kernel, dim3(g), dim3(256), 0, s, x, sgn, n);
            st.kase = 2;
            st.isave[0] = 4;
            return FLX_SUCCESS;
        }
        case 4: {
            const int jlast = st.isave[1];
            int i;
            FLX_RET(iamax(x, n, &i));
            st.isave[1] = i + 1;
            double xj, xi;
            FLX_RET(get(x, jlast - 1, &xj));
            FLX_RET(get(x, i, &xi));
            if (xj != std::fabs(xi) && st.isave[2] < itmax) {
                ++st.isave[2];
                main_loop = true;
            } else {
                complete = true;
            }
            break;
        }
        case 5: {
            double sx;
            FLX_RET(asum(x, n, &sx));
            const double temp = 2.0 * (sx / (double)(3 * n));
            if (temp > st.est) {
                hipLaunchKernelGGL(vcopy_kernel, dim3(g), dim3(256), 0, s, v, (const double *)x, n);
                st.est = temp;
            }
            st.kase = 0;
            return FLX_SUCCESS;
        }
        }
        if (main_loop) {
            hipLaunchKernelGGL(unit_vec_kernel, dim3(g), dim3(256), 0, s, x, n, st.isave[1] - 1);
            st.kase = 1;
            st.isave[0] = 3;
            return FLX_SUCCESS;
        }
        if (complete) {
            hipLaunchKernelGGL(altsgn_kernel, dim3(g), dim3(256), 0, s, x, n);
            st.kase = 1;
            st.isave[0] = 5;
            return FLX_SUCCESS;
        }
        st.kase = 0;
        return FLX_SUCCESS;
    }

    // identity interchanges for the unpermuted triangular solves of con
    int identity(int n)
    {
        if (b.ident_n >= n) return FLX_SUCCESS;
        std::vector<int> piv(n), perm(n);
        for (int i = 0; i < n; ++i) { piv[i] = i + 1; perm[i] = i; }
        FLX_TRY(ensure(h, b.ipiv1, sizeof(int) * n));
        FLX_TRY(ensure(h, b.perm0, sizeof(int) * n));
        FLX_TRY(hipMemcpyAsync(b.ipiv1.p, piv.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        FLX_TRY(hipMemcpyAsync(b.perm0.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        FLX_TRY(hipStreamSynchronize(s));
        b.ident_n = n;
        return FLX_SUCCESS;
    }

    // ---- con ---------------------------------------------------------------
    int gecon(char norm, int n, const double *LU, int lda, double anorm, double *rcond)
    {
        *rcond = 0.0;
        if (n == 0) { *rcond = 1.0; return FLX_SUCCESS; }
        if (anorm == 0.0) return FLX_SUCCESS;
        FLX_RET(need(n));
        FLX_RET(identity(n));
        double *x = (double *)b.x.p, *v = (double *)b.v.p;
        int *sgn = (int *)b.isgn.p;
        const int kase1 = norm == 'O' ? 1 : 2;
        // The solves are latrs without its rescaling, which it needs only
        // when a partial solution times a column's off-diagonal 1-norm nears
        // bignum = precision / safemin.  Every partial solution is bounded by
        // (1 + n max|LU|) max(|x_in|, |x_out|), so solutions below the limit
        // here are ones latrs computes unscaled; beyond it the estimate is 0.
        double lumax = 0.0;
        FLX_RET(lange('M', n, n, LU, lda, &lumax));
        const double limit = (PREC64 / SAFEMIN) / (16.0 * (1.0 + (double)n * lumax));
        Lacn2 st;
        for (;;) {
            FLX_RET(lacn2(st, n, v, x, sgn));
            if (st.kase == 0) break;
            const int tr = (st.kase == kase1) ? 0 : 1;
            FLX_RET(dgetrs_dev(h, tr, n, 1, LU, lda, (const int *)b.ipiv1.p,
                               (const int *)b.perm0.p, x, n, (const int *)b.perm0.p));
            int i;
            FLX_RET(iamax(x, n, &i));
            double xi;
            FLX_RET(get(x, i, &xi));
            if (!(std::fabs(xi) <= limit)) return FLX_SUCCESS;   // rcond = 0
        }
        if (st.est != 0.0) *rcond = (1.0 / st.est) / anorm;
        return FLX_SUCCESS;
    }

    // ---- rfs ---------------------------------------------------------------
    int residual(int trans, int n, const double *A, int lda, const double *x, const double *bj,
                 double *r, double *w, double safe1, double safe2, double *berr)
    {
        FLX_RET(init_slots());
        if (!trans) {
            const int S = (n + RR - 1) / RR;
            double *pr = (double *)b.part.p, *pw = pr + (size_t)S * n;
            hipLaunchKernelGGL(resid_nt_kernel, dim3((n + 511) / 512, S), dim3(256), 0, s, A,
                               (long)lda, n, x, pr, pw);
            hipLaunchKernelGGL(resid_nt_fin_kernel, dim3(grid1(n)), dim3(256), 0, s,
                               (const double *)pr, (const double *)pw, S, n, bj, r, w, safe1,
                               safe2, slot());
        } else {
            hipLaunchKernelGGL(resid_t_kernel, dim3(n), dim3(256), 0, s, A, (long)lda, n, x, bj,
                               r, w, safe1, safe2, slot());
        }
        FLX_TRY(hipGetLastError());
        unsigned long long hs[SL_N];
        FLX_RET(read_slots(hs));
        *berr = as_double(hs[SL_MAX]);
        return FLX_SUCCESS;
    }

    int gerfs(int trans, int n, int nrhs, const double *A, int lda, const double *AF, int ldaf,
              const int *ipiv, const double *B, int ldb, double *X, int ldx, double *ferr,
              double *berr)
    {
        if (n == 0 || nrhs == 0) {
            for (int j = 0; j < nrhs; ++j) { ferr[j] = 0.0; berr[j] = 0.0; }
            return FLX_SUCCESS;
        }
        FLX_RET(need(n));
        const int itmax = 5, nz = n + 1;
        const double eps = EPS64, safe1 = nz * SAFEMIN, safe2 = safe1 / eps;
        double *r = (double *)b.r.p, *w = (double *)b.w.p, *v = (double *)b.v.p;
        int *sgn = (int *)b.isgn.p;
        const int g = grid1(n);
        // the row orders of ipiv, once for all the solves below
        FLX_TRY(ensure(h, b.pf, sizeof(int) * (size_t)n));
        FLX_TRY(ensure(h, b.pb, sizeof(int) * (size_t)n));
        const int *pf = (const int *)b.pf.p, *pb = (const int *)b.pb.p;
        FLX_TRY(ipiv_to_perm(s, ipiv, (int *)b.pf.p, n, 0));
        FLX_TRY(ipiv_to_perm(s, ipiv, (int *)b.pb.p, n, 1));
        auto solve = [&](int t, double *y) -> int {
            return dgetrs_dev(h, t, n, 1, AF, ldaf, ipiv, pf, y, n, pb);
        };
        for (int j = 0; j < nrhs; ++j) {
            const double *bj = B + (size_t)j * ldb;
            double *xj = X + (size_t)j * ldx;
            int count = 1;
            double lstres = 3.0;
            for (;;) {
                FLX_RET(residual(trans, n, A, lda, xj, bj, r, w, safe1, safe2, &berr[j]));
                if (berr[j] > eps && 2.0 * berr[j] <= lstres && count <= itmax) {
                    FLX_RET(solve(trans, r));
                    hipLaunchKernelGGL(vadd_kernel, dim3(g), dim3(256), 0, s, xj,
                                       (const double *)r, n);
                    lstres = berr[j];
                    ++count;
                    continue;
                }
                break;
            }
            hipLaunchKernelGGL(ferr_weight_kernel, dim3(g), dim3(256), 0, s, (const double *)r, w,
                               n, (double)nz * eps, safe1, safe2);
            // lacn2 estimates || inv(op(A)) diag(w) ||_inf: kase 1 applies
            // diag(w) inv(op(A))^T, kase 2 inv(op(A)) diag(w)
            Lacn2 st;
            for (;;) {
                FLX_RET(lacn2(st, n, v, r, sgn));
                if (st.kase == 0) break;
                if (st.kase == 1) {
                    FLX_RET(solve(!trans, r));
                    hipLaunchKernelGGL(vmul_kernel, dim3(g), dim3(256), 0, s, r,
                                       (const double *)w, n);
                } else {
                    hipLaunchKernelGGL(vmul_kernel, dim3(g), dim3(256), 0, s, r,
                                       (const double *)w, n);
                    FLX_RET(solve(trans, r));
                }
            }
            ferr[j] = st.est;
            int i;
            FLX_RET(iamax(xj, n, &i));
            double xm;
            FLX_RET(get(xj, i, &xm));
            xm = std::fabs(xm);
            if (xm != 0.0 && xm == xm) ferr[j] /= xm;
        }
        FLX_TRY(hipGetLastError());
        return FLX_SUCCESS;
    }
};

int norm_code(char c)
{
    switch (c) {
    case 'M': case 'm': return 'M';
    case 'O': case 'o': case '1': return 'O';
    case 'I': case 'i': return 'I';
    case 'F': case 'f': case 'E': case 'e': return 'F';
    default: return 0;
    }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI (device pointers; scalar results on the host)
// ---------------------------------------------------------------------------
extern "C" int flx_dlange(flx_handle h, char norm, int m, int n, const double *A, int lda,
                          double *value)
{
    ProfBind pb_(h);
    const int nc = norm_code(norm);
    if (!h || !value || !nc || m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLX_INVALID_ARG;
    Aux a(h);
    return a.lange((char)nc, m, n, A, lda, value);
}

extern "C" int flx_dgeequ(flx_handle h, int m, int n, const double *A, int lda, double *r,
                          double *c, double *rowcnd, double *colcnd, double *amax, int *info)
{
    ProfBind pb_(h);
    if (!h || !rowcnd || !colcnd || !amax || !info || m < 0 || n < 0 ||
        lda < (m > 1 ? m : 1))
        return FLX_INVALID_ARG;
    Aux a(h);
    return a.geequ(m, n, A, lda, r, c, rowcnd, colcnd, amax, info);
}

extern "C" int flx_dlaqge(flx_handle h, int m, int n, double *A, int lda, const double *r,
                          const double *c, double rowcnd, double colcnd, double amax,
                          char *equed)
{
    ProfBind pb_(h);
    if (!h || !equed || m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLX_INVALID_ARG;
    Aux a(h);
    return a.laqge(m, n, A, lda, r, c, rowcnd, colcnd, amax, equed);
}

extern "C" int flx_dgecon(flx_handle h, char norm, int n, const double *LU, int lda,
                          double anorm, double *rcond)
{
    ProfBind pb_(h);
    const int nc = norm_code(norm);
    if (!h || !rcond || (nc != 'O' && nc != 'I') || n < 0 || lda < (n > 1 ? n : 1) ||
        anorm < 0.0)
        return FLX_INVALID_ARG;
    Aux a(h);
    FLX_RET(a.gecon((char)nc, n, LU, lda, anorm, rcond));
    return flx_sync(h);   // a bounded wait of the solves that gave up is reported here
}

extern "C" int flx_dgerfs(flx_handle h, int trans, int n, int nrhs, const double *A, int lda,
                          const double *AF, int ldaf, const int *ipiv, const double *B, int ldb,
                          double *X, int ldx, double *ferr, double *berr)
{
    ProfBind pb_(h);
    const int l = n > 1 ? n : 1;
    if (!h || (nrhs > 0 && (!ferr || !berr)) || n < 0 || nrhs < 0 || lda < l || ldaf < l ||
        ldb < l || ldx < l)
        return FLX_INVALID_ARG;
    Aux a(h);
    FLX_RET(a.gerfs(trans ? 1 : 0, n, nrhs, A, lda, AF, ldaf, ipiv, B, ldb, X, ldx, ferr, berr));
    return flx_sync(h);
}

extern "C" int flx_dgesvx(flx_handle h, char fact, int trans, int n, int nrhs, double *A,
                          int lda, double *AF, int ldaf, int *ipiv, char *equed, double *r,
                          double *c, double *B, int ldb, double *X, int ldx, double *rcond,
                          double *ferr, double *berr, double *rpvgrw, int *info)
{
    ProfBind pb_(h);
    const int l = n > 1 ? n : 1;
    if (!h || !equed || !rcond || !rpvgrw || !info || (nrhs > 0 && (!ferr || !berr)) || n < 0 ||
        nrhs < 0 || lda < l || ldaf < l || ldb < l || ldx < l)
        return FLX_INVALID_ARG;
    if (fact != 'N' && fact != 'E' && fact != 'F') return FLX_INVALID_ARG;
    if (n > LU_N_MAX) return FLX_UNSUPPORTED;
    Aux a(h);
    hipStream_t s = h->stream;
    const double small = SAFEMIN, big = 1.0 / small;
    *info = 0;
    bool rowequ = false, colequ = false;
    double rowcnd = 1.0, colcnd = 1.0;
    if (fact == 'N' || fact == 'E') {
        *equed = 'N';
    } else {
        rowequ = (*equed == 'R' || *equed == 'B');
        colequ = (*equed == 'C' || *equed == 'B');
    }
    if (n > 0) FLX_RET(a.need(n));
    unsigned long long hs[SL_N];
    auto vec_cond = [&](const double *v, double *cond) -> int {
        FLX_RET(a.init_slots(big));
        hipLaunchKernelGGL(vec_stats_kernel, dim3(grid1(n)), dim3(256), 0, s, v, n, big, a.slot());
        FLX_RET(a.read_slots(hs));
        const double mn = Aux::as_double(hs[SL_MIN]), mx = Aux::as_double(hs[SL_MAX]);
        // the reference asserts rcMin > 0 (svx.tcc:104-119)
        if (hs[SL_FLAG] || !(mn > 0.0)) return FLX_INVALID_ARG;
        *cond = n > 0 ? std::max(mn, small) / std::min(mx, big) : 1.0;
        return FLX_SUCCESS;
    };
    if (rowequ && n > 0) FLX_RET(vec_cond(r, &rowcnd));
    if (colequ && n > 0) FLX_RET(vec_cond(c, &colcnd));
    if (fact == 'E') {
        double amax = 0.0;
        int einfo = 0;
        FLX_RET(a.geequ(n, n, A, lda, r, c, &rowcnd, &colcnd, &amax, &einfo));
        if (einfo == 0) {
            FLX_RET(a.laqge(n, n, A, lda, r, c, rowcnd, colcnd, amax, equed));
            rowequ = (*equed == 'R' || *equed == 'B');
            colequ = (*equed == 'C' || *equed == 'B');
        }
    }
    if (!trans) {
        if (rowequ && nrhs > 0)
            hipLaunchKernelGGL(scale_rows_kernel, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, B,
                               (long)ldb, n, nrhs, (const double *)r);
    } else if (colequ && nrhs > 0) {
        hipLaunchKernelGGL(scale_rows_kernel, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, B,
                           (long)ldb, n, nrhs, (const double *)c);
    }
    FLX_TRY(hipGetLastError());
    // max |U| of the leading k x k block of AF (upper triangle)
    auto umax = [&](int k, double *out) -> int {
        FLX_RET(a.init_slots());
        hipLaunchKernelGGL(upper_absmax_kernel, dim3(col_grid(k)), dim3(256), 0, s,
                           (const double *)AF, (long)ldaf, k, a.slot());
        FLX_RET(a.read_slots(hs));
        *out = Aux::as_double(hs[SL_MAX]);
        return FLX_SUCCESS;
    };
    if (fact == 'N' || fact == 'E') {
        if (n > 0) {
            FLX_TRY(hipMemcpy2DAsync(AF, sizeof(double) * ldaf, A, sizeof(double) * lda,
                                     sizeof(double) * n, n, hipMemcpyDeviceToDevice, s));
            FLX_RET(dgetrf_dev(h, n, n, AF, ldaf, ipiv, (int *)a.b.x.p, true));
            int finfo = 0;
            FLX_TRY(hipMemcpyAsync(&finfo, a.b.x.p, sizeof(int), hipMemcpyDeviceToHost, s));
            FLX_TRY(hipStreamSynchronize(s));
            if (finfo > 0) {
                double g, am;
                FLX_RET(umax(finfo, &g));
                FLX_RET(a.lange('M', n, n, A, lda, &am));
                *rpvgrw = (g == 0.0) ? 1.0 : am / g;
                *rcond = 0.0;
                *info = finfo;
                return FLX_SUCCESS;
            }
        }
    }
    const char nrm = trans ? 'I' : 'O';
    double anorm = 0.0, g = 0.0, am = 0.0;
    FLX_RET(a.lange(nrm, n, n, A, lda, &anorm));
    if (n > 0) FLX_RET(umax(n, &g));
    FLX_RET(a.lange('M', n, n, A, lda, &am));
    *rpvgrw = (g == 0.0) ? 1.0 : am / g;
    FLX_RET(a.gecon(nrm, n, AF, ldaf, anorm, rcond));
    if (n > 0 && nrhs > 0) {
        FLX_TRY(hipMemcpy2DAsync(X, sizeof(double) * ldx, B, sizeof(double) * ldb,
                                 sizeof(double) * n, nrhs, hipMemcpyDeviceToDevice, s));
        FLX_RET(dgetrs_dev(h, trans ? 1 : 0, n, nrhs, AF, ldaf, ipiv, nullptr, X, ldx));
    }
    FLX_RET(a.gerfs(trans ? 1 : 0, n, nrhs, A, lda, AF, ldaf, ipiv, B, ldb, X, ldx, ferr, berr));
    const double *sc = nullptr;
    double cnd = 1.0;
    if (!trans) { if (colequ) { sc = c; cnd = colcnd; } }
    else if (rowequ) { sc = r; cnd = rowcnd; }
    if (sc && n > 0) {
        if (nrhs > 0)
            hipLaunchKernelGGL(scale_rows_kernel, dim3(grid1((long)n * nrhs)), dim3(256), 0, s, X,
                               (long)ldx, n, nrhs, sc);
        for (int j = 0; j < nrhs; ++j) ferr[j] /= cnd;
    }
    FLX_TRY(hipGetLastError());
    FLX_RET(flx_sync(h));
    if (*rcond < EPS64) *info = n + 1;
    return FLX_SUCCESS;
}
