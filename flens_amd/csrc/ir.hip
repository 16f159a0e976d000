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
//     block, each waiting (bounded) on per-block ready flags of the blocks
//     it depends on, which it consumes as they appear; the diagonal block
//     is applied through its precomputed inverse;
//   * vector kernels for GMRES (several dot products at once, w -= V h).
// The host-side loops (classic refinement with dsgesv's stopping rule,
// netlib dsgesv.f:199-296, and GMRES-IR) are in api.hip.
#include <cfloat>
#include <climits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

// max of non-negative doubles through their bit patterns
FLX_DEV void atomic_max_nonneg(unsigned long long *slot, double v)
{
    if (v > 0.0 || v != v) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

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
    if (i + 1 < m && (lda & 1) == 0) {
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

// power of two 2^-e with v = f 2^e, f in [0.5, 1): scales v into [0.5, 1)
FLX_DEV double pow2_recip(double v)
{
    if (!(v > 0.0) || !(v <= DBL_MAX)) return 1.0;
    int e;
    frexp(v, &e);
    return ldexp(1.0, -e);
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
// Inverses of the 64 x 64 diagonal blocks of the FP32 factor (row-major W):
// inv_l[b] = L_bb^{-1} (unit lower), inv_u[b] = U_bb^{-1}; a partial last
// block is completed by the identity.  One workgroup (64 threads) per block
// and triangle, thread c computes column c.
// ---------------------------------------------------------------------------
constexpr int TV = 64;

template <int UPPER>
__global__ __launch_bounds__(64) void tri_inv_rm_kernel(const float *__restrict__ W, long ldw,
                                                        int n, float *__restrict__ inv)
{
    __shared__ float a[TV][TV + 1];
    __shared__ float x[TV][TV + 1];   // x[r][c]: column c of the inverse
    const int b = blockIdx.x, r0 = b * TV, rows = min(TV, n - r0);
    const int c = threadIdx.x;
    for (int r = 0; r < TV; ++r)
        a[r][c] = (r < rows && c < rows) ? W[(long)(r0 + r) * ldw + r0 + c] : (r == c ? 1.0f : 0.0f);
    __syncthreads();
    if (!UPPER) {
        for (int r = 0; r < TV; ++r) x[r][c] = 0.0f;
        x[c][c] = 1.0f;
        for (int r = c + 1; r < TV; ++r) {
            float v = 0.0f;
            for (int q = c; q < r; ++q) v = fmaf(-a[r][q], x[q][c], v);
            x[r][c] = v;
        }
    } else {
        for (int r = 0; r < TV; ++r) x[r][c] = 0.0f;
        x[c][c] = 1.0f / a[c][c];
        for (int r = c - 1; r >= 0; --r) {
            float v = 0.0f;
            for (int q = r + 1; q <= c; ++q) v = fmaf(-a[r][q], x[q][c], v);
            x[r][c] = v / a[r][r];
        }
    }
    __syncthreads();
    float *o = inv + (long)b * TV * TV;
    for (int r = 0; r < TV; ++r) o[r * TV + c] = x[r][c];
}

hipError_t tri_inv_rm(hipStream_t s, const float *W, long ldw, int n, float *inv_l, float *inv_u)
{
    const int nbk = (n + TV - 1) / TV;
    if (nbk <= 0) return hipSuccess;
    hipLaunchKernelGGL(tri_inv_rm_kernel<0>, dim3(nbk), dim3(64), 0, s, W, ldw, n, inv_l);
    hipLaunchKernelGGL(tri_inv_rm_kernel<1>, dim3(nbk), dim3(64), 0, s, W, ldw, n, inv_u);
    return hipGetLastError();
}

size_t trsv_inv_floats(int n) { return (size_t)((n + TV - 1) / TV) * TV * TV; }

// ---------------------------------------------------------------------------
// y <- T^{-1} y for T the unit-lower (UPPER = 0) or upper triangle of W.
// Workgroup g owns block row bi = g (lower) or nbk-1-g (upper), so every
// block it waits for belongs to a workgroup launched before it.  Thread
// (rr, q) = (tid / 4, tid % 4) accumulates row rr over columns q*16..+15 of
// each finished block.  A finished block publishes its 64 values and then
// flags[bi] = epoch (agent scope); waits are bounded (err[0] |= 2 on give-up).
// ---------------------------------------------------------------------------
constexpr int TV_SPIN_LIMIT = 1 << 22;

FLX_DEV float ld_agent_f(const float *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV void st_agent_f(float *p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV unsigned ld_agent_u(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV void st_agent_u(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int UPPER>
__global__ __launch_bounds__(256) void trsv_rm_kernel(const float *__restrict__ W, long ldw, int n,
                                                      const float *__restrict__ inv,
                                                      float *__restrict__ y,
                                                      unsigned *__restrict__ flags,
                                                      unsigned epoch, int *__restrict__ err)
{
    __shared__ int s_ready[2];
    __shared__ float s_rhs[TV];
    const int nbk = (n + TV - 1) / TV;
    const int g = blockIdx.x;
    const int bi = UPPER ? nbk - 1 - g : g;
    const int r0 = bi * TV, rows = min(TV, n - r0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rr = tid >> 2, q = tid & 3;
    const int ndep = UPPER ? nbk - 1 - bi : bi;
    const bool row_ok = rr < rows;
    const float *wrow = W + (long)(r0 + (row_ok ? rr : 0)) * ldw;
    // the inverse of the diagonal block: row rr, columns q*16..+15
    float iv[16];
    {
        const float4 *pi = reinterpret_cast<const float4 *>(inv + (long)bi * TV * TV + rr * TV + q * 16);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float4 v = pi[c];
            iv[4 * c] = v.x; iv[4 * c + 1] = v.y; iv[4 * c + 2] = v.z; iv[4 * c + 3] = v.w;
        }
    }
    const float rhs0 = row_ok ? y[r0 + rr] : 0.0f;
    float acc = 0.0f;
    int done = 0, it = 0, spins = 0;
    while (done < ndep) {
        if (wave == 0) {
            const int d = done + lane;
            bool ready = false;
            if (d < ndep) {
                const int jb = UPPER ? nbk - 1 - d : d;
                ready = ld_agent_u(&flags[jb]) == epoch;
            }
            // finished blocks form a prefix of the dependency order
            const uint64_t bal = __ballot(ready);
            int take = (bal == ~0ull) ? 64 : (int)__builtin_ctzll(~bal);
            take = min(take, ndep - done);
            if (lane == 0) {
                if (take == 0 && ++spins > TV_SPIN_LIMIT) {
                    atomicOr(err, 2);
                    take = ndep - done;   // give up: finish with what is there
                }
                s_ready[it & 1] = take;
            }
        }
        __syncthreads();
        const int take = s_ready[it & 1];
        ++it;
        if (take == 0) {
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        for (int d = done; d < done + take; ++d) {
            const int jb = UPPER ? nbk - 1 - d : d;
            const int c0 = jb * TV + q * 16;
            float xv[16], wv[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const bool ok = c0 + c < n;
                xv[c] = ok ? ld_agent_f(&y[c0 + c]) : 0.0f;
                wv[c] = (ok && row_ok) ? wrow[c0 + c] : 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) acc = fmaf(wv[c], xv[c], acc);
        }
        done += take;
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (q == 0) s_rhs[rr] = row_ok ? rhs0 - acc : 0.0f;
    __syncthreads();
    float v = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; ++c) v = fmaf(iv[c], s_rhs[q * 16 + c], v);
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    if (q == 0 && row_ok) st_agent_f(&y[r0 + rr], v);
    __threadfence();
    __syncthreads();
    if (tid == 0) st_agent_u(&flags[bi], epoch);
}

hipError_t trsv_rm(hipStream_t s, int upper, const float *W, long ldw, int n, const float *inv,
                   float *y, unsigned *flags, unsigned epoch, int *err)
{
    const int nbk = (n + TV - 1) / TV;
    if (nbk <= 0) return hipSuccess;
    ProfScope ps(tl_prof, s, PROF_TRS, (double)n * n, 2.0 * n * (double)n);
    if (upper)
        hipLaunchKernelGGL(trsv_rm_kernel<1>, dim3(nbk), dim3(256), 0, s, W, ldw, n, inv, y, flags,
                           epoch, err);
    else
        hipLaunchKernelGGL(trsv_rm_kernel<0>, dim3(nbk), dim3(256), 0, s, W, ldw, n, inv, y, flags,
                           epoch, err);
    return hipGetLastError();
}

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

__global__ void zero_u32_n_kernel(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// a fresh flag value per substitution launch; flags are cleared on wrap
static unsigned next_epoch(hipStream_t s, const PrecFactor &f)
{
    if (++*f.epoch == 0u) {
        const int nbk = (f.n + TV - 1) / TV;
        hipLaunchKernelGGL(zero_u32_n_kernel, dim3((nbk + 255) / 256), dim3(256), 0, s, f.flags, nbk);
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
    hipLaunchKernelGGL(prec_in_kernel, dim3((n + 255) / 256), dim3(256), 0, s, v,
                       (const int *)f.perm, f.dr, (const unsigned long long *)vnrm, f.y32, n);
    FLX_CHECK(trsv_rm(s, 0, f.W, f.ldw, n, f.inv_l, f.y32, f.flags, next_epoch(s, f), f.err));
    FLX_CHECK(trsv_rm(s, 1, f.W, f.ldw, n, f.inv_u, f.y32, f.flags, next_epoch(s, f), f.err));
    hipLaunchKernelGGL(prec_out_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       (const float *)f.y32, f.dc, (const unsigned long long *)vnrm, z, add, n,
                       znrm);
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
