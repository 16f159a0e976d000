// rbt.hip -- recursive random butterfly transforms of depth 2 for the
// mixed-precision solver's pivot-free factor.
//
// The low-precision LU only preconditions FP64 refinement, so it need not
// reproduce the reference's partial pivoting (flens/lapack/ge/tf2.tcc:
// 88-133).  With A' = U^T A V for random recursive butterflies U and V
// (Parker, "Random butterfly transformations with applications in
// computational linear algebra", UCLA CSD-950023, 1995; Baboulin, Dongarra,
// Herrmann, Tomov, "Accelerating linear system solutions using
// randomization techniques", ACM TOMS 39(2), 2013) Gaussian elimination
// without pivoting is stable with high probability, and its panel needs no
// pivot search across workgroups.  The preconditioner is then
//     M^{-1} = D_c V (L U)^{-1} U^T D_r
// for A' = U^T D_r A D_c V = L U (D_r, D_c: the optional power-of-two
// equilibration of the FP16 factor).
//
// A butterfly of order n (even) is B = [R0 R1; R0 -R1] / sqrt(2) with R0,
// R1 random diagonal; depth 2 is U = diag(B', B'') B with B', B'' of order
// n / 2.  For n = 4q it couples only the indices {i, i+q, i+2q, i+3q}
// (i < q), so U is a direct sum of 4 x 4 blocks up to a permutation and
// every transform below works on such groups.  Storage: per butterfly two
// arrays of n doubles, r1 (level 1: R0 at [a], R1 at [a + n/2]) and r2
// (level 2: the halves' R0 at [a], R1 at [a + n/4], and [n/2 + a],
// [3n/4 + a]); a pair (a, b) of a level takes R0 = r[a], R1 = r[b].  The
// 1/sqrt(2) of each level is folded into its arrays.
#include <cfloat>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

namespace {

// x <- U^T x on the group of index i (U from its level arrays r1, r2)
FLX_DEV void bfly_t(const double *__restrict__ r1, const double *__restrict__ r2, int i, int q,
                    double x[4])
{
    const int p0 = i, p1 = i + q, p2 = i + 2 * q, p3 = i + 3 * q;
    // level 2 transposed: pairs (0, 1) and (2, 3)
    const double a0 = r2[p0] * (x[0] + x[1]), a1 = r2[p1] * (x[0] - x[1]);
    const double a2 = r2[p2] * (x[2] + x[3]), a3 = r2[p3] * (x[2] - x[3]);
    // level 1 transposed: pairs (0, 2) and (1, 3)
    x[0] = r1[p0] * (a0 + a2);
    x[2] = r1[p2] * (a0 - a2);
    x[1] = r1[p1] * (a1 + a3);
    x[3] = r1[p3] * (a1 - a3);
}

// x <- V x on the group of index i
FLX_DEV void bfly_n(const double *__restrict__ r1, const double *__restrict__ r2, int i, int q,
                    double x[4])
{
    const int p0 = i, p1 = i + q, p2 = i + 2 * q, p3 = i + 3 * q;
    // level 1: pairs (0, 2) and (1, 3)
    const double a0 = r1[p0] * x[0] + r1[p2] * x[2], a2 = r1[p0] * x[0] - r1[p2] * x[2];
    const double a1 = r1[p1] * x[1] + r1[p3] * x[3], a3 = r1[p1] * x[1] - r1[p3] * x[3];
    // level 2: pairs (0, 1) and (2, 3)
    x[0] = r2[p0] * a0 + r2[p1] * a1;
    x[1] = r2[p0] * a0 - r2[p1] * a1;
    x[2] = r2[p2] * a2 + r2[p3] * a3;
    x[3] = r2[p2] * a2 - r2[p3] * a3;
}

constexpr int RT = 16;   // tile of group indices per workgroup side

// W[r][c] = float((U^T D_r A D_c V)[r][c]), A column-major FP64, W row-major
// FP32; one 16 x 16 tile of group indices (i, j) per workgroup, i.e. the 16
// tiles A[i0 + a q.., j0 + b q..] for a, b < 4.  *flag |= 1 when a finite
// value exceeds FP32's range.
__global__ __launch_bounds__(256) void rbt_conv_kernel(const double *__restrict__ A, long lda,
                                                       int n, const double *__restrict__ dr,
                                                       const double *__restrict__ dc, Rbt rb,
                                                       float *__restrict__ W, long ldw,
                                                       int *__restrict__ flag)
{
    __shared__ double s[4][4][RT][RT + 1];
    const int q = n / 4;
    const int t = threadIdx.x;
    const int i0 = blockIdx.x * RT, j0 = blockIdx.y * RT;
    {
        // loads along the column (contiguous in A)
        const int ii = t % RT, jj = t / RT;
        const bool in = i0 + ii < q && j0 + jj < q;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                s[a][b][ii][jj] = in ? A[(long)(i0 + ii + a * q) + (long)(j0 + jj + b * q) * lda]
                                     : 0.0;
    }
    __syncthreads();
    {
        const int ii = t / RT, jj = t % RT;
        const int i = i0 + ii, j = j0 + jj;
        if (i < q && j < q) {
            double S[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    double v = s[a][b][ii][jj];
                    if (dr) v *= dr[i + a * q];
                    if (dc) v *= dc[j + b * q];
                    S[a][b] = v;
                }
            // rows: S[a][:] <- V^T S[a][:]  (S V); columns: U^T
#pragma unroll
            for (int a = 0; a < 4; ++a) bfly_t(rb.v1, rb.v2, j, q, S[a]);
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                double c[4] = {S[0][b], S[1][b], S[2][b], S[3][b]};
                bfly_t(rb.u1, rb.u2, i, q, c);
#pragma unroll
                for (int a = 0; a < 4; ++a) s[a][b][ii][jj] = c[a];
            }
        }
    }
    __syncthreads();
    {
        // stores along the row (contiguous in W)
        const int jj = t % RT, ii = t / RT;
        if (i0 + ii >= q || j0 + jj >= q) return;
        bool ovf = false;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const double v = s[a][b][ii][jj];
                ovf = ovf || (fabs(v) > (double)FLT_MAX && fabs(v) <= DBL_MAX);
                W[(long)(i0 + ii + a * q) * ldw + (j0 + jj + b * q)] = (float)v;
            }
        if (ovf) atomicOr(flag, 1);
    }
}

// y = round32(2^-e U^T D_r v), 2^e from ||v||_inf (as prec_in_kernel)
__global__ void rbt_in_kernel(const double *__restrict__ v, const double *__restrict__ dr,
                              const unsigned long long *nrm, Rbt rb, float *__restrict__ y32,
                              double *__restrict__ y64, int n)
{
    const int q = n / 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= q) return;
    const double sc = nrm ? pow2_recip(__longlong_as_double((long long)*nrm)) : 1.0;
    double x[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int p = i + a * q;
        x[a] = dr ? v[p] * dr[p] : v[p];
    }
    bfly_t(rb.u1, rb.u2, i, q, x);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        if (y32) y32[i + a * q] = (float)(x[a] * sc);
        else y64[i + a * q] = x[a];
    }
}

// out = (add ? out : 0) + D_c V y 2^e; ||out||_inf -> onrm (optional)
__global__ void rbt_out_kernel(const float *__restrict__ y32, const double *__restrict__ y64,
                               const double *__restrict__ dc, const unsigned long long *nrm,
                               Rbt rb, double *__restrict__ out, int add, int n,
                               unsigned long long *onrm)
{
    const int q = n / 4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (i < q) {
        const double sc = nrm ? 1.0 / pow2_recip(__longlong_as_double((long long)*nrm)) : 1.0;
        double x[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) x[a] = y32 ? (double)y32[i + a * q] * sc : y64[i + a * q];
        bfly_n(rb.v1, rb.v2, i, q, x);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int p = i + a * q;
            double t = dc ? x[a] * dc[p] : x[a];
            t = add ? out[p] + t : t;
            out[p] = t;
            m = fmax(m, fabs(t));
        }
    }
    if (onrm) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) atomic_max_nonneg(onrm, m);
    }
}

} // namespace

bool rbt_supported(int n) { return n >= 4 && n % 4 == 0; }

// splitmix64 stream -> exp(r / 10) with r uniform in [-1/2, 1/2), over sqrt(2)
void rbt_values(int n, unsigned long long seed, double *out)
{
    unsigned long long z = seed;
    for (int k = 0; k < n; ++k) {
        z += 0x9E3779B97F4A7C15ull;
        unsigned long long x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        const double r = (double)(x >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        out[k] = exp(r / 10.0) * 0.70710678118654752440;
    }
}

hipError_t rbt_conv(hipStream_t s, const double *A, long lda, int n, const double *dr,
                    const double *dc, const Rbt &rb, float *W, long ldw, int *flag)
{
    if (!rbt_supported(n)) return hipErrorInvalidValue;
    const int q = n / 4;
    ProfScope ps(tl_prof, s, PROF_TRANSPOSE, 0.0, 12.0 * n * (double)n);
    dim3 grid((q + RT - 1) / RT, (q + RT - 1) / RT);
    hipLaunchKernelGGL(rbt_conv_kernel, grid, dim3(256), 0, s, A, lda, n, dr, dc, rb, W, ldw,
                       flag);
    return hipGetLastError();
}

hipError_t rbt_in(hipStream_t s, const double *v, const double *dr,
                  const unsigned long long *nrm, const Rbt &rb, float *y32, double *y64, int n)
{
    const int q = n / 4;
    hipLaunchKernelGGL(rbt_in_kernel, dim3((q + 255) / 256), dim3(256), 0, s, v, dr, nrm, rb,
                       y32, y64, n);
    return hipGetLastError();
}

hipError_t rbt_out(hipStream_t s, const float *y32, const double *y64, const double *dc,
                   const unsigned long long *nrm, const Rbt &rb, double *out, int add, int n,
                   unsigned long long *onrm)
{
    const int q = n / 4;
    hipLaunchKernelGGL(rbt_out_kernel, dim3((q + 255) / 256), dim3(256), 0, s, y32, y64, dc,
                       nrm, rb, out, add, n, onrm);
    return hipGetLastError();
}

} // namespace flx
