/*
 * flens_oracle.c -- CPU oracle for the dense GEMM + LU-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * FLENS generic path (cxxblas gemm/trsm/ger/iamax, FLENS-LAPACK
 * trf/tf2/laswp/trs/sv) and of netlib dsgesv's mixed-precision refinement.
 * It is built into oracle/liboracle.so by oracle/Makefile and is loaded
 * only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * as the checker and as the timed single-core CPU baseline.  The product
 * library (flens_amd/libflx.so) never links or calls it.
 *
 * Compiled with -O2 -ffp-contract=off: the reference's generic path is bit
 * reproducible only without fused multiply-adds (the reference compiled
 * with -march=native -O3 contracts and changes its own low bits).
 *
 * Pinning (tests/test_oracle.py): bit-exact against oracle/_ref (the
 * reference compiled from its own headers) and the golden fixtures under
 * tests/golden/, the 4x4 known-answer example of
 * flens/examples/lapack-getrf.cc:43-46, exact-zero / tie / permutation /
 * diagonal special cases, and brute-force long-double LU for n <= 8.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- RNG --- */
/*
 * The build's own counter-based generator (not glibc rand(), which FLENS's
 * fillRandom uses at flens/auxiliary/randomvalue.tcc:43-45): element idx of
 * a stream with seed s is splitmix64(s + (idx+1)*golden), mapped exactly to
 * the dyadic grid k*2^-52 - 1 in [-1,1).  Identical in C, numpy and HIP.
 */
static inline uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_fill_uniform(double *a, int m, int n, int lda, uint64_t seed,
                      uint64_t offset)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            uint64_t idx = offset + (uint64_t)j * (uint64_t)m + (uint64_t)i;
            uint64_t z = orc_mix64(seed + (idx + 1) * 0x9E3779B97F4A7C15ULL);
            int64_t k = (int64_t)(z >> 11) - ((int64_t)1 << 52);
            a[(size_t)i + (size_t)j * lda] = (double)k * 0x1p-52;
        }
}

/* ---------------------------------------------------------- LU bodies --- */
#define REAL double
#define FN(x) orc_d##x
#define RABS fabs
#define RSAFEMIN DBL_MIN      /* flens/lapack/la/lamch.tcc:72-81 */
#include "lu_body.h"
#undef REAL
#undef FN
#undef RABS
#undef RSAFEMIN

#define REAL float
#define FN(x) orc_s##x
#define RABS fabsf
#define RSAFEMIN FLT_MIN
#include "lu_body.h"
#undef REAL
#undef FN
#undef RABS
#undef RSAFEMIN

/* --------------------------------------------------------------- GEMM --- */
/*
 * cxxblas/level3/gemm.tcc:149-299 (gemm_generic) for real data, all four
 * transpose combinations (0 = NoTrans, 1 = Trans).  The ColMajor call is
 * rewritten as RowMajor with A<->B swapped (gemm.tcc:166-173), C is scaled
 * by beta first (gescal, :175 -- so beta=0 computes 0*C and NaN/Inf in C
 * propagate), alpha==0 returns (:176-178), and every C(l,i) then receives
 *   s = sum_p op(A)(l,p)*op(B)(p,i)  (dotu_generic, dot.tcc:42-52)
 *   C(l,i) += alpha*s                (gemv.tcc:65-68 / :104-107).
 */
void orc_dgemm(int transa, int transb, int m, int n, int k, double alpha,
               const double *a, int lda, const double *b, int ldb,
               double beta, double *c, int ldc)
{
    if (m == 0 || n == 0) return;
    for (int i = 0; i < n; ++i)
        for (int l = 0; l < m; ++l) c[(size_t)l + (size_t)i * ldc] *= beta;
    if (alpha == 0.0) return;
    for (int l = 0; l < m; ++l) {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int p = 0; p < k; ++p) {
                double av = transa ? a[(size_t)p + (size_t)l * lda]
                                   : a[(size_t)l + (size_t)p * lda];
                double bv = transb ? b[(size_t)i + (size_t)p * ldb]
                                   : b[(size_t)p + (size_t)i * ldb];
                s += bv * av;
            }
            c[(size_t)l + (size_t)i * ldc] += alpha * s;
        }
    }
}

/* --------------------------------------------------------------- TRSM --- */
/*
 * cxxblas/level2/trsv.tcc:42-... (trsv_generic) for real data in RowMajor
 * storage (the order trsm_generic always reaches it in): each x_i first
 * subtracts the dot product of its already solved neighbours, formed by
 * dotu_generic (cxxblas/level1/dot.tcc:42-52: sum from 0, ascending), then
 * divides by the diagonal unless diag is Unit.  Conj / ConjTrans are NoTrans
 * / Trans for real data (dot_generic conjugates nothing there).
 */
static void orc_trsv_rm(int upper, int trans, int unit, int n, const double *a,
                        int lda, double *x, int incx)
{
#define AR_(i, j) a[(size_t)(i) * (size_t)lda + (size_t)(j)]
#define X_(i) x[(size_t)(i) * (size_t)incx]
    if (!trans && upper) {            /* backward, row i right of the diagonal */
        for (int i = n - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = 0; k < n - i - 1; ++k) s += AR_(i, i + 1 + k) * X_(i + 1 + k);
            X_(i) -= s;
            if (!unit) X_(i) /= AR_(i, i);
        }
    } else if (!trans) {              /* Lower: forward, row i left of the diagonal */
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < i; ++k) s += AR_(i, k) * X_(k);
            X_(i) -= s;
            if (!unit) X_(i) /= AR_(i, i);
        }
    } else if (upper) {               /* Trans, Upper: forward, column i above */
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int k = 0; k < i; ++k) s += AR_(k, i) * X_(k);
            X_(i) -= s;
            if (!unit) X_(i) /= AR_(i, i);
        }
    } else {                          /* Trans, Lower: backward, column i below */
        for (int i = n - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = 0; k < n - i - 1; ++k) s += AR_(i + 1 + k, i) * X_(i + 1 + k);
            X_(i) -= s;
            if (!unit) X_(i) /= AR_(i, i);
        }
    }
#undef AR_
#undef X_
}

/*
 * cxxblas/level3/trsm.tcc:40-68 (trsm_generic), the routine blas::sm
 * reaches (flens/blas/level3/sm.tcc:63).  order 0 RowMajor / 1 ColMajor,
 * right 0 Left / 1 Right, upper, trans, unit as 0/1.  A ColMajor call is
 * the RowMajor one with side and uplo flipped and m, n swapped; Right solves
 * every row of B with trsv(trans ^ Trans), Left every column (stride ldb)
 * with trsv(trans); only then is B scaled by alpha (gescal,
 * cxxblas/level1extensions/gescal.tcc:40-62, i.e. scal's b *= alpha).
 */
void orc_dtrsm(int order, int right, int upper, int trans, int unit, int m, int n,
               double alpha, const double *a, int lda, double *b, int ldb)
{
    if (order == 1) {
        right = !right;
        upper = !upper;
        int t = m; m = n; n = t;
    }
    if (right) {
        for (int i = 0; i < m; ++i)
            orc_trsv_rm(upper, !trans, unit, n, a, lda, b + (size_t)i * ldb, 1);
    } else {
        for (int j = 0; j < n; ++j)
            orc_trsv_rm(upper, trans, unit, m, a, lda, b + j, ldb);
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) b[(size_t)i * ldb + j] *= alpha;
}

/* Single-precision GEMM with the same order (used by the sgetrf restatement
 * only through orc_sgetrf; exported for completeness of the mixed oracle). */

/* ------------------------------------------------------- LU front-ends --- */
int orc_dgetrf_nb(int m, int n, double *a, int lda, int *ipiv, int nb)
{
    return orc_dgetrf(m, n, a, lda, ipiv, nb);
}

int orc_sgetrf_nb(int m, int n, float *a, int lda, int *ipiv, int nb)
{
    return orc_sgetrf(m, n, a, lda, ipiv, nb);
}

void orc_dgetrs_c(int trans, int n, int nrhs, const double *a, int lda,
                  const int *ipiv, double *b, int ldb)
{
    orc_dgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb);
}

void orc_sgetrs_c(int trans, int n, int nrhs, const float *a, int lda,
                  const int *ipiv, float *b, int ldb)
{
    orc_sgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb);
}

int orc_dgesv_nb(int n, int nrhs, double *a, int lda, int *ipiv, double *b,
                 int ldb, int nb)
{
    return orc_dgesv(n, nrhs, a, lda, ipiv, b, ldb, nb);
}

/* laswp front-end (flens/lapack/la/laswp.tcc:54-93; LAPACK dlaswp argument
 * convention: k1..k2 1-based, incx = +1 forward / -1 backward). */
void orc_dlaswp(int n, double *a, int lda, int k1, int k2, const int *ipiv,
                int incx)
{
    int *p0 = (int *)malloc(sizeof(int) * (size_t)(k2 + 1));
    for (int i = 0; i < k2; ++i) p0[i] = ipiv[i] - 1;
    if (incx > 0) orc_dlaswp_fwd(n, a, lda, k1 - 1, k2, p0);
    else          orc_dlaswp_bwd(n, a, lda, k1 - 1, k2, p0);
    free(p0);
}

/* ----------------------------------------------------------- checkers --- */
/*
 * LAPACK dget01-style ratio ||L*U - P*A||_1 / (n * ||A||_1 * eps) with
 * eps = 2^-53 (lamch(Eps), flens/lapack/la/lamch.tcc:69).  lu holds the
 * factors as returned by getrf, a0 the original matrix.  The product is
 * formed in long double so the checker adds no error of its own.
 */
double orc_dget01(int m, int n, const double *a0, int lda0, const double *lu,
                  int ldlu, const int *ipiv)
{
    const int mn = m < n ? m : n;
    long double *r = (long double *)calloc((size_t)m * n, sizeof(long double));
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double s = 0;
            int kmax = i < j ? i : j;
            if (kmax > mn - 1) kmax = mn - 1;
            for (int k = 0; k <= kmax; ++k) {
                long double l = (k == i) ? 1.0L : (long double)lu[i + (size_t)k * ldlu];
                s += l * (long double)lu[k + (size_t)j * ldlu];
            }
            r[i + (size_t)j * m] = s;
        }
    /* apply P^T: undo the interchanges, last to first (dlaswp incx=-1). */
    for (int i = mn - 1; i >= 0; --i) {
        int ip = ipiv[i] - 1;
        if (ip != i)
            for (int j = 0; j < n; ++j) {
                long double t = r[i + (size_t)j * m];
                r[i + (size_t)j * m] = r[ip + (size_t)j * m];
                r[ip + (size_t)j * m] = t;
            }
    }
    long double anorm = 0, rnorm = 0;
    for (int j = 0; j < n; ++j) {
        long double ca = 0, cr = 0;
        for (int i = 0; i < m; ++i) {
            ca += fabsl((long double)a0[i + (size_t)j * lda0]);
            cr += fabsl(r[i + (size_t)j * m] - (long double)a0[i + (size_t)j * lda0]);
        }
        if (ca > anorm) anorm = ca;
        if (cr > rnorm) rnorm = cr;
    }
    free(r);
    const double eps = DBL_EPSILON * 0.5;
    if (anorm == 0) return rnorm == 0 ? 0.0 : 1.0 / eps;
    return (double)(rnorm / ((long double)n * anorm * eps));
}

/*
 * LAPACK dget02-style ratio ||B - A*X||_1 / (||A||_1 * ||X||_1 * eps),
 * maximised over right-hand sides, in long double.
 */
double orc_dget02(int n, int nrhs, const double *a, int lda, const double *x,
                  int ldx, const double *b, int ldb)
{
    long double anorm = 0;
    for (int j = 0; j < n; ++j) {
        long double s = 0;
        for (int i = 0; i < n; ++i) s += fabsl((long double)a[i + (size_t)j * lda]);
        if (s > anorm) anorm = s;
    }
    const double eps = DBL_EPSILON * 0.5;
    double worst = 0;
    for (int c = 0; c < nrhs; ++c) {
        long double rn = 0, xn = 0;
        for (int i = 0; i < n; ++i) {
            long double s = (long double)b[i + (size_t)c * ldb];
            for (int k = 0; k < n; ++k)
                s -= (long double)a[i + (size_t)k * lda] * (long double)x[k + (size_t)c * ldx];
            rn += fabsl(s);
            xn += fabsl((long double)x[i + (size_t)c * ldx]);
        }
        double ratio;
        if (anorm == 0 || xn == 0) ratio = rn == 0 ? 0.0 : 1.0 / eps;
        else ratio = (double)(rn / anorm / xn / eps);
        if (ratio > worst) worst = ratio;
    }
    return worst;
}

/*
 * netlib dblat3 DMMCH criterion: max over entries of |C - Cexact| /
 * (eps * G) with G = |alpha| * sum_p |op(A)(i,p)||op(B)(p,j)| + |beta||C0|,
 * Cexact in long double.  c0 is C before the call, c the result.
 */
double orc_dmmch(int transa, int transb, int m, int n, int k, double alpha,
                 const double *a, int lda, const double *b, int ldb,
                 double beta, const double *c0, const double *c, int ldc)
{
    const double eps = DBL_EPSILON * 0.5;
    double worst = 0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double s = 0, g = 0;
            for (int p = 0; p < k; ++p) {
                long double av = transa ? a[p + (size_t)i * lda] : a[i + (size_t)p * lda];
                long double bv = transb ? b[j + (size_t)p * ldb] : b[p + (size_t)j * ldb];
                s += av * bv;
                g += fabsl(av * bv);
            }
            long double cz = c0[i + (size_t)j * ldc];
            long double ex = (long double)alpha * s + (long double)beta * cz;
            g = fabsl((long double)alpha) * g + fabsl((long double)beta) * fabsl(cz);
            long double err = fabsl((long double)c[i + (size_t)j * ldc] - ex);
            double r;
            if (g == 0) r = err == 0 ? 0.0 : 1.0 / eps;
            else r = (double)(err / (eps * g));
            if (r > worst) worst = r;
        }
    return worst;
}

/* -------------------------------------------------- mixed precision IR --- */
/*
 * netlib LAPACK 3.3.1 dsgesv.f (cxxlapack/netlib/lapack.tgz; wrapper
 * cxxlapack/interface/sgesv.tcc, declaration
 * cxxlapack/netlib/interface/lapack.in.h:10611-10623), restated:
 *   anrm = ||A||_inf, cte = anrm*eps*sqrt(n)*BWDMAX (BWDMAX = 1),
 *   SA = float(A), SX = float(B)  (overflow -> ITER = -2, fall back),
 *   SGETRF (info > 0 -> ITER = -3, fall back), SGETRS, X = double(SX),
 *   R = B - A*X (DGEMM), converged when for every rhs
 *   max|R(:,j)| <= max|X(:,j)| * cte;
 *   up to ITERMAX = 30 corrections X += double(SGETRS(float(R)));
 *   ITER = -31 when not converged; fallback = DGETRF + DGETRS in FP64.
 * The single-precision LU uses nb = 64 (ilaenv for SGETRF).
 * Returns info; *iter receives ITER.
 */
int orc_dsgesv(int n, int nrhs, double *a, int lda, int *ipiv,
               const double *b, int ldb, double *x, int ldx, int itermax,
               int *iter)
{
    const double eps = DBL_EPSILON * 0.5;
    double anrm = 0;
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += fabs(a[i + (size_t)j * lda]);
        if (s > anrm || isnan(s)) anrm = s;
    }
    const double cte = anrm * eps * sqrt((double)n) * 1.0;
    float *sa = (float *)malloc(sizeof(float) * (size_t)n * n);
    float *sx = (float *)malloc(sizeof(float) * (size_t)n * (nrhs > 0 ? nrhs : 1));
    double *r = (double *)malloc(sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1));
    const double rmax = (double)FLT_MAX;
    int ok = 1, info = 0;
    *iter = 0;
    if (n == 0 || nrhs == 0) { free(sa); free(sx); free(r); return 0; }
    for (int j = 0; j < nrhs && ok; ++j)
        for (int i = 0; i < n; ++i) {
            double v = b[i + (size_t)j * ldb];
            if (v < -rmax || v > rmax) { ok = 0; break; }
            sx[i + (size_t)j * n] = (float)v;
        }
    if (!ok) { *iter = -2; goto fallback; }
    for (int j = 0; j < n && ok; ++j)
        for (int i = 0; i < n; ++i) {
            double v = a[i + (size_t)j * lda];
            if (v < -rmax || v > rmax) { ok = 0; break; }
            sa[i + (size_t)j * n] = (float)v;
        }
    if (!ok) { *iter = -2; goto fallback; }
    if (orc_sgetrf(n, n, sa, n, ipiv, 64) != 0) { *iter = -3; goto fallback; }
    orc_sgetrs(0, n, nrhs, sa, n, ipiv, sx, n);
    for (int j = 0; j < nrhs; ++j)
        for (int i = 0; i < n; ++i) x[i + (size_t)j * ldx] = (double)sx[i + (size_t)j * n];
    for (int it = 0; it <= itermax; ++it) {
        /* R = B - A*X via the generic DGEMM order */
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i) r[i + (size_t)j * n] = b[i + (size_t)j * ldb];
        orc_dgemm(0, 0, n, nrhs, n, -1.0, a, lda, x, ldx, 1.0, r, n);
        int conv = 1;
        for (int j = 0; j < nrhs; ++j) {
            int ix = orc_diamax(n, x + (size_t)j * ldx, 1);
            int ir = orc_diamax(n, r + (size_t)j * n, 1);
            double xn = fabs(x[ix + (size_t)j * ldx]);
            double rn = fabs(r[ir + (size_t)j * n]);
            if (rn > xn * cte) { conv = 0; break; }
        }
        if (conv) { *iter = it; free(sa); free(sx); free(r); return 0; }
        if (it == itermax) break;
        ok = 1;
        for (int j = 0; j < nrhs && ok; ++j)
            for (int i = 0; i < n; ++i) {
                double v = r[i + (size_t)j * n];
                if (v < -rmax || v > rmax) { ok = 0; break; }
                sx[i + (size_t)j * n] = (float)v;
            }
        if (!ok) { *iter = -2; goto fallback; }
        orc_sgetrs(0, n, nrhs, sa, n, ipiv, sx, n);
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i)
                x[i + (size_t)j * ldx] += (double)sx[i + (size_t)j * n];
    }
    *iter = -itermax - 1;
fallback:
    for (int j = 0; j < nrhs; ++j)
        for (int i = 0; i < n; ++i) x[i + (size_t)j * ldx] = b[i + (size_t)j * ldb];
    info = orc_dgetrf(n, n, a, lda, ipiv, 64);
    if (info == 0) orc_dgetrs(0, n, nrhs, a, lda, ipiv, x, ldx);
    free(sa); free(sx); free(r);
    return info;
}

/* Pivot-search diagnostics: the blocked getrf with, per column j, the
 * relative gap (|pivot| - |runner-up|) / |pivot| among the candidates the
 * reference's iamax compared.  A tiny gap marks a near-tie where another
 * summation order may legitimately pick a different row. */
int orc_dgetrf_gaps(int m, int n, double *a, int lda, int *ipiv, int nb,
                    double *gap)
{
    return orc_dgetrf_gap(m, n, a, lda, ipiv, nb, gap);
}

/* norms, equilibration, condition estimate, refinement, expert driver */
#include "aux_body.h"
