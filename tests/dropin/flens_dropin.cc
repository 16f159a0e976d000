// A FLENS program that uses only FLENS's public API (GeMatrix, blas::mm,
// blas::sm, lapack::trf / trs / sv).  Compiled with the reference's own switches for
// an external BLAS/LAPACK (-DWITH_REFBLAS -DUSE_CXXLAPACK
// -DALWAYS_USE_CXXLAPACK) and linked against libflx instead of a CPU
// BLAS/LAPACK, its GEMM, triangular-solve and LU calls land in libflx's
// cblas_dgemm, cblas_dtrsm, dgetrf_, dgetrs_ and dgesv_ (flens/lapack/lapack.h:48-55,
// flens/lapack/ge/trf.tcc:164-173, cxxblas/level3/gemm.tcc:327-349).
//
// Prints one line per check, "ok <name> <value>" or "FAIL <name> <value>".
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <flens/flens.cxx>

using namespace flens;

// FLENS with HAVE_CBLAS also sends level-1 copies (matrix assignment) to the
// CBLAS.  libflx replaces only the GEMM / LU path, so a program links it
// before its CPU BLAS; this stand-alone program has no CPU BLAS and supplies
// the one level-1 routine it needs (cxxblas/drivers/cblas.h, cblas_dcopy).
extern "C" void cblas_dcopy(int n, const double *x, int incx, double *y, int incy)
{
    if (n <= 0) return;
    long ix = incx < 0 ? long(1 - n) * incx : 0, iy = incy < 0 ? long(1 - n) * incy : 0;
    for (int i = 0; i < n; ++i, ix += incx, iy += incy) y[iy] = x[ix];
}

typedef GeMatrix<FullStorage<double, ColMajor> > DMatrix;
typedef DenseVector<Array<int> > IntVector;

static uint64_t mix(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void fill(DMatrix &A, uint64_t seed)
{
    for (int j = 1; j <= A.numCols(); ++j)
        for (int i = 1; i <= A.numRows(); ++i)
            A(i, j) = double(mix(seed * 1000003ull + uint64_t(j) * 65537ull + i) >> 11)
                      * std::ldexp(1.0, -52) - 1.0;
}

static int report(const char *name, bool ok, double v)
{
    std::printf("%s %s %.6e\n", ok ? "ok" : "FAIL", name, v);
    return ok ? 0 : 1;
}

int main()
{
    int bad = 0;
    // 4x4 example of flens/examples/lapack-getrf.cc: piv = (2,3,4,4), x = (1,9,9,9)
    {
        DMatrix Ab(4, 5);
        Ab = 2, 3, -1, 0, 20,
            -6, -5, 0, 2, -33,
             2, -5, 6, -6, -43,
             4, 6, 2, -3, 49;
        IntVector piv(4);
        int info = lapack::trf(Ab, piv);
        bool ok = info == 0 && piv(1) == 2 && piv(2) == 3 && piv(3) == 4 && piv(4) == 4;
        bad += report("kat_trf_pivots", ok, info);
        // the example's second half: blas::sm(Left, NoTrans, 1, A.upper(), B)
        // (flens/examples/lapack-getrf.cc:56-61) -> cxxblas::trsm -> cblas_dtrsm
        const Underscore<int> _;
        const auto A = Ab(_, _(1, 4));
        auto B = Ab(_, _(5, 5));
        blas::sm(Left, NoTrans, 1.0, A.upper(), B);
        double dx = std::fabs(B(1, 1) - 1) + std::fabs(B(2, 1) - 9) + std::fabs(B(3, 1) - 9) +
                    std::fabs(B(4, 1) - 9);
        bad += report("kat_sm_solution", dx < 1e-12, dx);
    }
    // blas::mm: C = 1.5 A B - 0.5 C against a scalar loop
    {
        const int m = 300, n = 200, k = 250;
        DMatrix A(m, k), B(k, n), C(m, n), C0(m, n);
        fill(A, 1); fill(B, 2); fill(C, 3);
        C0 = C;
        blas::mm(NoTrans, NoTrans, 1.5, A, B, -0.5, C);
        double err = 0, scale = 0;
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= m; ++i) {
                double s = 0, a = 0;
                for (int p = 1; p <= k; ++p) { s += A(i, p) * B(p, j); a += std::fabs(A(i, p) * B(p, j)); }
                double want = 1.5 * s - 0.5 * C0(i, j);
                err = std::fmax(err, std::fabs(C(i, j) - want));
                scale = std::fmax(scale, 1.5 * a + 0.5 * std::fabs(C0(i, j)));
            }
        double ratio = err / (4.0 * k * std::ldexp(1.0, -53) * scale);
        bad += report("mm_bound_ratio", ratio <= 1.0, ratio);
    }
    // lapack::sv, then trf + trs (NoTrans and Trans): scaled residuals
    {
        const int n = 700, nrhs = 3;
        DMatrix A(n, n), A0(n, n), B(n, nrhs), B0(n, nrhs);
        fill(A, 4); fill(B, 5);
        A0 = A; B0 = B;
        IntVector piv(n);
        int info = lapack::sv(A, piv, B);
        double anorm = 0, xnorm = 0, rnorm = 0;
        for (int j = 1; j <= n; ++j) {
            double s = 0;
            for (int i = 1; i <= n; ++i) s += std::fabs(A0(i, j));
            anorm = std::fmax(anorm, s);
        }
        for (int r = 1; r <= nrhs; ++r) {
            double rs = 0, xs = 0;
            for (int i = 1; i <= n; ++i) {
                double t = B0(i, r);
                for (int j = 1; j <= n; ++j) t -= A0(i, j) * B(j, r);
                rs += std::fabs(t); xs += std::fabs(B(i, r));
            }
            rnorm = std::fmax(rnorm, rs); xnorm = std::fmax(xnorm, xs);
        }
        double ratio = rnorm / (anorm * xnorm * n * std::ldexp(1.0, -53));
        bad += report("sv_dget02_ratio", info == 0 && ratio < 30, ratio);

        DMatrix LU(A0);
        IntVector p2(n);
        info = lapack::trf(LU, p2);
        DMatrix X(B0);
        lapack::trs(Trans, LU, p2, X);
        double r2 = 0, x2 = 0;
        for (int r = 1; r <= nrhs; ++r) {
            double rs = 0, xs = 0;
            for (int i = 1; i <= n; ++i) {
                double t = B0(i, r);
                for (int j = 1; j <= n; ++j) t -= A0(j, i) * X(j, r);
                rs += std::fabs(t); xs += std::fabs(X(i, r));
            }
            r2 = std::fmax(r2, rs); x2 = std::fmax(x2, xs);
        }
        double ratio2 = r2 / (anorm * x2 * n * std::ldexp(1.0, -53));
        bad += report("trs_trans_dget02_ratio", info == 0 && ratio2 < 30, ratio2);
        bool same = true;
        for (int i = 1; i <= n; ++i) same = same && (piv(i) == p2(i));
        bad += report("sv_trf_same_pivots", same, 0.0);
    }
    std::printf("%s\n", bad ? "DROPIN FAILED" : "DROPIN PASSED");
    return bad ? 1 : 0;
}
