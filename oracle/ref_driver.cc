// ref_driver.cc -- TEST INFRASTRUCTURE ONLY.
//
// A thin C entry layer over the reference's own generic FLENS path, compiled
// directly from the reference's headers (g++ -I<reference root>) into
// oracle/_ref/libflensref.so by oracle/Makefile.  Nothing of the reference is
// copied into this repository; this file only calls it.  It exists so the
// C restatement in flens_oracle.c can be checked bit for bit against the
// reference itself, and so golden fixtures can be generated from it
// (tests/golden/make_golden.py).  The product library never links it.
//
// The GETRF block size is swept through the reference's own ilaenv hook
// (flens/lapack/la/ilaenv.tcc:453-509: USE_ILAENV_WITH_UNDERSCORE plus a
// caller-supplied ilaenv_), the same hook the netlib test suite uses.

#define USE_ILAENV_WITH_UNDERSCORE 1
#define USE_NATIVE_ILAENV 1
#define INTEGER int
#include <flens/flens.cxx>

static int g_getrf_nb = 64;

extern "C" int ilaenv_(const int *spec, const char *, const char *,
                       const int *, const int *, const int *, const int *,
                       int, int)
{
    if (*spec == 1) return g_getrf_nb;   // optimal block size (GETRF)
    if (*spec == 2) return 2;            // minimal block size
    return 1;                            // crossover
}

using namespace flens;

typedef FullStorageView<double, ColMajor>       DView;
typedef ConstFullStorageView<double, ColMajor>  DConstView;
typedef GeMatrix<DView>                          DGeView;
typedef GeMatrix<DConstView>                     DGeConstView;
typedef DenseVector<ArrayView<int> >             IView;

extern "C" {

// flens::blas::mm (flens/blas/level3/mm.tcc:380-479) -> cxxblas::gemm
// generic (cxxblas/level3/gemm.tcc:149-299).  trans: 0 NoTrans, 1 Trans.
void ref_dgemm(int transa, int transb, int m, int n, int k, double alpha,
               const double *a, int lda, const double *b, int ldb,
               double beta, double *c, int ldc)
{
    int ar = transa ? k : m, ac = transa ? m : k;
    int br = transb ? n : k, bc = transb ? k : n;
    DGeConstView A = DConstView(ar, ac, a, lda);
    DGeConstView B = DConstView(br, bc, b, ldb);
    DGeView      C = DView(m, n, c, ldc);
    blas::mm(transa ? Trans : NoTrans, transb ? Trans : NoTrans,
             alpha, A, B, beta, C);
}

// flens::lapack::trf (flens/lapack/ge/trf.tcc:183-261); ipiv 1-based.
int ref_dgetrf(int m, int n, double *a, int lda, int *ipiv, int nb)
{
    g_getrf_nb = nb;
    DGeView A = DView(m, n, a, lda);
    IView piv = ArrayView<int>(std::min(m, n), ipiv, 1);
    return lapack::trf(A, piv);
}

// flens::lapack::trs (flens/lapack/ge/trs.tcc:151-252).
void ref_dgetrs(int trans, int n, int nrhs, const double *a, int lda,
                const int *ipiv, double *b, int ldb)
{
    DGeConstView A = DConstView(n, n, a, lda);
    IView piv = ArrayView<int>(n, const_cast<int *>(ipiv), 1);
    DGeView B = DView(n, nrhs, b, ldb);
    lapack::trs(trans ? Trans : NoTrans, A, piv, B);
}

// flens::lapack::sv (flens/lapack/ge/sv.tcc:115-219).
int ref_dgesv(int n, int nrhs, double *a, int lda, int *ipiv, double *b,
              int ldb, int nb)
{
    g_getrf_nb = nb;
    DGeView A = DView(n, n, a, lda);
    IView piv = ArrayView<int>(n, ipiv, 1);
    DGeView B = DView(n, nrhs, b, ldb);
    return lapack::sv(A, piv, B);
}

// flens::lapack::laswp (flens/lapack/la/laswp.tcc:126-186) over rows
// k1..k2 (1-based), forward for incx>0, piv.reverse() for incx<0.
void ref_dlaswp(int m, int n, double *a, int lda, int k1, int k2,
                const int *ipiv, int incx)
{
    DGeView A = DView(m, n, a, lda);
    int len = k2 - k1 + 1;
    IView piv = ArrayView<int>(len, const_cast<int *>(ipiv) + (k1 - 1), 1, k1);
    if (incx > 0) {
        lapack::laswp(A, piv);
    } else {
        lapack::laswp(A, piv.reverse());
    }
}

// cxxblas::trsm (cxxblas/level3/trsm.tcc:40-82, generic: no CBLAS here), the
// routine blas::sm calls (flens/blas/level3/sm.tcc:63).  order 0 RowMajor /
// 1 ColMajor; right, upper, trans, unit as 0/1.
void ref_dtrsm(int order, int right, int upper, int trans, int unit, int m, int n,
               double alpha, const double *a, int lda, double *b, int ldb)
{
    cxxblas::trsm(order ? cxxblas::ColMajor : cxxblas::RowMajor,
                  right ? cxxblas::Right : cxxblas::Left,
                  upper ? cxxblas::Upper : cxxblas::Lower,
                  trans ? cxxblas::Trans : cxxblas::NoTrans,
                  unit ? cxxblas::Unit : cxxblas::NonUnit,
                  m, n, alpha, a, lda, b, ldb);
}

// ---- FLENS-LAPACK auxiliaries: norms, equilibration, condition estimate,
// refinement, expert driver (generic implementations) --------------------

static lapack::Norm ref_norm(char c)
{
    switch (c) {
    case 'M': case 'm': return lapack::MaximumNorm;
    case 'O': case 'o': case '1': return lapack::OneNorm;
    case 'I': case 'i': return lapack::InfinityNorm;
    default: return lapack::FrobeniusNorm;
    }
}

typedef DenseVector<ArrayView<double> > DVecView;
typedef DenseVector<Array<double> >     DVec;
typedef DenseVector<Array<int> >        IVec;

// flens::lapack::lan, GeMatrix variant (flens/lapack/la/lan.tcc:55-130)
double ref_dlange(char norm, int m, int n, const double *a, int lda)
{
    DGeConstView A = DConstView(m, n, a, std::max(lda, 1));
    DVec work(std::max(m, 1));
    return lapack::lan(ref_norm(norm), A, work);
}

// flens::lapack::equ (flens/lapack/ge/equ.tcc:62-194)
int ref_dgeequ(int m, int n, const double *a, int lda, double *r, double *c,
               double *rowcnd, double *colcnd, double *amax)
{
    DGeConstView A = DConstView(m, n, a, std::max(lda, 1));
    DVecView R = ArrayView<double>(m, r, 1);
    DVecView Cv = ArrayView<double>(n, c, 1);
    return lapack::equ(A, R, Cv, *rowcnd, *colcnd, *amax);
}

// flens::lapack::laq (flens/lapack/la/laq.tcc:57-118); returns 'N','R','C','B'
int ref_dlaqge(int m, int n, double *a, int lda, const double *r, const double *c,
               double rowcnd, double colcnd, double amax)
{
    DGeView A = DView(m, n, a, std::max(lda, 1));
    DVecView R = ArrayView<double>(m, const_cast<double *>(r), 1);
    DVecView Cv = ArrayView<double>(n, const_cast<double *>(c), 1);
    return (int)lapack::laq(A, R, Cv, rowcnd, colcnd, amax);
}

// flens::lapack::con, real variant (flens/lapack/ge/cond.tcc:62-165) on an
// LU factor as trf leaves it
double ref_dgecon(char norm, int n, const double *lu, int lda, double anorm)
{
    DGeConstView A = DConstView(n, n, lu, std::max(lda, 1));
    DVec work(4 * std::max(n, 1));
    IVec iwork(std::max(n, 1));
    double rcond = 0;
    lapack::con(ref_norm(norm), A, anorm, rcond, work, iwork);
    return rcond;
}

// flens::lapack::rfs (flens/lapack/ge/rfs.tcc:63-261)
void ref_dgerfs(int trans, int n, int nrhs, const double *a, int lda, const double *af,
                int ldaf, const int *ipiv, const double *b, int ldb, double *x, int ldx,
                double *ferr, double *berr)
{
    DGeConstView A = DConstView(n, n, a, std::max(lda, 1));
    DGeConstView AF = DConstView(n, n, af, std::max(ldaf, 1));
    IView piv = ArrayView<int>(n, const_cast<int *>(ipiv), 1);
    DGeConstView B = DConstView(n, nrhs, b, std::max(ldb, 1));
    DGeView X = DView(n, nrhs, x, std::max(ldx, 1));
    DVecView F = ArrayView<double>(nrhs, ferr, 1);
    DVecView Bk = ArrayView<double>(nrhs, berr, 1);
    DVec work(3 * std::max(n, 1));
    IVec iwork(std::max(n, 1));
    lapack::rfs(trans ? Trans : NoTrans, A, AF, piv, B, X, F, Bk, work, iwork);
}

// flens::lapack::svx (flens/lapack/ge/svx.tcc:60-262).  fact: 'N','E','F';
// equed is an input only (the reference takes it by value); work0 receives
// the reciprocal pivot growth (work(1)).
int ref_dgesvx(char fact, int trans, int n, int nrhs, double *a, int lda, double *af,
               int ldaf, int *ipiv, char equed, double *r, double *c, double *b, int ldb,
               double *x, int ldx, double *rcond, double *ferr, double *berr, double *work0,
               int nb)
{
    g_getrf_nb = nb;
    DGeView A = DView(n, n, a, std::max(lda, 1));
    DGeView AF = DView(n, n, af, std::max(ldaf, 1));
    IView piv = ArrayView<int>(n, ipiv, 1);
    DVecView R = ArrayView<double>(n, r, 1);
    DVecView Cv = ArrayView<double>(n, c, 1);
    DGeView B = DView(n, nrhs, b, std::max(ldb, 1));
    DGeView X = DView(n, nrhs, x, std::max(ldx, 1));
    DVecView F = ArrayView<double>(nrhs, ferr, 1);
    DVecView Bk = ArrayView<double>(nrhs, berr, 1);
    DVec work(4 * std::max(n, 1));
    IVec iwork(std::max(n, 1));
    const int info = lapack::svx(lapack::SVX::Fact(fact), trans ? Trans : NoTrans, A, AF, piv,
                                 lapack::SVX::Equilibration(equed), R, Cv, B, X, *rcond, F, Bk, work,
                                 iwork);
    *work0 = work(1);
    return info;
}

// flens::lapack::tri for GeMatrix (flens/lapack/ge/tri.tcc:64-187)
int ref_dgetri(int n, double *a, int lda, const int *ipiv)
{
    DGeView A = DView(n, n, a, std::max(lda, 1));
    IView piv = ArrayView<int>(n, const_cast<int *>(ipiv), 1);
    DVec work;
    return lapack::tri(A, piv, work);
}

// flens::save / flens::load (flens/io/fullstorage/save.tcc:49-82,
// load.tcc:52-80; flens/io/array/save.tcc:43-66)
int ref_save_gematrix(const char *path, int m, int n, const double *a, int lda)
{
    DGeConstView A = DConstView(m, n, a, std::max(lda, 1));
    return flens::save(std::string(path), A) ? 0 : 1;
}

int ref_load_gematrix(const char *path, int *m, int *n, double *out, int cap)
{
    GeMatrix<FullStorage<double, ColMajor> > A;
    if (!flens::load(std::string(path), A)) return 1;
    *m = A.numRows();
    *n = A.numCols();
    if ((long)A.numRows() * A.numCols() > cap) return 2;
    for (int j = 1; j <= A.numCols(); ++j)
        for (int i = 1; i <= A.numRows(); ++i) out[(i - 1) + (long)(j - 1) * A.numRows()] = A(i, j);
    return 0;
}

int ref_save_ivector(const char *path, int n, const int *x)
{
    IView v = ArrayView<int>(n, const_cast<int *>(x), 1);
    return flens::save(std::string(path), v) ? 0 : 1;
}

} // extern "C"
