// netlib_checks.cc -- TEST PROGRAM: the netlib BLAS-3 GEMM test (dblat3's
// DCHK1 grid with its DMMCH check) and the LAPACK LU test (dchkge with its
// dget01/02/03/04/06/07 ratios over DLATMS-style matrix types), restated
// over libflx's Fortran seams dgemm_, dgetrf_, dgetrs_, dgecon_, dgerfs_,
// dgetri_, plus the error-exit checks of dblat3's DCHKE / LAPACK's derrge
// through xerbla_.  The reference runs these suites through its Fortran
// interface (netlib thresholds: 16 for dblat3, 30 for dchkge);
// here they run on the device behind the same symbols.
//
// Exit status 0 when every ratio is below its threshold; one line per
// group on stdout.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

extern "C" {
void dgemm_(const char *, const char *, const int *, const int *, const int *, const double *,
            const double *, const int *, const double *, const int *, const double *, double *,
            const int *);
void dgetrf_(const int *, const int *, double *, const int *, int *, int *);
void dgetrs_(const char *, const int *, const int *, const double *, const int *, const int *,
             double *, const int *, int *);
void dgecon_(const char *, const int *, const double *, const int *, const double *, double *,
             double *, int *, int *);
void dgerfs_(const char *, const int *, const int *, const double *, const int *, const double *,
             const int *, const int *, const double *, const int *, double *, const int *,
             double *, double *, double *, int *, int *);
void dgetri_(const int *, double *, const int *, const int *, double *, const int *, int *);
double dlange_(const char *, const int *, const int *, const double *, const int *, double *);

// error hook: records the last call (overrides libflx's weak default)
static std::string g_sr;
static int g_info = 0, g_calls = 0;
void xerbla_(const char *sr, const int *info, int len)
{
    g_sr.assign(sr, (size_t)len);
    while (!g_sr.empty() && g_sr.back() == ' ') g_sr.pop_back();
    g_info = *info;
    ++g_calls;
}
}

static const double EPS = 1.1102230246251565e-16;   // 2^-53 (dlamch('E'))
static int g_fail = 0;

static void report(const char *group, int tests, double worst, double thresh)
{
    const bool ok = worst < thresh;
    if (!ok) ++g_fail;
    printf("%-28s tests %6d  worst ratio %10.3g  threshold %5.1f  %s\n", group, tests, worst,
           thresh, ok ? "PASS" : "FAIL");
}

// ------------------------------------------------------------ dblat3 ----
// DMMCH (dblat3.f): expected CT = alpha op(A) op(B) + beta C in long double,
// gauge G = |alpha| |op(A)| |op(B)| + |beta| |C|; ratio = max |CT - CC| / (eps G)
static double dmmch(char ta, char tb, int m, int n, int k, double alpha, const double *a,
                    int lda, const double *b, int ldb, double beta, const double *c0, int ldc,
                    const double *cc)
{
    double worst = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            long double ct = 0.0L, g = 0.0L;
            for (int l = 0; l < k; ++l) {
                const double av = (ta == 'N') ? a[i + (size_t)l * lda] : a[l + (size_t)i * lda];
                const double bv = (tb == 'N') ? b[l + (size_t)j * ldb] : b[j + (size_t)l * ldb];
                ct += (long double)av * bv;
                g += fabsl((long double)av * bv);
            }
            ct = alpha * ct + (long double)beta * c0[i + (size_t)j * ldc];
            g = fabsl((long double)alpha) * g + fabsl((long double)beta * c0[i + (size_t)j * ldc]);
            const long double err = fabsl(ct - (long double)cc[i + (size_t)j * ldc]);
            double r = 0.0;
            if (g > 0.0L) r = (double)(err / (EPS * g));
            else if (err > 0.0L) r = 1e30;
            if (r > worst) worst = r;
        }
    return worst;
}

static void gemm_grid(std::mt19937_64 &rng)
{
    std::uniform_real_distribution<double> U(-0.5, 0.5);
    const int dims[] = {0, 1, 2, 3, 5, 9, 35};
    const double alphas[] = {0.0, 1.0, 0.7};
    const double betas[] = {0.0, 1.0, 1.3};
    const char tr[] = {'N', 'T', 'C'};
    int tests = 0;
    double worst = 0.0;
    bool untouched = true;
    for (int m : dims)
        for (int n : dims)
            for (int k : dims)
                for (char ta : tr)
                    for (char tb : tr) {
                        const int ar = ta == 'N' ? m : k, ac = ta == 'N' ? k : m;
                        const int br = tb == 'N' ? k : n, bc = tb == 'N' ? n : k;
                        // dblat3 pads the leading dimensions by one
                        const int lda = (ar > 1 ? ar : 1) + 1, ldb = (br > 1 ? br : 1) + 1;
                        const int ldc = (m > 1 ? m : 1) + 1;
                        std::vector<double> a((size_t)lda * (ac > 0 ? ac : 1)),
                            b((size_t)ldb * (bc > 0 ? bc : 1)), c((size_t)ldc * (n > 0 ? n : 1));
                        for (auto &x : a) x = U(rng);
                        for (auto &x : b) x = U(rng);
                        for (double alpha : alphas)
                            for (double beta : betas) {
                                for (auto &x : c) x = U(rng);
                                const std::vector<double> a0 = a, b0 = b, c0 = c;
                                dgemm_(&ta, &tb, &m, &n, &k, &alpha, a.data(), &lda, b.data(), &ldb,
                                       &beta, c.data(), &ldc);
                                ++tests;
                                untouched &= (a == a0) && (b == b0);
                                for (int j = 0; j < n; ++j)   // padding rows of C untouched
                                    for (int i = m; i < ldc; ++i)
                                        untouched &= c[i + (size_t)j * ldc] == c0[i + (size_t)j * ldc];
                                const double r = dmmch(ta, tb, m, n, k, alpha, a.data(), lda, b.data(),
                                                       ldb, beta, c0.data(), ldc, c.data());
                                if (r > worst) worst = r;
                            }
                    }
    report("dgemm_ (dblat3 DCHK1 grid)", tests, worst, 16.0);
    report("dgemm_ inputs unchanged", tests, untouched ? 0.0 : 1e30, 1.0);
}

static void gemm_error_exits()
{
    // dblat3 DCHKE: each illegal argument raises xerbla with its position
    double a[4] = {0}, b[4] = {0}, c[4] = {0}, one = 1.0;
    int two = 2, one_i = 1, neg = -1, zero = 0;
    struct Case { char ta, tb; int *m, *n, *k, *lda, *ldb, *ldc; int want; };
    const Case cases[] = {
        {'/', 'N', &zero, &zero, &zero, &one_i, &one_i, &one_i, 1},
        {'N', '/', &zero, &zero, &zero, &one_i, &one_i, &one_i, 2},
        {'N', 'N', &neg, &zero, &zero, &one_i, &one_i, &one_i, 3},
        {'N', 'N', &zero, &neg, &zero, &one_i, &one_i, &one_i, 4},
        {'N', 'N', &zero, &zero, &neg, &one_i, &one_i, &one_i, 5},
        {'N', 'N', &two, &zero, &zero, &one_i, &one_i, &two, 8},
        {'N', 'N', &zero, &zero, &two, &one_i, &one_i, &one_i, 10},
        {'N', 'N', &two, &zero, &zero, &two, &one_i, &one_i, 13},
        {'T', 'T', &zero, &zero, &two, &one_i, &one_i, &one_i, 8},
        {'T', 'T', &zero, &two, &zero, &one_i, &one_i, &one_i, 10},
    };
    int bad = 0, tests = 0;
    for (const Case &q : cases) {
        g_sr.clear();
        g_info = 0;
        dgemm_(&q.ta, &q.tb, q.m, q.n, q.k, &one, a, q.lda, b, q.ldb, &one, c, q.ldc);
        ++tests;
        if (g_sr != "DGEMM" || g_info != q.want) {
            printf("  dgemm_ error exit: want %d got %s %d\n", q.want, g_sr.c_str(), g_info);
            ++bad;
        }
    }
    // dgetrf_ / dgetrs_ (derrge)
    int info = 0, piv[2];
    g_sr.clear();
    dgetrf_(&neg, &zero, a, &one_i, piv, &info);
    tests += 3;
    if (!(g_sr == "DGETRF" && g_info == 1 && info == -1)) ++bad;
    dgetrf_(&two, &one_i, a, &one_i, piv, &info);
    if (!(g_sr == "DGETRF" && g_info == 4 && info == -4)) ++bad;
    const char badt = '/';
    dgetrs_(&badt, &zero, &zero, a, &one_i, piv, b, &one_i, &info);
    if (!(g_sr == "DGETRS" && g_info == 1 && info == -1)) ++bad;
    report("error exits (DCHKE / derrge)", tests, bad ? 1e30 : 0.0, 1.0);
}

// ------------------------------------------------------------ dchkge ----
typedef std::vector<double> Mat;   // column-major, ld = rows

static double norm1(int m, int n, const Mat &a)
{
    double v = 0.0;
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += fabs(a[i + (size_t)j * m]);
        v = s > v ? s : v;
    }
    return v;
}

// random orthogonal n x n (Gram-Schmidt twice on a Gaussian matrix)
static Mat orthogonal(int n, std::mt19937_64 &rng)
{
    std::normal_distribution<double> N(0.0, 1.0);
    Mat q((size_t)n * n);
    for (auto &x : q) x = N(rng);
    for (int j = 0; j < n; ++j)
        for (int pass = 0; pass < 2; ++pass) {
            for (int k = 0; k < j; ++k) {
                double d = 0.0;
                for (int i = 0; i < n; ++i) d += q[i + (size_t)k * n] * q[i + (size_t)j * n];
                for (int i = 0; i < n; ++i) q[i + (size_t)j * n] -= d * q[i + (size_t)k * n];
            }
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += q[i + (size_t)j * n] * q[i + (size_t)j * n];
            s = sqrt(s);
            for (int i = 0; i < n; ++i) q[i + (size_t)j * n] /= s;
        }
    return q;
}

// DLATMS-style test matrix of dchkge type 1..11 (m x n)
static Mat make_matrix(int type, int m, int n, std::mt19937_64 &rng, int *zero_from)
{
    const int mn = m < n ? m : n;
    const double badc1 = sqrt(0.1 / EPS), badc2 = 0.1 / EPS;
    const double small = 2.2250738585072014e-308 / EPS * 1e10, large = 1.0 / small;
    double cond = 2.0, scale = 1.0;
    if (type == 8) cond = badc1;
    if (type == 9) cond = badc2;
    if (type == 10) scale = small;
    if (type == 11) scale = large;
    std::vector<double> sig(mn > 0 ? mn : 1);
    for (int i = 0; i < mn; ++i) sig[i] = mn > 1 ? pow(cond, -(double)i / (mn - 1)) : 1.0;
    Mat a((size_t)m * n, 0.0);
    if (type == 1 || type == 2 || type == 3) {
        std::uniform_real_distribution<double> U(-1.0, 1.0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) {
                const bool keep = (type == 1) ? i == j : (type == 2) ? i <= j : i >= j;
                if (!keep) continue;
                a[i + (size_t)j * m] = (i == j) ? sig[i] * (U(rng) < 0 ? -1.0 : 1.0)
                                                : 0.5 * U(rng) / n;
            }
    } else {
        const Mat u = orthogonal(m, rng), v = orthogonal(n, rng);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) {
                long double s = 0.0L;
                for (int k = 0; k < mn; ++k) s += (long double)u[i + (size_t)k * m] * sig[k] * v[j + (size_t)k * n];
                a[i + (size_t)j * m] = (double)s * scale;
            }
    }
    *zero_from = 0;
    if (type >= 5 && type <= 7 && n > 0) {
        int c0 = type == 5 ? 0 : type == 6 ? n - 1 : n / 2;
        int c1 = type == 7 ? n : c0 + 1;
        for (int j = c0; j < c1; ++j)
            for (int i = 0; i < m; ++i) a[i + (size_t)j * m] = 0.0;
        *zero_from = c0 + 1;
    }
    return a;
}

static void lu_suite(std::mt19937_64 &rng)
{
    const int dims[] = {0, 1, 2, 3, 5, 10, 50};
    const int nrhss[] = {1, 2, 15};
    double w01 = 0, w02 = 0, w03 = 0, w04 = 0, w06 = 0, w07 = 0;
    int t01 = 0, t02 = 0, t03 = 0, t04 = 0, t06 = 0, t07 = 0, bad_info = 0;
    for (int m : dims)
        for (int n : dims)
            for (int type = 1; type <= 11; ++type) {
                if ((type == 2 || type == 3 || type >= 8) && m != n) continue;
                int zf = 0;
                const Mat a = make_matrix(type, m, n, rng, &zf);
                const int lda = m > 1 ? m : 1;
                Mat lu = a;
                std::vector<int> piv((m < n ? m : n) > 0 ? (m < n ? m : n) : 1);
                int info = 0;
                dgetrf_(&m, &n, lu.data(), &lda, piv.data(), &info);
                const int mn = m < n ? m : n;
                if (zf > 0 && zf <= mn && info <= 0) ++bad_info;   // a zero column must show
                // dget01: || P L U - A ||_1 / (n ||A||_1 eps)
                if (mn > 0) {
                    Mat l((size_t)m * mn, 0.0), u((size_t)mn * n, 0.0);
                    for (int j = 0; j < mn; ++j)
                        for (int i = 0; i < m; ++i)
                            l[i + (size_t)j * m] = i == j ? 1.0 : (i > j ? lu[i + (size_t)j * m] : 0.0);
                    for (int j = 0; j < n; ++j)
                        for (int i = 0; i < mn && i <= j; ++i) u[i + (size_t)j * mn] = lu[i + (size_t)j * m];
                    Mat pa = a;   // rows of A interchanged as the pivots say
                    for (int i = 0; i < mn; ++i) {
                        const int p = piv[i] - 1;
                        if (p != i)
                            for (int j = 0; j < n; ++j) std::swap(pa[i + (size_t)j * m], pa[p + (size_t)j * m]);
                    }
                    Mat r((size_t)m * n);
                    for (int j = 0; j < n; ++j)
                        for (int i = 0; i < m; ++i) {
                            long double s = 0.0L;
                            for (int k = 0; k < mn; ++k) s += (long double)l[i + (size_t)k * m] * u[k + (size_t)j * mn];
                            r[i + (size_t)j * m] = (double)(s - pa[i + (size_t)j * m]);
                        }
                    const double an = norm1(m, n, a);
                    const double ratio = an > 0 ? norm1(m, n, r) / ((double)n * an * EPS) : norm1(m, n, r) / EPS;
                    w01 = ratio > w01 ? ratio : w01;
                    ++t01;
                }
                if (m != n || info != 0 || n == 0) continue;
                // dget03 via dgetri_
                Mat inv = lu;
                std::vector<double> work((size_t)64 * n);
                const int lw = (int)work.size();
                int iinfo = 0;
                dgetri_(&n, inv.data(), &lda, piv.data(), work.data(), &lw, &iinfo);
                const double an = norm1(n, n, a), ain = norm1(n, n, inv);
                {
                    Mat r((size_t)n * n);
                    for (int j = 0; j < n; ++j)
                        for (int i = 0; i < n; ++i) {
                            long double s = (i == j) ? 1.0L : 0.0L;
                            for (int k = 0; k < n; ++k) s -= (long double)a[i + (size_t)k * n] * inv[k + (size_t)j * n];
                            r[i + (size_t)j * n] = (double)s;
                        }
                    const double ratio = norm1(n, n, r) / ((double)n * an * ain * EPS);
                    w03 = ratio > w03 ? ratio : w03;
                    ++t03;
                }
                const double rcondc = 1.0 / (an * ain);
                // dget06 via dgecon_
                {
                    double rc = 0.0, anorm = an;
                    std::vector<double> cw(4 * (size_t)n);
                    std::vector<int> ciw(n);
                    int cinfo = 0;
                    const char one = 'O';
                    dgecon_(&one, &n, lu.data(), &lda, &anorm, &rc, cw.data(), ciw.data(), &cinfo);
                    double ratio = 0.0;
                    if (rc > 0 && rcondc > 0) ratio = std::max(rc / rcondc, rcondc / rc) - (1.0 - EPS);
                    else if (rc > 0 || rcondc > 0) ratio = 1.0 / EPS;
                    // dget06 scales by 1/eps only when one estimate is 0; the
                    // estimate may exceed the true rcond by at most n (lacn2)
                    w06 = ratio > w06 ? ratio : w06;
                    ++t06;
                }
                for (int nrhs : nrhss)
                    for (int tr = 0; tr < 2; ++tr) {
                        const char t = tr ? 'T' : 'N';
                        std::uniform_real_distribution<double> U(-1.0, 1.0);
                        Mat xt((size_t)n * nrhs), b((size_t)n * nrhs);
                        for (auto &x : xt) x = U(rng);
                        for (int j = 0; j < nrhs; ++j)
                            for (int i = 0; i < n; ++i) {
                                long double s = 0.0L;
                                for (int k = 0; k < n; ++k)
                                    s += (long double)(tr ? a[k + (size_t)i * n] : a[i + (size_t)k * n]) * xt[k + (size_t)j * n];
                                b[i + (size_t)j * n] = (double)s;
                            }
                        Mat x = b;
                        int sinfo = 0;
                        dgetrs_(&t, &n, &nrhs, lu.data(), &lda, piv.data(), x.data(), &lda, &sinfo);
                        // dget02: ||b - op(A) x||_1 / (||A||_1 ||x||_1 eps), per column max
                        double r02 = 0.0, r04 = 0.0;
                        for (int j = 0; j < nrhs; ++j) {
                            double rn = 0.0, xn = 0.0, dn = 0.0, xtn = 0.0;
                            for (int i = 0; i < n; ++i) {
                                long double s = b[i + (size_t)j * n];
                                for (int k = 0; k < n; ++k)
                                    s -= (long double)(tr ? a[k + (size_t)i * n] : a[i + (size_t)k * n]) * x[k + (size_t)j * n];
                                rn += fabs((double)s);
                                xn += fabs(x[i + (size_t)j * n]);
                                dn = std::max(dn, fabs(x[i + (size_t)j * n] - xt[i + (size_t)j * n]));
                                xtn = std::max(xtn, fabs(xt[i + (size_t)j * n]));
                            }
                            const double an_t = tr ? dlange_("I", &n, &n, a.data(), &lda, nullptr) : an;
                            r02 = std::max(r02, xn > 0 ? rn / (an_t * xn * EPS) : 0.0);
                            r04 = std::max(r04, xtn > 0 ? dn / (xtn * (1.0 / rcondc) * EPS) : 0.0);
                        }
                        w02 = std::max(w02, r02);
                        ++t02;
                        w04 = std::max(w04, r04);
                        ++t04;
                        // dget07: refinement bounds from dgerfs_
                        std::vector<double> ferr(nrhs), berr(nrhs), rw(3 * (size_t)n);
                        std::vector<int> riw(n);
                        int rinfo = 0;
                        dgerfs_(&t, &n, &nrhs, a.data(), &lda, lu.data(), &lda, piv.data(), b.data(),
                                &lda, x.data(), &lda, ferr.data(), berr.data(), rw.data(), riw.data(),
                                &rinfo);
                        double r07 = 0.0;
                        for (int j = 0; j < nrhs; ++j) {
                            double dn = 0.0, xn = 0.0;
                            for (int i = 0; i < n; ++i) {
                                dn = std::max(dn, fabs(x[i + (size_t)j * n] - xt[i + (size_t)j * n]));
                                xn = std::max(xn, fabs(x[i + (size_t)j * n]));
                            }
                            // forward error within the bound, backward error ~ eps
                            if (xn > 0 && ferr[j] > 0) r07 = std::max(r07, dn / (xn * ferr[j]));
                            r07 = std::max(r07, berr[j] / ((n + 1) * EPS));
                        }
                        w07 = std::max(w07, r07);
                        ++t07;
                    }
            }
    report("dgetrf_ dget01 (LU residual)", t01, w01, 30.0);
    report("dgetri_ dget03 (inverse)", t03, w03, 30.0);
    report("dgecon_ dget06 (rcond)", t06, w06, 30.0);
    report("dgetrs_ dget02 (residual)", t02, w02, 30.0);
    report("dgetrs_ dget04 (fwd error)", t04, w04, 30.0);
    report("dgerfs_ dget07 (bounds)", t07, w07, 30.0);
    report("dgetrf_ info on zero columns", t01, bad_info ? 1e30 : 0.0, 1.0);
}

int main()
{
    std::mt19937_64 rng(20240531);
    gemm_error_exits();
    gemm_grid(rng);
    lu_suite(rng);
    printf("%s\n", g_fail ? "FAILED" : "ALL PASSED");
    return g_fail ? 1 : 0;
}
