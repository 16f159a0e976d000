/* aux_body.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the FLENS-LAPACK routines around the LU that estimate,
 * scale and refine: lan (GeMatrix norms), equ + laq (equilibration), lacn2 +
 * latrs + rscl + con (1-norm condition estimate of an LU factor), rfs
 * (iterative refinement with error bounds) and svx (expert driver).  Each
 * function follows the generic implementation it cites, operation for
 * operation (compiled with -ffp-contract=off, like the rest of the oracle),
 * so it can be compared bit for bit with oracle/_ref built from the
 * reference.  Included at the end of flens_oracle.c: it uses that file's
 * getrf/getrs restatement (orc_dgetrf_nb, orc_dgetrs_c).
 */

#define A_(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)lda]
#define AUX_EPS (DBL_EPSILON * 0.5)   /* lamch Eps, flens/lapack/la/lamch.tcc:69 */
#define AUX_PREC DBL_EPSILON          /* lamch Precision = Eps * base */
#define AUX_SAFEMIN DBL_MIN           /* lamch SafeMin: max(min, (1/max)(1+eps)) */

/* std::max / std::min: a NaN second operand never wins */
static double aux_max(double a, double b) { return (a < b) ? b : a; }
static double aux_min(double a, double b) { return (b < a) ? b : a; }

/* flens/lapack/auxiliary/sign.tcc:48-62: sign(1, y), -1 for y = -0 */
static double aux_sign1(double y)
{
    if (y > 0.0) return 1.0;
    if (y < 0.0) return -1.0;
    return signbit(y) ? -1.0 : 1.0;
}

/* flens/lapack/auxiliary/nint.tcc:42-47 */
static double aux_nint(double x) { return floor(x + 0.5); }

/* cxxblas/level1/asum.tcc:44-54 */
static double aux_asum(int n, const double *x, int incx)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += fabs(x[(size_t)i * incx]);
    return s;
}

/* cxxblas/level1/iamax.tcc:42-58 (0-based, first maximum; n >= 1) */
static int aux_iamax(int n, const double *x)
{
    int im = 0;
    double m = fabs(x[0]);
    for (int i = 0; i < n; ++i)
        if (fabs(x[i]) > m) { im = i; m = fabs(x[i]); }
    return im;
}

/* cxxblas/level1/dot.tcc:40-52: sum from 0, ascending */
static double aux_dot(int n, const double *x, int incx, const double *y, int incy)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[(size_t)i * incx] * y[(size_t)i * incy];
    return s;
}

/* flens/lapack/la/lassq.tcc:58-80 */
static void aux_lassq(int n, const double *x, double *scale, double *sumsq)
{
    for (int i = 0; i < n; ++i) {
        if (x[i] != 0.0) {
            const double a = fabs(x[i]);
            if (*scale < a) {
                *sumsq = 1.0 + *sumsq * pow(*scale / a, 2.0);
                *scale = a;
            } else {
                *sumsq += pow(a / *scale, 2.0);
            }
        }
    }
}

/* flens/lapack/la/lan.tcc:55-130 (GeMatrix); norm 'M', 'O'/'1', 'I', 'F' */
double orc_dlange(char norm, int m, int n, const double *a, int lda)
{
    if ((m < n ? m : n) == 0) return 0.0;
    if (norm == 'M' || norm == 'm') {
        double v = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) v = aux_max(v, fabs(A_(i, j)));
        return v;
    }
    if (norm == 'O' || norm == 'o' || norm == '1') {
        double v = 0.0;
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < m; ++i) s += fabs(A_(i, j));
            v = aux_max(v, s);
        }
        return v;
    }
    if (norm == 'I' || norm == 'i') {
        double *w = (double *)calloc((size_t)m, sizeof(double));
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) w[i] += fabs(A_(i, j));
        double v = 0.0;
        for (int i = 0; i < m; ++i) v = aux_max(v, w[i]);
        free(w);
        return v;
    }
    double scale = 0.0, sum = 1.0;
    for (int j = 0; j < n; ++j) aux_lassq(m, &A_(0, j), &scale, &sum);
    return scale * sqrt(sum);
}

/* flens/lapack/ge/equ.tcc:62-194; r (m), c (n); returns info.  rowcnd /
 * colcnd are left untouched on the paths that do not set them. */
int orc_dgeequ(int m, int n, const double *a, int lda, double *r, double *c, double *rowcnd,
               double *colcnd, double *amax)
{
    if (m == 0 || n == 0) {
        *rowcnd = 1.0; *colcnd = 1.0; *amax = 0.0;
        return 0;
    }
    const double smlnum = AUX_SAFEMIN, bignum = 1.0 / smlnum;
    for (int i = 0; i < m; ++i) r[i] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) r[i] = aux_max(r[i], fabs(A_(i, j)));
    double rcmin = bignum, rcmax = 0.0;
    for (int i = 0; i < m; ++i) { rcmax = aux_max(rcmax, r[i]); rcmin = aux_min(rcmin, r[i]); }
    *amax = rcmax;
    if (rcmin == 0.0) {
        for (int i = 0; i < m; ++i)
            if (r[i] == 0.0) return i + 1;
    } else {
        for (int i = 0; i < m; ++i) r[i] = 1.0 / aux_min(aux_max(r[i], smlnum), bignum);
        *rowcnd = aux_max(rcmin, smlnum) / aux_min(rcmax, bignum);
    }
    for (int j = 0; j < n; ++j) c[j] = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) c[j] = aux_max(c[j], fabs(A_(i, j)) * r[i]);
    rcmin = bignum; rcmax = 0.0;
    for (int j = 0; j < n; ++j) { rcmin = aux_min(rcmin, c[j]); rcmax = aux_max(rcmax, c[j]); }
    if (rcmin == 0.0) {
        for (int j = 0; j < n; ++j)
            if (c[j] == 0.0) return m + j + 1;
    } else {
        for (int j = 0; j < n; ++j) c[j] = 1.0 / aux_min(aux_max(c[j], smlnum), bignum);
        *colcnd = aux_max(rcmin, smlnum) / aux_min(rcmax, bignum);
    }
    return 0;
}

/* flens/lapack/la/laq.tcc:57-118; returns 'N', 'R', 'C' or 'B' */
int orc_dlaqge(int m, int n, double *a, int lda, const double *r, const double *c,
               double rowcnd, double colcnd, double amax)
{
    if (m == 0 || n == 0) return 'N';
    const double thresh = 0.1;
    const double small = AUX_SAFEMIN / AUX_PREC, large = 1.0 / small;
    if (rowcnd >= thresh && amax >= small && amax <= large) {
        if (colcnd >= thresh) return 'N';
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) A_(i, j) *= c[j];
        return 'C';
    }
    if (colcnd >= thresh) {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) A_(i, j) *= r[i];
        return 'R';
    }
    for (int j = 0; j < n; ++j) {
        const double cj = c[j];
        for (int i = 0; i < m; ++i) A_(i, j) *= cj * r[i];
    }
    return 'B';
}

/* flens/lapack/la/lacn2.tcc:55-185 (real).  isave[0..2] = ISAVE(1..3),
 * ISAVE(2) kept 1-based as in the reference. */
static void aux_lacn2(int n, double *v, double *x, int *sgn, double *est, int *kase, int *isave)
{
    const int itmax = 5;
    if (*kase == 0) {
        for (int i = 0; i < n; ++i) x[i] = 1.0 / (double)n;
        *kase = 1;
        isave[0] = 1;
        return;
    }
    int jlast;
    double estold;
    switch (isave[0]) {
    case 1:
        if (n == 1) {
            v[0] = x[0];
            *est = fabs(v[0]);
            goto quit;
        }
        *est = aux_asum(n, x, 1);
        for (int i = 0; i < n; ++i) {
            x[i] = aux_sign1(x[i]);
            sgn[i] = (int)aux_nint(x[i]);
        }
        *kase = 2;
        isave[0] = 2;
        return;
    case 2:
        isave[1] = aux_iamax(n, x) + 1;
        isave[2] = 2;
    main_loop:
        for (int i = 0; i < n; ++i) x[i] = 0.0;
        x[isave[1] - 1] = 1.0;
        *kase = 1;
        isave[0] = 3;
        return;
    case 3: {
        for (int i = 0; i < n; ++i) v[i] = x[i];
        estold = *est;
        *est = aux_asum(n, v, 1);
        int converged = 1;
        for (int i = 0; i < n; ++i)
            if ((int)aux_nint(aux_sign1(x[i])) != sgn[i]) { converged = 0; break; }
        if (converged) goto complete;
        if (*est <= estold) goto complete;
        for (int i = 0; i < n; ++i) {
            x[i] = aux_sign1(x[i]);
            sgn[i] = (int)aux_nint(x[i]);
        }
        *kase = 2;
        isave[0] = 4;
        return;
    }
    case 4:
        jlast = isave[1];
        isave[1] = aux_iamax(n, x) + 1;
        if (x[jlast - 1] != fabs(x[isave[1] - 1]) && isave[2] < itmax) {
            ++isave[2];
            goto main_loop;
        }
    complete: {
        double alt = 1.0;
        for (int i = 0; i < n; ++i) {
            x[i] = alt * (1.0 + (double)i / (double)(n - 1));
            alt = -alt;
        }
        *kase = 1;
        isave[0] = 5;
        return;
    }
    case 5: {
        const double temp = 2.0 * (aux_asum(n, x, 1) / (double)(3 * n));
        if (temp > *est) {
            for (int i = 0; i < n; ++i) v[i] = x[i];
            *est = temp;
        }
    }
    }
quit:
    *kase = 0;
}

/* flens/lapack/ge/rscl.tcc:52-110: x <- x / sa without intermediate
 * over/underflow (labad is the identity for IEEE double) */
static void aux_rscl(int n, double sa, double *x)
{
    if (n == 0) return;
    const double smlnum = AUX_SAFEMIN, bignum = 1.0 / smlnum;
    double cden = sa, cnum = 1.0;
    int done = 0;
    do {
        const double cden1 = cden * smlnum, cnum1 = cnum / bignum;
        double mul;
        if (fabs(cden1) > fabs(cnum) && cnum != 0.0) {
            mul = smlnum; done = 0; cden = cden1;
        } else if (fabs(cnum1) > fabs(cden)) {
            mul = bignum; done = 0; cnum = cnum1;
        } else {
            mul = cnum / cden; done = 1;
        }
        for (int i = 0; i < n; ++i) x[i] *= mul;
    } while (!done);
}

/* flens/lapack/la/latrs.tcc:60-530 (real): solve op(T) x = scale * b with
 * the growth bound deciding between the plain substitution (blas::sv ->
 * cxxblas/level2/trsv.tcc) and the scaled column / dot-product solve.
 * T is the upper (upper = 1, non-unit) or unit-lower (upper = 0) triangle of
 * an LU factor; cnorm (n) is computed when normin == 0. */
static void aux_latrs(int upper, int trans, int normin, int n, const double *a, int lda,
                      double *x, double *scale, double *cnorm)
{
    const int nounit = upper;   /* U: non-unit; L: unit */
    const double smlnum = AUX_SAFEMIN / AUX_PREC, bignum = 1.0 / smlnum;
    *scale = 1.0;
    if (n == 0) return;
    if (!normin) {
        if (upper) {
            for (int j = 0; j < n; ++j) cnorm[j] = aux_asum(j, &A_(0, j), 1);
        } else {
            for (int j = 0; j < n - 1; ++j) cnorm[j] = aux_asum(n - j - 1, &A_(j + 1, j), 1);
            cnorm[n - 1] = 0.0;
        }
    }
    const int imax = aux_iamax(n, cnorm);
    const double tmax = cnorm[imax];
    double tscal;
    if (tmax <= bignum) {
        tscal = 1.0;
    } else {
        tscal = 1.0 / (smlnum * tmax);
        for (int j = 0; j < n; ++j) cnorm[j] *= tscal;
    }
    double xmax = fabs(x[aux_iamax(n, x)]);
    double xbnd = xmax, grow;
    int jfirst, jlast, jinc;
    if (!trans) {
        if (upper) { jfirst = n - 1; jlast = 0; jinc = -1; }
        else { jfirst = 0; jlast = n - 1; jinc = 1; }
    } else {
        if (upper) { jfirst = 0; jlast = n - 1; jinc = 1; }
        else { jfirst = n - 1; jlast = 0; jinc = -1; }
    }
    const int jend = jlast + jinc;
    if (tscal != 1.0) {
        grow = 0.0;
    } else if (!trans) {
        if (nounit) {
            grow = 1.0 / aux_max(xbnd, smlnum);
            xbnd = grow;
            int small = 0;
            for (int j = jfirst; j != jend; j += jinc) {
                if (grow <= smlnum) { small = 1; break; }
                const double tjj = fabs(A_(j, j));
                xbnd = aux_min(xbnd, aux_min(1.0, tjj) * grow);
                if (tjj + cnorm[j] >= smlnum) grow *= tjj / (tjj + cnorm[j]);
                else grow = 0.0;
            }
            if (!small) grow = xbnd;
        } else {
            grow = aux_min(1.0, 1.0 / aux_max(xbnd, smlnum));
            for (int j = jfirst; j != jend; j += jinc) {
                if (grow <= smlnum) break;
                grow *= 1.0 / (1.0 + cnorm[j]);
            }
        }
    } else {
        if (nounit) {
            grow = 1.0 / aux_max(xbnd, smlnum);
            xbnd = grow;
            for (int j = jfirst; j != jend; j += jinc) {
                /* the reference leaves its "too small" flag false here
                 * (latrs.tcc:248-252), so the final min below always runs */
                if (grow <= smlnum) break;
                const double xj = 1.0 + cnorm[j];
                grow = aux_min(grow, xbnd / xj);
                const double tjj = fabs(A_(j, j));
                if (xj > tjj) xbnd *= tjj / xj;
            }
            grow = aux_min(grow, xbnd);
        } else {
            grow = aux_min(1.0, 1.0 / aux_max(xbnd, smlnum));
            for (int j = jfirst; j != jend; j += jinc) {
                if (grow <= smlnum) break;
                const double xj = 1.0 + cnorm[j];
                grow /= xj;
            }
        }
    }
    if (grow * tscal > smlnum) {
        /* blas::sv: cxxblas/level2/trsv.tcc (ColMajor mapped to RowMajor) */
        if (!trans) {
            if (upper) {
                for (int i = n - 1; i >= 0; --i) {
                    double t = 0.0;
                    for (int k = i + 1; k < n; ++k) t += A_(i, k) * x[k];
                    x[i] -= t;
                    x[i] /= A_(i, i);
                }
            } else {
                for (int i = 0; i < n; ++i) {
                    double t = 0.0;
                    for (int k = 0; k < i; ++k) t += A_(i, k) * x[k];
                    x[i] -= t;
                }
            }
        } else {
            if (upper) {
                for (int i = 0; i < n; ++i) {
                    const double t = aux_dot(i, &A_(0, i), 1, x, 1);
                    x[i] -= t;
                    x[i] /= A_(i, i);
                }
            } else {
                for (int i = n - 1; i >= 0; --i) {
                    const double t = aux_dot(n - i - 1, &A_(i + 1, i), 1, x + i + 1, 1);
                    x[i] -= t;
                }
            }
        }
        return;
    }
    /* scaled solve */
    if (xmax > bignum) {
        *scale = bignum / xmax;
        for (int i = 0; i < n; ++i) x[i] *= *scale;
        xmax = bignum;
    }
    if (!trans) {
        for (int j = jfirst; j != jend; j += jinc) {
            double xj = fabs(x[j]);
            double tjjs = 0.0;
            int skip = 0;
            if (nounit) tjjs = A_(j, j) * tscal;
            else { tjjs = tscal; if (tscal == 1.0) skip = 1; }
            if (!skip) {
                const double tjj = fabs(tjjs);
                if (tjj > smlnum) {
                    if (tjj < 1.0 && xj > tjj * bignum) {
                        const double rec = 1.0 / xj;
                        for (int i = 0; i < n; ++i) x[i] *= rec;
                        *scale *= rec; xmax *= rec;
                    }
                    x[j] /= tjjs;
                    xj = fabs(x[j]);
                } else if (tjj > 0.0) {
                    if (xj > tjj * bignum) {
                        double rec = (tjj * bignum) / xj;
                        if (cnorm[j] > 1.0) rec /= cnorm[j];
                        for (int i = 0; i < n; ++i) x[i] *= rec;
                        *scale *= rec; xmax *= rec;
                    }
                    x[j] /= tjjs;
                    xj = fabs(x[j]);
                } else {
                    for (int i = 0; i < n; ++i) x[i] = 0.0;
                    x[j] = 1.0; xj = 1.0; *scale = 0.0; xmax = 0.0;
                }
            }
            if (xj > 1.0) {
                double rec = 1.0 / xj;
                if (cnorm[j] > (bignum - xmax) * rec) {
                    rec *= 0.5;
                    for (int i = 0; i < n; ++i) x[i] *= rec;
                    *scale *= rec;
                }
            } else if (xj * cnorm[j] > (bignum - xmax)) {
                for (int i = 0; i < n; ++i) x[i] *= 0.5;
                *scale *= 0.5;
            }
            if (upper) {
                if (j > 0) {
                    const double f = x[j] * tscal;
                    for (int i = 0; i < j; ++i) x[i] -= f * A_(i, j);
                    xmax = fabs(x[aux_iamax(j, x)]);
                }
            } else if (j < n - 1) {
                const double f = x[j] * tscal;
                for (int i = j + 1; i < n; ++i) x[i] -= f * A_(i, j);
                xmax = fabs(x[j + 1 + aux_iamax(n - j - 1, x + j + 1)]);
            }
        }
    } else {
        for (int j = jfirst; j != jend; j += jinc) {
            double xj = fabs(x[j]);
            double uscal = tscal;
            double rec = 1.0 / aux_max(xmax, 1.0);
            double tjjs = 0.0;
            if (cnorm[j] > (bignum - xj) * rec) {
                rec *= 0.5;
                tjjs = nounit ? A_(j, j) * tscal : tscal;
                const double tjj = fabs(tjjs);
                if (tjj > 1.0) {
                    rec = aux_min(1.0, rec * tjj);
                    uscal /= tjjs;
                }
                if (rec < 1.0) {
                    for (int i = 0; i < n; ++i) x[i] *= rec;
                    *scale *= rec; xmax *= rec;
                }
            }
            double sumj = 0.0;
            if (uscal == 1.0) {
                if (upper) sumj = aux_dot(j, &A_(0, j), 1, x, 1);
                else if (j < n - 1) sumj = aux_dot(n - j - 1, &A_(j + 1, j), 1, x + j + 1, 1);
            } else {
                if (upper) {
                    for (int i = 0; i < j; ++i) sumj += (A_(i, j) * uscal) * x[i];
                } else if (j < n - 1) {
                    for (int i = j + 1; i < n; ++i) sumj += (A_(i, j) * uscal) * x[i];
                }
            }
            if (uscal == tscal) {
                x[j] -= sumj;
                xj = fabs(x[j]);
                int skip = 0;
                if (nounit) tjjs = A_(j, j) * tscal;
                else { tjjs = tscal; if (tscal == 1.0) skip = 1; }
                if (!skip) {
                    const double tjj = fabs(tjjs);
                    if (tjj > smlnum) {
                        if (tjj < 1.0 && xj > tjj * bignum) {
                            rec = 1.0 / xj;
                            for (int i = 0; i < n; ++i) x[i] *= rec;
                            *scale *= rec; xmax *= rec;
                        }
                        x[j] /= tjjs;
                    } else if (tjj > 0.0) {
                        if (xj > tjj * bignum) {
                            const double r2 = (tjj * bignum) / xj;
                            for (int i = 0; i < n; ++i) x[i] *= r2;
                            *scale *= r2; xmax *= r2;
                        }
                        x[j] /= tjjs;
                    } else {
                        for (int i = 0; i < n; ++i) x[i] = 0.0;
                        x[j] = 1.0; *scale = 0.0; xmax = 0.0;
                    }
                }
            } else {
                x[j] = x[j] / tjjs - sumj;
            }
            xmax = aux_max(xmax, fabs(x[j]));
        }
    }
    *scale /= tscal;
    if (tscal != 1.0)
        for (int j = 0; j < n; ++j) cnorm[j] *= 1.0 / tscal;
}

/* flens/lapack/ge/cond.tcc:62-165 (real): reciprocal condition number of
 * the matrix whose LU factor (trf layout) is lu, in the 1-norm ('O') or the
 * infinity norm ('I'), given that norm of the matrix. */
double orc_dgecon(char norm, int n, const double *lu, int lda, double anorm)
{
    if (n == 0) return 1.0;
    if (anorm == 0.0) return 0.0;
    const double smlnum = AUX_SAFEMIN;
    double *work = (double *)calloc((size_t)4 * n, sizeof(double));
    int *iwork = (int *)calloc((size_t)n, sizeof(int));
    double *w1 = work, *w2 = work + n, *w3 = work + 2 * n, *w4 = work + 3 * n;
    double ainvnm = 0.0, rcond = 0.0;
    int normin = 0, kase = 0, isave[3] = {0, 0, 0};
    const int kase1 = (norm == 'O' || norm == 'o' || norm == '1') ? 1 : 2;
    int ok = 1;
    for (;;) {
        aux_lacn2(n, w2, w1, iwork, &ainvnm, &kase, isave);
        if (kase == 0) break;
        double sl, su;
        if (kase == kase1) {
            aux_latrs(0, 0, normin, n, lu, lda, w1, &sl, w3);
            aux_latrs(1, 0, normin, n, lu, lda, w1, &su, w4);
        } else {
            aux_latrs(1, 1, normin, n, lu, lda, w1, &su, w4);
            aux_latrs(0, 1, normin, n, lu, lda, w1, &sl, w3);
        }
        const double scale = sl * su;
        normin = 1;
        if (scale != 1.0) {
            const int ix = aux_iamax(n, w1);
            if (scale < fabs(w1[ix]) * smlnum || scale == 0.0) { ok = 0; break; }
            aux_rscl(n, scale, w1);
        }
    }
    if (ok && ainvnm != 0.0) rcond = (1.0 / ainvnm) / anorm;
    free(work);
    free(iwork);
    return rcond;
}

/* flens/lapack/ge/rfs.tcc:63-261.  trans 0/1; x (n x nrhs) refined in
 * place; ferr, berr (nrhs). */
void orc_dgerfs(int trans, int n, int nrhs, const double *a, int lda, const double *af,
                int ldaf, const int *ipiv, const double *b, int ldb, double *x, int ldx,
                double *ferr, double *berr)
{
    if (n == 0 || nrhs == 0) {
        for (int j = 0; j < nrhs; ++j) { ferr[j] = 0.0; berr[j] = 0.0; }
        return;
    }
    const int itmax = 5;
    const int nz = n + 1;
    const double eps = AUX_EPS, safmin = AUX_SAFEMIN;
    const double safe1 = nz * safmin, safe2 = safe1 / eps;
    double *work = (double *)calloc((size_t)3 * n, sizeof(double));
    int *iwork = (int *)calloc((size_t)n, sizeof(int));
    double *w1 = work, *w2 = work + n, *w3 = work + 2 * n;
    for (int j = 0; j < nrhs; ++j) {
        const double *bj = b + (size_t)j * ldb;
        double *xj = x + (size_t)j * ldx;
        int count = 1;
        double lstres = 3.0;
        for (;;) {
            /* r = b - op(A) x: blas::mv(trans, -1, A, x, 1, r), generic gemv
             * (cxxblas/level2/gemv.tcc): y = beta y, then each y_i +=
             * alpha * dot(row i of op(A), x) */
            for (int i = 0; i < n; ++i) w2[i] = bj[i];
            if (!trans) {
                for (int i = 0; i < n; ++i) {
                    const double t = aux_dot(n, &A_(i, 0), lda, xj, 1);
                    w2[i] += -1.0 * t;
                }
            } else {
                for (int i = 0; i < n; ++i) {
                    const double t = aux_dot(n, &A_(0, i), 1, xj, 1);
                    w2[i] += -1.0 * t;
                }
            }
            for (int i = 0; i < n; ++i) w1[i] = fabs(bj[i]);
            if (!trans) {
                for (int k = 0; k < n; ++k) {
                    const double xk = fabs(xj[k]);
                    for (int i = 0; i < n; ++i) w1[i] += fabs(A_(i, k)) * xk;
                }
            } else {
                for (int k = 0; k < n; ++k) {
                    double s = 0.0;
                    for (int i = 0; i < n; ++i) s += fabs(A_(i, k)) * fabs(xj[i]);
                    w1[k] += s;
                }
            }
            double s = 0.0;
            for (int i = 0; i < n; ++i) {
                if (w1[i] > safe2) s = aux_max(s, fabs(w2[i]) / w1[i]);
                else s = aux_max(s, (fabs(w2[i]) + safe1) / (w1[i] + safe1));
            }
            berr[j] = s;
            if (berr[j] > eps && 2.0 * berr[j] <= lstres && count <= itmax) {
                orc_dgetrs_c(trans, n, 1, af, ldaf, ipiv, w2, n);
                for (int i = 0; i < n; ++i) xj[i] += w2[i];
                lstres = berr[j];
                ++count;
                continue;
            }
            break;
        }
        for (int i = 0; i < n; ++i) {
            if (w1[i] > safe2) w1[i] = fabs(w2[i]) + nz * eps * w1[i];
            else w1[i] = fabs(w2[i]) + nz * eps * w1[i] + safe1;
        }
        int kase = 0, isave[3] = {0, 0, 0};
        for (;;) {
            aux_lacn2(n, w3, w2, iwork, &ferr[j], &kase, isave);
            if (kase == 0) break;
            if (kase == 1) {
                orc_dgetrs_c(!trans, n, 1, af, ldaf, ipiv, w2, n);
                for (int i = 0; i < n; ++i) w2[i] *= w1[i];
            } else {
                for (int i = 0; i < n; ++i) w2[i] *= w1[i];
                orc_dgetrs_c(trans, n, 1, af, ldaf, ipiv, w2, n);
            }
        }
        lstres = 0.0;
        for (int i = 0; i < n; ++i) lstres = aux_max(lstres, fabs(xj[i]));
        if (lstres != 0.0) ferr[j] /= lstres;
    }
    free(work);
    free(iwork);
}

/* flens/lapack/ge/svx.tcc:60-262.  fact 'N', 'E' or 'F'; *equed is read for
 * fact 'F' and, unlike the reference (which takes it by value), written back
 * for 'E'.  rpvgrw receives work(1).  Returns info. */
int orc_dgesvx(char fact, int trans, int n, int nrhs, double *a, int lda, double *af,
               int ldaf, int *ipiv, char *equed, double *r, double *c, double *b, int ldb,
               double *x, int ldx, double *rcond, double *ferr, double *berr, double *rpvgrw,
               int nb)
{
    const double smlnum = AUX_SAFEMIN, bignum = 1.0 / smlnum;
    int info = 0, rowequ, colequ;
    double rowcnd = 1.0, colcnd = 1.0;
    if (fact == 'N' || fact == 'E') {
        *equed = 'N';
        rowequ = colequ = 0;
    } else {
        rowequ = (*equed == 'R' || *equed == 'B');
        colequ = (*equed == 'C' || *equed == 'B');
    }
    if (rowequ) {
        double mn = bignum, mx = 0.0;
        for (int j = 0; j < n; ++j) { mn = aux_min(mn, r[j]); mx = aux_max(mx, r[j]); }
        rowcnd = n > 0 ? aux_max(mn, smlnum) / aux_min(mx, bignum) : 1.0;
    }
    if (colequ) {
        double mn = bignum, mx = 0.0;
        for (int j = 0; j < n; ++j) { mn = aux_min(mn, c[j]); mx = aux_max(mx, c[j]); }
        colcnd = n > 0 ? aux_max(mn, smlnum) / aux_min(mx, bignum) : 1.0;
    }
    if (fact == 'E') {
        double amax = 0.0;
        if (orc_dgeequ(n, n, a, lda, r, c, &rowcnd, &colcnd, &amax) == 0) {
            *equed = (char)orc_dlaqge(n, n, a, lda, r, c, rowcnd, colcnd, amax);
            rowequ = (*equed == 'R' || *equed == 'B');
            colequ = (*equed == 'C' || *equed == 'B');
        }
    }
    if (!trans) {
        if (rowequ)
            for (int j = 0; j < nrhs; ++j)
                for (int i = 0; i < n; ++i) b[i + (size_t)j * ldb] *= r[i];
    } else if (colequ) {
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i) b[i + (size_t)j * ldb] *= c[i];
    }
    if (fact == 'N' || fact == 'E') {
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) af[i + (size_t)j * ldaf] = A_(i, j);
        info = orc_dgetrf_nb(n, n, af, ldaf, ipiv, nb);
        if (info > 0) {
            double g = 0.0;   /* max |U(1:info,1:info)| */
            for (int j = 0; j < info; ++j)
                for (int i = 0; i <= j; ++i) g = aux_max(g, fabs(af[i + (size_t)j * ldaf]));
            *rpvgrw = (g == 0.0) ? 1.0 : orc_dlange('M', n, n, a, lda) / g;
            *rcond = 0.0;
            return info;
        }
    }
    const char nrm = trans ? 'I' : 'O';
    const double anorm = orc_dlange(nrm, n, n, a, lda);
    double g = 0.0;   /* max |U| */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) g = aux_max(g, fabs(af[i + (size_t)j * ldaf]));
    *rpvgrw = (g == 0.0) ? 1.0 : orc_dlange('M', n, n, a, lda) / g;
    *rcond = orc_dgecon(nrm, n, af, ldaf, anorm);
    for (int j = 0; j < nrhs; ++j)
        for (int i = 0; i < n; ++i) x[i + (size_t)j * ldx] = b[i + (size_t)j * ldb];
    orc_dgetrs_c(trans, n, nrhs, af, ldaf, ipiv, x, ldx);
    orc_dgerfs(trans, n, nrhs, a, lda, af, ldaf, ipiv, b, ldb, x, ldx, ferr, berr);
    if (!trans) {
        if (colequ) {
            for (int j = 0; j < nrhs; ++j)
                for (int i = 0; i < n; ++i) x[i + (size_t)j * ldx] *= c[i];
            for (int j = 0; j < nrhs; ++j) ferr[j] /= colcnd;
        }
    } else if (rowequ) {
        for (int j = 0; j < nrhs; ++j)
            for (int i = 0; i < n; ++i) x[i + (size_t)j * ldx] *= r[i];
        for (int j = 0; j < nrhs; ++j) ferr[j] /= rowcnd;
    }
    if (*rcond < AUX_EPS) info = n + 1;
    return info;
}

/* flens/lapack/ge/tri.tcc:64-187 with trtri (flens/lapack/tr/tri.tcc) on
 * their unblocked paths -- ti2 (flens/lapack/tr/ti2.tcc:58-90, upper,
 * non-unit, trmv per cxxblas/level2/trmv.tcc) and the column-by-column solve
 * of inv(A) L = inv(U) (gemv per cxxblas/level2/gemv.tcc) -- which the
 * reference runs for n <= 64 (ilaenv GETRI = TRTRI = 64).  Returns info
 * (first zero U(i,i), A untouched). */
int orc_dgetri(int n, double *a, int lda, const int *ipiv)
{
    if (n == 0) return 0;
    for (int i = 0; i < n; ++i)
        if (A_(i, i) == 0.0) return i + 1;
    /* ti2: inv(U) in place */
    for (int j = 0; j < n; ++j) {
        A_(j, j) = 1.0 / A_(j, j);
        const double ajj = -A_(j, j);
        for (int i = 0; i < j; ++i) {   /* u12 <- U11 u12 (U11 already inverted) */
            double t = 0.0;
            for (int k = i; k < j; ++k) t += A_(i, k) * A_(k, j);
            A_(i, j) = t;
        }
        for (int i = 0; i < j; ++i) A_(i, j) *= ajj;
    }
    /* inv(A) L = inv(U), columns n..1 */
    double *work = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int j = n - 1; j >= 0; --j) {
        for (int i = j + 1; i < n; ++i) { work[i] = A_(i, j); A_(i, j) = 0.0; }
        if (j < n - 1) {
            for (int i = 0; i < n; ++i) {
                const double t = aux_dot(n - j - 1, &A_(i, j + 1), lda, work + j + 1, 1);
                A_(i, j) += -1.0 * t;
            }
        }
    }
    free(work);
    for (int j = n - 2; j >= 0; --j) {
        const int jp = ipiv[j] - 1;
        if (jp != j)
            for (int i = 0; i < n; ++i) {
                const double t = A_(i, j);
                A_(i, j) = A_(i, jp);
                A_(i, jp) = t;
            }
    }
    return 0;
}

#undef A_
