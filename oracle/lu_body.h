/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY.  Never linked into the product
 * library; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker / timed CPU baseline.
 *
 * Type-generic body of the LU restatement, included twice by
 * flens_oracle.c with REAL = double (prefix orc_d) and REAL = float
 * (prefix orc_s).  Every routine follows the loop and summation order of
 * the FLENS generic path it cites, so that compiled with
 * -ffp-contract=off it reproduces the reference bit for bit.
 * Storage is column-major, indices 0-based internally; pivot vectors are
 * 1-based on output, as in LAPACK/FLENS.
 */

#define A_(i, j) a[(size_t)(i) + (size_t)(j) * (size_t)lda]

/* cxxblas/level1/iamax.tcc:42-58: first index of max |x|, strict '>' */
static int FN(iamax)(int n, const REAL *x, int incx)
{
    int best = 0;
    REAL bv = RABS(x[0]);
    for (int i = 0; i < n; ++i) {
        REAL v = RABS(x[(size_t)i * incx]);
        if (v > bv) { best = i; bv = v; }
    }
    return best;
}

/*
 * flens/lapack/ge/tf2.tcc:56-136 (unblocked LU, LAPACK getf2).  piv[j] is
 * the 0-based row relative to a's first row.  Returns info (1-based
 * column of the first exact-zero pivot, 0 if none).
 */
static int FN(tf2)(int m, int n, REAL *a, int lda, int *piv, double *gap)
{
    int info = 0;
    if (m == 0 || n == 0) return 0;
    const REAL safemin = RSAFEMIN;
    const int mn = m < n ? m : n;
    for (int j = 0; j < mn; ++j) {
        /* tf2.tcc:101-102 */
        int jp = j + FN(iamax)(m - j, &A_(j, j), 1);
        piv[j] = jp;
        if (gap) {
            /* relative gap between the chosen |pivot| and the runner-up */
            double best = (double)RABS(A_(jp, j)), second = 0;
            for (int i = j; i < m; ++i)
                if (i != jp && (double)RABS(A_(i, j)) > second)
                    second = (double)RABS(A_(i, j));
            gap[j] = best > 0 ? (best - second) / best : 0;
        }
        if (A_(jp, j) != (REAL)0) {
            /* tf2.tcc:107-109: swap whole rows of the panel view */
            if (jp != j) {
                for (int c = 0; c < n; ++c) {
                    REAL t = A_(j, c); A_(j, c) = A_(jp, c); A_(jp, c) = t;
                }
            }
            /* tf2.tcc:112-121 */
            if (j < m - 1) {
                if (RABS(A_(j, j)) >= safemin) {
                    REAL r = (REAL)1 / A_(j, j);
                    for (int i = j + 1; i < m; ++i) A_(i, j) *= r;
                } else {
                    for (int i = j + 1; i < m; ++i) A_(i, j) /= A_(j, j);
                }
            }
        } else if (info == 0) {
            info = j + 1;                       /* tf2.tcc:122-126 */
        }
        /* tf2.tcc:130-132: ru(-1, A(j+1:m,j), A(j,j+1:n), A(j+1:m,j+1:n))
         * -> cxxblas/level2/ger.tcc:42-65: per column c, axpy with
         *    scalar (alpha*y_c) (cxxblas/level1/axpy.tcc:43-55). */
        if (j < mn - 1) {
            for (int c = j + 1; c < n; ++c) {
                REAL t = (REAL)(-1) * A_(j, c);
                for (int i = j + 1; i < m; ++i) A_(i, c) += t * A_(i, j);
            }
        }
    }
    return info;
}

/*
 * flens/lapack/la/laswp.tcc:54-93 with a forward pivot range k1..k2-1
 * (0-based, absolute rows).  Columns are processed in blocks of 32 like
 * the reference; the result is pure data movement either way.
 */
static void FN(laswp_fwd)(int ncols, REAL *a, int lda, int k1, int k2,
                          const int *piv0)
{
    for (int i = k1; i < k2; ++i) {
        int ip = piv0[i];
        if (ip != i) {
            for (int c = 0; c < ncols; ++c) {
                REAL t = A_(i, c); A_(i, c) = A_(ip, c); A_(ip, c) = t;
            }
        }
    }
}

static void FN(laswp_bwd)(int ncols, REAL *a, int lda, int k1, int k2,
                          const int *piv0)
{
    for (int i = k2 - 1; i >= k1; --i) {
        int ip = piv0[i];
        if (ip != i) {
            for (int c = 0; c < ncols; ++c) {
                REAL t = A_(i, c); A_(i, c) = A_(ip, c); A_(ip, c) = t;
            }
        }
    }
}

/*
 * trsm Left/Lower/NoTrans/Unit on column-major data: cxxblas/level3/
 * trsm.tcc:40-68 maps ColMajor to RowMajor/Right/Upper and calls trsv per
 * column with transA=Trans; cxxblas/level2/trsv.tcc (Trans, Upper, Unit
 * branch) is the dot-form forward substitution
 *   x_i -= sum_{k<i} L(i,k) x_k   (sum from 0, k ascending).
 * The final gescal by alpha=1 is exact and omitted.
 */
static void FN(trsm_llnu)(int m, int n, const REAL *a, int lda, REAL *b,
                          int ldb)
{
    for (int c = 0; c < n; ++c) {
        REAL *x = b + (size_t)c * ldb;
        for (int i = 0; i < m; ++i) {
            REAL s = 0;
            for (int k = 0; k < i; ++k) s += A_(i, k) * x[k];
            x[i] -= s;
        }
    }
}

/* trsm Left/Upper/NoTrans/NonUnit: trsv (RowMajor, Lower, Trans, NonUnit)
 * branch: x_i -= sum_{c>i} U(i,c) x_c (c ascending from i+1); x_i /= U(i,i). */
static void FN(trsm_lunn)(int m, int n, const REAL *a, int lda, REAL *b,
                          int ldb)
{
    for (int c = 0; c < n; ++c) {
        REAL *x = b + (size_t)c * ldb;
        for (int i = m - 1; i >= 0; --i) {
            REAL s = 0;
            for (int k = i + 1; k < m; ++k) s += A_(i, k) * x[k];
            x[i] -= s;
            x[i] /= A_(i, i);
        }
    }
}

/* trsm Left/Upper/Trans/NonUnit: trsv (RowMajor, Lower, NoTrans, NonUnit):
 * x_i -= sum_{k<i} U(k,i) x_k; x_i /= U(i,i). */
static void FN(trsm_lutn)(int m, int n, const REAL *a, int lda, REAL *b,
                          int ldb)
{
    for (int c = 0; c < n; ++c) {
        REAL *x = b + (size_t)c * ldb;
        for (int i = 0; i < m; ++i) {
            REAL s = 0;
            for (int k = 0; k < i; ++k) s += A_(k, i) * x[k];
            x[i] -= s;
            x[i] /= A_(i, i);
        }
    }
}

/* trsm Left/Lower/Trans/Unit: trsv (RowMajor, Upper, NoTrans, Unit):
 * x_i -= sum_{k>i} L(k,i) x_k, i descending. */
static void FN(trsm_lltu)(int m, int n, const REAL *a, int lda, REAL *b,
                          int ldb)
{
    for (int c = 0; c < n; ++c) {
        REAL *x = b + (size_t)c * ldb;
        for (int i = m - 1; i >= 0; --i) {
            REAL s = 0;
            for (int k = i + 1; k < m; ++k) s += A_(k, i) * x[k];
            x[i] -= s;
        }
    }
}

/*
 * GEMM NN with alpha=-1, beta=1 as called by trf.tcc:144-146; follows
 * cxxblas/level3/gemm.tcc:149-184 -> gemv.tcc:44-70 -> dot.tcc:42-52:
 *   C(l,i) = C(l,i)*1 ; s = sum_p A(l,p) B(p,i) (from 0, p ascending) ;
 *   C(l,i) += (-1)*s.
 */
static void FN(gemm_nn_minus)(int m, int n, int k, const REAL *a, int lda,
                              const REAL *b, int ldb, REAL *c, int ldc)
{
    for (int l = 0; l < m; ++l) {
        for (int i = 0; i < n; ++i) {
            REAL s = 0;
            for (int p = 0; p < k; ++p)
                s += b[(size_t)p + (size_t)i * ldb] * A_(l, p);
            c[(size_t)l + (size_t)i * ldc] += (REAL)(-1) * s;
        }
    }
}

/*
 * flens/lapack/ge/trf.tcc:59-152 (blocked right-looking LU, LAPACK 3.2
 * dgetrf).  nb plays the role of ilaenv(1,"GETRF") (64 in
 * flens/lapack/la/ilaenv.tcc:91-99).  ipiv is 1-based on output.
 */
int FN(getrf_gap)(int m, int n, REAL *a, int lda, int *ipiv, int nb,
                  double *gap)
{
    int info = 0;
    if (m == 0 || n == 0) return 0;
    const int mn = m < n ? m : n;
    int *piv0 = (int *)malloc(sizeof(int) * (size_t)mn);
    if (nb <= 1 || nb >= mn) {
        info = FN(tf2)(m, n, a, lda, piv0, gap);
    } else {
        for (int j = 0; j < mn; j += nb) {
            int jb = mn - j < nb ? mn - j : nb;
            int info_ = FN(tf2)(m - j, jb, &A_(j, j), lda, piv0 + j,
                                gap ? gap + j : (double *)0);
            if (info == 0 && info_ > 0) info = info_ + j;
            for (int i = j; i < (m < j + jb ? m : j + jb); ++i) piv0[i] += j;
            /* trf.tcc:126: columns 0..j-1 */
            FN(laswp_fwd)(j, a, lda, j, j + jb, piv0);
            if (j + jb < n) {
                /* trf.tcc:132 */
                FN(laswp_fwd)(n - j - jb, &A_(0, j + jb), lda, j, j + jb, piv0);
                /* trf.tcc:136-138 */
                FN(trsm_llnu)(jb, n - j - jb, &A_(j, j), lda, &A_(j, j + jb),
                              lda);
                /* trf.tcc:144-146 */
                if (j + jb < m)
                    FN(gemm_nn_minus)(m - j - jb, n - j - jb, jb,
                                      &A_(j + jb, j), lda, &A_(j, j + jb), lda,
                                      &A_(j + jb, j + jb), lda);
            }
        }
    }
    for (int i = 0; i < mn; ++i) ipiv[i] = piv0[i] + 1;
    free(piv0);
    return info;
}

int FN(getrf)(int m, int n, REAL *a, int lda, int *ipiv, int nb)
{
    return FN(getrf_gap)(m, n, a, lda, ipiv, nb, (double *)0);
}

/* The first nsteps block steps of getrf (trf.tcc:96-149), then stop: a
 * bounded sample of the same work, used to time the CPU baseline. */
void FN(getrf_steps)(int m, int n, REAL *a, int lda, int nb, int nsteps)
{
    const int mn = m < n ? m : n;
    int *piv0 = (int *)malloc(sizeof(int) * (size_t)mn);
    for (int s = 0, j = 0; s < nsteps && j < mn; ++s, j += nb) {
        int jb = mn - j < nb ? mn - j : nb;
        FN(tf2)(m - j, jb, &A_(j, j), lda, piv0 + j, (double *)0);
        for (int i = j; i < (m < j + jb ? m : j + jb); ++i) piv0[i] += j;
        FN(laswp_fwd)(j, a, lda, j, j + jb, piv0);
        if (j + jb < n) {
            FN(laswp_fwd)(n - j - jb, &A_(0, j + jb), lda, j, j + jb, piv0);
            FN(trsm_llnu)(jb, n - j - jb, &A_(j, j), lda, &A_(j, j + jb), lda);
            if (j + jb < m)
                FN(gemm_nn_minus)(m - j - jb, n - j - jb, jb, &A_(j + jb, j), lda,
                                  &A_(j, j + jb), lda, &A_(j + jb, j + jb), lda);
        }
    }
    free(piv0);
}

/* flens/lapack/ge/trs.tcc:59-109; trans: 0 = NoTrans, 1 = Trans. */
void FN(getrs)(int trans, int n, int nrhs, const REAL *a, int lda,
               const int *ipiv, REAL *b, int ldb)
{
    if (n == 0 || nrhs == 0) return;
    int *piv0 = (int *)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) piv0[i] = ipiv[i] - 1;
    if (!trans) {
        FN(laswp_fwd)(nrhs, b, ldb, 0, n, piv0);         /* trs.tcc:84 */
        FN(trsm_llnu)(n, nrhs, a, lda, b, ldb);          /* trs.tcc:88 */
        FN(trsm_lunn)(n, nrhs, a, lda, b, ldb);          /* trs.tcc:92 */
    } else {
        FN(trsm_lutn)(n, nrhs, a, lda, b, ldb);          /* trs.tcc:99 */
        FN(trsm_lltu)(n, nrhs, a, lda, b, ldb);          /* trs.tcc:103 */
        FN(laswp_bwd)(nrhs, b, ldb, 0, n, piv0);         /* trs.tcc:107 */
    }
    free(piv0);
}

/* flens/lapack/ge/sv.tcc:58-77: trf, then trs when info==0. */
int FN(gesv)(int n, int nrhs, REAL *a, int lda, int *ipiv, REAL *b, int ldb,
             int nb)
{
    int info = FN(getrf)(n, n, a, lda, ipiv, nb);
    if (info == 0) FN(getrs)(0, n, nrhs, a, lda, ipiv, b, ldb);
    return info;
}

#undef A_
