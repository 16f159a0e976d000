/*
 * flx.h -- C ABI of libflx, an MI355X-native (gfx950, HIP) dense GEMM and
 * LU-solve library that is a drop-in for FLENS's GEMM / LU path.
 *
 * Two layers, both plain C (pointers and sizes, no C++ or torch types):
 *
 *  1. The FLENS plugin seams.  FLENS reaches an external BLAS through the
 *     CBLAS symbols (cxxblas/drivers/cblas.h) when HAVE_CBLAS is set, and an
 *     external LAPACK through the Fortran symbols (cxxlapack/netlib/
 *     interface/lapack.in.h) when USE_CXXLAPACK + ALWAYS_USE_CXXLAPACK are
 *     set (flens/lapack/lapack.h:48-55).  libflx exports exactly those
 *     symbols with the same signatures, so unmodified FLENS code runs its
 *     GEMM, trf, trs and sv on the GPU by linking libflx instead of a CPU
 *     BLAS/LAPACK.  They take HOST pointers, stage through device memory,
 *     and are synchronous, like the CPU routines they replace.
 *
 *  2. flx_* device-pointer entry points: the same operations on data that
 *     already lives in device memory (column-major, caller-owned), queued
 *     on the handle's HIP stream.  This is the fast path.
 *
 * Conventions (identical to the reference): column-major storage with a
 * leading dimension, 1-based int32 pivot vectors ("row i was interchanged
 * with row ipiv[i]"), A overwritten by L\U with L's unit diagonal not
 * stored, B overwritten by X, INFO = 0 success, INFO = j > 0 first exact
 * zero pivot U(j,j) (the factorisation completes, dgesv_ skips the solve),
 * INFO = -i invalid i-th argument after calling xerbla_.
 */
#ifndef FLX_H
#define FLX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum flx_status {
    FLX_SUCCESS = 0,
    FLX_INVALID_ARG = 1,
    FLX_HIP_ERROR = 2,
    FLX_ALLOC_FAILED = 3,
    FLX_SYNC_TIMEOUT = 4,     /* a bounded device-side wait gave up */
    FLX_UNSUPPORTED = 5,
    FLX_SCRATCH_LIMIT = 6     /* larger than flx_dgetrf_noalloc's fixed scratch */
};

typedef struct flx_context *flx_handle;

/* Handle with its own scratch memory; bound to the current device and to
 * the null stream until flx_set_stream.  One handle per host thread. */
int flx_create(flx_handle *h);
int flx_destroy(flx_handle h);
int flx_set_stream(flx_handle h, void *hip_stream);
int flx_sync(flx_handle h);                /* waits, then reports device-side errors:
                                              FLX_SYNC_TIMEOUT when a bounded wait of
                                              a queued launch gave up since the last
                                              report (asynchronous entry points such
                                              as flx_dgetrs report it here or in the
                                              next synchronous entry point) */
const char *flx_status_string(int status);
/* LU blocking: nb = outer block (trailing-update K), ib = panel width
 * (<= 64).  The reference's equivalent knob is ilaenv(1,"GETRF") = 64,
 * flens/lapack/la/ilaenv.tcc:91-99. */
int flx_set_lu_blocking(flx_handle h, int nb, int ib);
/* Lookahead depth of the LU (0 or 1, default 1).  With 1 the row
 * interchanges of an outer block reach the columns outside it in one pass
 * after its last panel, the next block's columns are updated first, and the
 * rest of the trailing update runs on a second stream beside the next
 * block's panels.  Results are identical for both depths.  No counterpart in
 * the reference (flens/lapack/ge/trf.tcc:94-139 is the right-looking
 * blocked LU without lookahead). */
int flx_set_lu_lookahead(flx_handle h, int depth);

/* Workgroups the lookahead's second stream may occupy with the rest of a
 * trailing update, so that the next block's panels keep compute units:
 * which = 0 the FP64 LU (0: the default, three quarters of the compute
 * units, seven eighths for the first blocks while more than 0.8 of the rows
 * remain -- 0.6 at nb >= 256; another value holds for every block),
 * which = 1 the mixed-precision factor (0: the default, no cap).
 * Tuning only; results do not depend on it.  No counterpart in the
 * reference. */
int flx_set_lu_side_groups(flx_handle h, int which, int groups);

/* Test hook for the error path: raises the handle's bounded-wait failure
 * flag on its stream, exactly as a device-side wait that gave up would, so
 * tests can check which call reports FLX_SYNC_TIMEOUT without making a
 * kernel time out.  Every synchronous entry point (and flx_sync) reports a
 * raised flag as FLX_SYNC_TIMEOUT and re-arms it; the FLENS seams abort
 * with that status.  No counterpart in the reference. */
int flx_test_raise_sync_flag(flx_handle h);

/* Launch profiler (no counterpart in the reference, which has no timers:
 * flens/debug only logs calls).  Between start and stop every library
 * launch of a profiled kernel class is bracketed by HIP events on the
 * stream it runs on.  stop() synchronises and writes FLX_PROF_NCLS x 4
 * doubles: {launches, milliseconds, algorithmic flops, algorithmic bytes}
 * per class, classes in the order of enum flx_prof_class. */
enum flx_prof_class {
    FLX_PROF_GEMM = 0, FLX_PROF_PANEL = 1, FLX_PROF_ROWPERM = 2,
    FLX_PROF_TRSM_RM = 3, FLX_PROF_TRANSPOSE = 4, FLX_PROF_TRS = 5,
    FLX_PROF_OTHER = 6, FLX_PROF_NCLS = 7
};
int flx_profile_start(flx_handle h);
int flx_profile_stop(flx_handle h, double *out);
/* After flx_profile_start: record only the classes whose bit (1 << class)
 * is set in mask, so a timed region can carry events on one kernel class
 * without the cost of events on every launch. */
int flx_profile_classes(flx_handle h, unsigned mask);

/* On-device peak probes (no counterpart in the reference: measurement
 * utilities for roofline denominators).  flx_probe_mfma times a grid that
 * keeps every SIMD issuing independent MFMA chains of one shape -- kind 0
 * v_mfma_f64_16x16x4f64, 1 v_mfma_f32_16x16x4f32, 2
 * v_mfma_f32_16x16x32_bf16, 3 v_mfma_f32_16x16x32_f16 -- for iters
 * iterations (8 instructions per wave each) and reports dense TFLOP/s, best
 * of 3 after one untimed run.  flx_probe_copy streams bytes (a multiple of
 * 32, both pointers 32-byte aligned device memory) from src to dst and
 * reports read + write GB/s, best of 5.  Both synchronise the handle's
 * stream. */
int flx_probe_mfma(flx_handle h, int kind, int iters, double *tflops);
int flx_probe_copy(flx_handle h, const void *src, void *dst, size_t bytes, double *gbps);

/* ------------------------------------------------- device-pointer entries */
/* C <- beta*C + alpha*op(A)*op(B), trans = 0 NoTrans / 1 Trans.
 * Replaces cxxblas::gemm (cxxblas/level3/gemm.tcc:149-299, reached from
 * flens::blas::mm, flens/blas/level3/mm.tcc:380-479).  Like the reference's
 * generic path, beta is applied to C first, so beta = 0 computes 0*C. */
int flx_dgemm(flx_handle h, int transa, int transb, int m, int n, int k,
              double alpha, const double *A, int lda, const double *B,
              int ldb, double beta, double *C, int ldc);

/* LU with partial pivoting in place; ipiv (min(m,n) ints) and info (one
 * int) are device pointers.  Replaces flens::lapack::trf
 * (flens/lapack/ge/trf.tcc:183-261).  m <= 32768 (FLX_UNSUPPORTED above:
 * the panel's cooperating workgroups hold 512 rows each, at most 64); the
 * same limit holds for the solvers and estimators built on it. */
int flx_dgetrf(flx_handle h, int m, int n, double *A, int lda, int *ipiv,
               int *info);

/* Solve op(A) X = B with the factors of flx_dgetrf; trans 0/1.  Replaces
 * flens::lapack::trs (flens/lapack/ge/trs.tcc:151-252). */
int flx_dgetrs(flx_handle h, int trans, int n, int nrhs, const double *A,
               int lda, const int *ipiv, double *B, int ldb);

/* trf + trs; *info_host (host int) receives INFO and the solve is skipped
 * when INFO > 0 (B left as it was), as in flens::lapack::sv
 * (flens/lapack/ge/sv.tcc:58-77).  The factors and pivots are flx_dgetrf's;
 * B is carried through the factorisation as extra columns (L^{-1} P B from
 * its own interchanges and block updates), then the U solve of flx_dgetrs
 * runs, so X agrees with trf + trs to rounding (not bit for bit).  This
 * entry synchronises the stream to read INFO. */
int flx_dgesv(flx_handle h, int n, int nrhs, double *A, int lda, int *ipiv,
              double *B, int ldb, int *info_host);

/* flx_dgetrf without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments, kernels and results as flx_dgetrf (run without lookahead; the
 * factors and pivots do not depend on it); asynchronous on the handle's
 * stream.  Its scratch is one static device array of
 * FLX_NOALLOC_SCRATCH_MIB MiB per device (held until the process ends):
 * flx_noalloc_bytes(m, n) bytes are needed -- about 8 m ceil32(n) bytes, so
 * square matrices up to n = 2160 -- and larger shapes return
 * FLX_SCRATCH_LIMIT without launching anything.  All calls on one device
 * share that array: calls that overlap on a device race.  *info (device) =
 * INT_MIN when a bounded device-side wait of the panel gave up. */
enum { FLX_NOALLOC_SCRATCH_MIB = 36 };
int flx_dgetrf_noalloc(flx_handle h, int m, int n, double *A, int lda, int *ipiv,
                       int *info);
size_t flx_noalloc_bytes(int m, int n);

/* Row interchanges k1..k2 (1-based), incx = +1 / -1.  Replaces
 * flens::lapack::laswp (flens/lapack/la/laswp.tcc:126-186). */
int flx_dlaswp(flx_handle h, int n, double *A, int lda, int k1, int k2,
               const int *ipiv, int incx);

/* Fill an m x n column-major block with the library's seeded uniform
 * [-1,1) stream (element (i,j) = stream[offset + j*m + i]). */
int flx_fill_uniform(flx_handle h, double *A, int m, int n, int lda,
                     uint64_t seed, uint64_t offset);

/* ------------------------------------------- mixed-precision solvers */
/* Mixed-precision solve of A X = B (device pointers, column-major; A and B
 * are not modified unless the FP64 fallback runs, which overwrites A with
 * its FP64 factors).  The factor is computed in FP32 with tournament
 * pivoting and a trailing update in the chosen precision; FP64 is used for
 * residuals, updates and norms.  Replaces the mixed-precision route of the
 * reference, cxxlapack::sgesv -> dsgesv_ (cxxlapack/netlib/interface/
 * lapack.in.h:10611-10623, netlib dsgesv.f:131-337), and extends it with
 * BF16/FP16 factors and GMRES-IR. */
enum flx_mixed_prec { FLX_MIXED_FP32 = 0, FLX_MIXED_BF16 = 1, FLX_MIXED_FP16 = 2 };
/* FLX_REFINE_AUTO: after the factor, kappa_inf is estimated from two solves
 * with it (||A||_inf times a lower bound of ||A^{-1}||_inf); with u the
 * factor's unit roundoff, kappa u < 2 -> classic (at least 60 steps),
 * < 100 -> GMRES-IR, else the FP64 LU at once (stats.iter = -4). */
enum flx_refine { FLX_REFINE_CLASSIC = 0, FLX_REFINE_GMRES = 1, FLX_REFINE_AUTO = 2 };
/* Pivoting of the low-precision factor.  TOURNAMENT: communication-avoiding
 * partial pivoting (CALU) of each 32-column panel.  RBT: the factor of
 * U^T A V for two random recursive butterflies U, V of depth 2 (fixed seeds),
 * without pivoting -- panels with no pivot search, the preconditioner
 * applied as V (LU)^{-1} U^T; needs n % 4 == 0, else TOURNAMENT is used.
 * Either way refinement works on A itself and the FP64 LU is the fallback. */
enum flx_pivot { FLX_PIVOT_TOURNAMENT = 0, FLX_PIVOT_RBT = 1 };
typedef struct flx_mixed_opts {
    int prec;          /* enum flx_mixed_prec (FP16 adds power-of-two row and
                          column equilibration and per-step operand scaling) */
    int refine;        /* enum flx_refine */
    int itermax;       /* outer refinement steps (30 in dsgesv) */
    int gmres_m;       /* Krylov vectors per GMRES solve (0: 40) */
    double tol;        /* stop at ||r||/(||A|| ||x|| + ||b||) <= tol
                          (0: sqrt(n) 2^-53, the level of dsgesv's test) */
    double gmres_tol;  /* relative GMRES residual (0: 1e-10) */
    int fallback;      /* 1: FP64 LU + solve when refinement fails */
    int nb;            /* outer block of the factorisation (0: 256) */
    int graph;         /* 1: replay the factorisation's launches from a
                          captured HIP graph (per shape), 0: launch eagerly */
    int pivot;         /* enum flx_pivot (0: tournament) */
} flx_mixed_opts;
typedef struct flx_mixed_stats {
    int iter;              /* refinement steps; < 0: failed (dsgesv ITER codes) */
    int inner_iters;       /* GMRES iterations in total */
    int fallback;          /* 1 when the FP64 fallback produced X */
    int info;              /* INFO of the FP64 fallback factorisation */
    double backward_error; /* max over columns, inf-norm, at exit */
    double anorm;          /* ||A||_inf */
    double kappa_est;      /* condition estimate (FLX_REFINE_AUTO), else 0 */
    double factor_ms, refine_ms, fallback_ms;
} flx_mixed_stats;
int flx_dgesv_mixed(flx_handle h, int n, int nrhs, double *A, int lda,
                    const double *B, int ldb, double *X, int ldx,
                    const flx_mixed_opts *opt, flx_mixed_stats *stats);

/* dsgesv semantics on device pointers: FP32 factor + classic refinement
 * (ITERMAX 30, stop when ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)),
 * *iter / *info (host ints) as LAPACK's ITER / INFO: ITER = -2 FP32
 * overflow, -3 zero pivot in the FP32 factor, -31 no convergence, each
 * followed by the FP64 LU + solve.  ipiv (device, n ints) receives the
 * interchanges of the factor that produced X. */
int flx_dsgesv(flx_handle h, int n, int nrhs, double *A, int lda, int *ipiv,
               const double *B, int ldb, double *X, int ldx, int *iter,
               int *info);

/* ------------------- norms, equilibration, condition, refinement -----
 * Device pointers for matrices and vectors; scalar results (value, rowcnd,
 * rcond, ferr, berr, ...) are written to host memory and are valid when the
 * call returns (these routines synchronise the handle's stream). */

/* ||A|| for norm 'M' (max |a_ij|), 'O' or '1' (max column sum), 'I' (max
 * row sum), 'F' (Frobenius).  Replaces flens::lapack::lan for GeMatrix
 * (flens/lapack/la/lan.tcc:55-130).  'O' and 'I' sum in the reference's
 * order (bit-identical); 'M' is exact; NaN entries are skipped by 'M', 'O'
 * and 'I' as the reference's std::max does, and propagate through 'F'. */
int flx_dlange(flx_handle h, char norm, int m, int n, const double *A, int lda,
               double *value);

/* Row and column scalings r (m) and c (n, device) that equilibrate A, with
 * *rowcnd, *colcnd, *amax (host) and *info (host: 0, i <= m for a zero row
 * i, m + j for a zero column j).  Replaces flens::lapack::equ
 * (flens/lapack/ge/equ.tcc:62-194); results bit-identical. */
int flx_dgeequ(flx_handle h, int m, int n, const double *A, int lda, double *r,
               double *c, double *rowcnd, double *colcnd, double *amax,
               int *info);

/* Applies the scalings when they help (thresholds 0.1 and the safe range):
 * *equed (host) = 'N', 'R', 'C' or 'B'.  Replaces flens::lapack::laq
 * (flens/lapack/la/laq.tcc:57-118). */
int flx_dlaqge(flx_handle h, int m, int n, double *A, int lda, const double *r,
               const double *c, double rowcnd, double colcnd, double amax,
               char *equed);

/* Reciprocal condition number estimate of A from its LU factor (as
 * flx_dgetrf leaves it) and ||A|| in norm 'O' ('1') or 'I'.  Replaces
 * flens::lapack::con (flens/lapack/ge/cond.tcc:62-165) with lacn2
 * (flens/lapack/la/lacn2.tcc:55-185); the triangular solves are the LU's,
 * i.e. latrs (flens/lapack/la/latrs.tcc) without its rescaling, so a
 * solution growing toward the overflow threshold yields *rcond = 0. */
int flx_dgecon(flx_handle h, char norm, int n, const double *LU, int lda,
               double anorm, double *rcond);

/* Iterative refinement of X (n x nrhs, device) for op(A) X = B with
 * forward and backward error bounds ferr, berr (nrhs, host).  AF, ipiv:
 * the factor of A (device).  Replaces flens::lapack::rfs
 * (flens/lapack/ge/rfs.tcc:63-261): at most 5 steps per right-hand side,
 * stop when berr <= eps or it stops halving. */
int flx_dgerfs(flx_handle h, int trans, int n, int nrhs, const double *A,
               int lda, const double *AF, int ldaf, const int *ipiv,
               const double *B, int ldb, double *X, int ldx, double *ferr,
               double *berr);

/* Expert driver: fact 'N' (factor A), 'E' (equilibrate, then factor) or
 * 'F' (AF, ipiv and, per *equed, r / c given); solves op(A) X = B, with
 * *rcond, ferr, berr, *rpvgrw (reciprocal pivot growth, LAPACK's WORK(1))
 * and *info (0; i <= n: U(i,i) = 0; n + 1: rcond < eps) on the host.
 * Replaces flens::lapack::svx (flens/lapack/ge/svx.tcc:60-262); unlike the
 * reference, which takes EQUED by value, *equed reports the scaling used. */
int flx_dgesvx(flx_handle h, char fact, int trans, int n, int nrhs, double *A,
               int lda, double *AF, int ldaf, int *ipiv, char *equed,
               double *r, double *c, double *B, int ldb, double *X, int ldx,
               double *rcond, double *ferr, double *berr, double *rpvgrw,
               int *info);

/* inv(A) in place from the LU factor (A, ipiv as flx_dgetrf leaves them);
 * *info (host) = i > 0 when U(i,i) = 0, A then untouched.  Replaces
 * flens::lapack::tri for GeMatrix (flens/lapack/ge/tri.tcc:64-187, with
 * trtri flens/lapack/tr/tri.tcc); triangular inverses by recursive
 * doubling on DGEMM, inv(A) = inv(U) inv(L) P^T. */
int flx_dgetri(flx_handle h, int n, double *A, int lda, const int *ipiv, int *info);

/* ------------------------------------------------------------------- io */
/* FLENS's binary files (flens::save / flens::load, IndexType int):
 * GeMatrix flens/io/fullstorage/save.tcc:49-82 + load.tcc:52-80 -- int32
 * numRows, numCols, firstRow, firstCol, then the elements row by row;
 * DenseVector flens/io/array/save.tcc:43-66 + load.tcc:43-68 -- int32
 * length, firstIndex, then the elements.  Matrices and vectors are on the
 * device (column-major matrices); the row order is produced on the device.
 * FLENS's default first index is 1. */
int flx_save_gematrix(flx_handle h, const char *path, int m, int n, const double *A,
                      int lda, int first_row, int first_col);
int flx_load_gematrix_header(const char *path, int *m, int *n, int *first_row,
                             int *first_col);
int flx_load_gematrix(flx_handle h, const char *path, double *A, int lda);
/* elem = 8 for double, 4 for int (pivot vectors) */
int flx_save_densevector(flx_handle h, const char *path, int n, const void *x, int elem,
                         int first_index);
int flx_load_densevector(flx_handle h, const char *path, void *x, int capacity, int elem,
                         int *n, int *first_index);

/* ------------------------------------------------------ batched plans */
/* Thousands of independent small problems in one persistent launch each
 * (no counterpart in the reference, whose trf / trs / mm handle one matrix
 * per call: flens/lapack/ge/trf.tcc:183-261, trs.tcc:151-252,
 * flens/blas/level3/mm.tcc:380-479; the batch applies those routines'
 * semantics to every problem).  Inputs are packed column-major, ld = rows:
 *   solves: A_i (n_i x n_i) at offset sum_{j<i} n_j^2, B_i (n_i x nrhs_i)
 *           at sum_{j<i} n_j nrhs_j (overwritten by X_i), ipiv_i at
 *           sum_{j<i} n_j; 1 <= n_i <= 512, 0 <= nrhs_i <= 8; info[i] as
 *           LAPACK's INFO; the mode (FLX_BATCH_*) sets each system's
 *           precision ladder; iters[i] (optional) = refinement steps or a
 *           dsgesv-style negative code (-2 overflow in the low format, -3
 *           singular low-precision factor, -4 sent to FP64 by the selector
 *           or the rule, -itmax-1 no convergence), the system then solved
 *           by the FP64 LU.  Refinement stops at ||r_c||_inf <= sqrt(n)
 *           2^-53 (||A||_inf ||x_c||_inf + ||b_c||_inf) for every column.
 *   GEMMs:  C_i <- beta_i C_i + alpha_i A_i B_i (beta applied to C first,
 *           as the reference's generic gemm), A_i m_i x k_i at
 *           sum m_j k_j, B_i k_i x n_i at sum k_j n_j, C_i at sum m_j n_j.
 * A plan keeps the device pointers it was created with, picks its launch
 * shape (GEMM tile and persistent grid) by timing the candidates on copies
 * of the inputs when tune = 1, and captures the batch into a HIP graph;
 * flx_batch_execute replays it on the handle's stream. */
typedef struct flx_batch *flx_batch_t;
/* Solve plan modes.  The ladders: FP64 LU; FP32 LU + classic refinement
 * (FP64 residuals); BF16-operand LU + GMRES-IR; FP16-operand LU of the
 * row-scaled matrix (rows to 0.1 x the largest FP16 value) + GMRES-IR
 * (FP64 GMRES vectors, the low-precision factor as preconditioner).  In
 * the BF16 / FP16 LU the trailing updates multiply operands rounded to the
 * format with FP32 accumulation; factors are kept in FP32.
 *   FLX_BATCH_FP64, _FP32_IR, _BF16_GMRES, _FP16_GMRES: one ladder for all;
 *   FLX_BATCH_LEARNED: the selector network (extra->selector) reads each
 *     system's features before any factor -- log2 n / 9, nrhs / 8, the
 *     condition feature (-0.25: not yet estimated), log10 of the smallest
 *     diagonal-dominance ratio / 4, log10 of the max/min row-norm ratio /
 *     12 -- and picks a class = ladder + 4 x tile; after a low-precision
 *     factor it reads them again with log10 kappa_inf / 16 (kappa_inf =
 *     ||A||_inf times a lower bound of ||A^{-1}||_inf from two solves with
 *     the factor) and sends the system to the FP64 LU when it then picks
 *     the FP64 ladder;
 *   FLX_BATCH_RULE: the rule-based ladder: FP32 LU, classic refinement
 *     while kappa_inf <= 1e6, else FP64;
 *   FLX_BATCH_TABLE: extra->cfg[i] = ladder + 4 x tile per system.
 * Every failed or skipped refinement ends in the FP64 LU of A.  tile 0 /
 * 1: the trailing update works in strips of two / one 16-column tiles per
 * wave; with tune = 1 the autotuner picks it per size bin (n in [8 2^b,
 * 16 2^b), b = 0..5) from the in-kernel time of every system under both,
 * except where the selector or the table sets it. */
enum { FLX_BATCH_FP64 = 0, FLX_BATCH_FP32_IR = 1, FLX_BATCH_LEARNED = 2,
       FLX_BATCH_BF16_GMRES = 3, FLX_BATCH_FP16_GMRES = 4, FLX_BATCH_RULE = 5,
       FLX_BATCH_TABLE = 6 };
/* Optional per-system inputs / outputs of a solve plan (device pointers,
 * any may be NULL): the selector's parameters (FLX_BATCH_LEARNED), the
 * features as last evaluated (count x FLX_MLP_IN), the in-kernel time of
 * every system in 100 MHz ticks, the selector's pre-factor class, the
 * configuration table (FLX_BATCH_TABLE), the GMRES steps taken and the
 * ladder + 4 x tile each system ran. */
typedef struct flx_batch_extra {
    const float *selector;
    float *features;
    long long *ticks;
    int *choice;
    const int *cfg;
    int *inner;
    int *used;
} flx_batch_extra;
int flx_batch_gesv_create(flx_handle h, flx_batch_t *plan, int count,
                          const int *n, const int *nrhs, double *A,
                          double *B, int *ipiv, int *info, int *iters,
                          int mode, int itmax, int tune,
                          const flx_batch_extra *extra);
int flx_batch_gemm_create(flx_handle h, flx_batch_t *plan, int count,
                          const int *m, const int *n, const int *k,
                          const double *alpha, const double *beta,
                          const double *A, const double *B, double *C,
                          int tune);
int flx_batch_execute(flx_batch_t plan);
/* cfg[0] = GEMM tile shape (0: 32x32, 1: 64x64, 2: 128x64; -1: chosen
 * per size bin, see flx_batch_bin_tiles), cfg[1] = grid; *tuned_ms = the
 * winning grid's time (0 when not tuned).  With tune = 1 a GEMM plan times
 * every problem under each shape (per-tile in-kernel time) and gives each
 * size bin of max(m, n) -- [16 2^b, 32 2^b), b = 0..5 -- its fastest shape. */
int flx_batch_config(flx_batch_t plan, int *cfg, double *tuned_ms);
/* The tile chosen for each of the nbins >= 6 size bins (solve plans: the
 * LU strip width per n bin; GEMM plans: the shape per max(m, n) bin);
 * returns the number of bins (6). */
int flx_batch_bin_tiles(flx_batch_t plan, int *tiles, int nbins);
int flx_batch_destroy(flx_batch_t plan);

/* ------------------------------------------- precision selector network */
/* FP32 MLP  in (FLX_MLP_IN) -> 64 ReLU -> 64 ReLU -> FLX_MLP_OUT logits
 * (class = ladder + 4 x tile, ladders {FP64, FP32+IR, BF16+GMRES-IR,
 * FP16-scaled+GMRES-IR}, tiles as in FLX_BATCH_*), parameters in one
 * flat array: W1[64][IN], b1[64], W2[64][64], b2[64], W3[OUT][64], b3[OUT].
 * Forward, mean softmax cross-entropy gradient (deterministic) and an Adam
 * step, on device pointers.  No counterpart in the reference. */
enum { FLX_MLP_IN = 5, FLX_MLP_HIDDEN = 64, FLX_MLP_OUT = 8, FLX_MLP_PARAMS = 5064 };
int flx_mlp_forward(flx_handle h, const float *params, const float *X, int nb,
                    float *logits);
int flx_mlp_grad(flx_handle h, const float *params, const float *X,
                 const int *label, const float *weight, int nb, float *grad,
                 float *loss);
int flx_mlp_adam(flx_handle h, float *params, const float *grad, float *m,
                 float *v, int step, float lr, float weight_decay);

/* --------------------------------------------------- FLENS plugin seams */
/* CBLAS seam, cxxblas/drivers/cblas.h:798-805; enum values :22-27
 * (Order 101/102, Trans 111/112/113).  Caller: cxxblas::gemm<int> for
 * double, cxxblas/level3/gemm.tcc:327-349.  BLAS convention: beta = 0
 * does not read C. */
void cblas_dgemm(int order, int transA, int transB, int M, int N, int K,
                 double alpha, const double *A, int lda, const double *B,
                 int ldb, double beta, double *C, int ldc);

/* CBLAS seam, cxxblas/drivers/cblas.h:1044-1049 (Side 141/142, Uplo
 * 121/122, Diag 131/132).  Caller: cxxblas::trsm, cxxblas/level3/
 * trsm.tcc:107-126, reached from blas::sm (flens/blas/level3/sm.tcc:63). */
void cblas_dtrsm(int order, int side, int uplo, int transA, int diag, int M,
                 int N, double alpha, const double *A, int lda, double *B,
                 int ldb);

/* Fortran LAPACK seam, cxxlapack/netlib/interface/lapack.in.h:6751-6756.
 * Caller: cxxlapack::getrf (cxxlapack/interface/getrf.tcc:68-90) from
 * flens::lapack::external::trf_impl (flens/lapack/ge/trf.tcc:164-173);
 * same shape as FLENS's own export flens/lapack/interface/src/getrf.cc:9-45. */
void dgetrf_(const int *M, const int *N, double *A, const int *LDA, int *IPIV,
             int *INFO);

/* lapack.in.h:6770-6778; caller external::trs_impl (trs.tcc:124-140);
 * FLENS export flens/lapack/interface/src/getrs.cc:12-53. */
void dgetrs_(const char *TRANS, const int *N, const int *NRHS, const double *A,
             const int *LDA, const int *IPIV, double *B, const int *LDB,
             int *INFO);

/* lapack.in.h:6633-6640; caller external::sv_impl (sv.tcc:92-105);
 * FLENS export flens/lapack/interface/src/gesv.cc:12-49. */
void dgesv_(const int *N, const int *NRHS, double *A, const int *LDA,
            int *IPIV, double *B, const int *LDB, int *INFO);

/* lapack.in.h:9333-9339; caller external::laswp_impl
 * (flens/lapack/la/laswp.tcc:105-118). */
void dlaswp_(const int *N, double *A, const int *LDA, const int *K1,
             const int *K2, const int *IPIV, const int *INCX);

/* Fortran-LAPACK seam for the mixed-precision solver,
 * cxxlapack/netlib/interface/lapack.in.h:10611-10623.  WORK and SWORK are
 * accepted and unused (the library keeps its own device workspace). */
void dsgesv_(const int *N, const int *NRHS, double *A, const int *LDA,
             int *IPIV, double *B, const int *LDB, double *X, const int *LDX,
             double *WORK, float *SWORK, int *ITER, int *INFO);

/* Fortran BLAS GEMM with the conventions of FLENS's own export
 * (flens/blas/interface/blas/gemm.cc:121-229): info 1, 2, 3, 4, 5, 8, 10, 13
 * through xerbla_, C scaled by beta first (beta = 0 computes 0 * C, as the
 * generic cxxblas::gemm, cxxblas/level3/gemm.tcc:175). */
void dgemm_(const char *TRANSA, const char *TRANSB, const int *M, const int *N,
            const int *K, const double *ALPHA, const double *A, const int *LDA,
            const double *B, const int *LDB, const double *BETA, double *C,
            const int *LDC);

/* Fortran seams of the routines above (host pointers; WORK / IWORK are
 * accepted and unused except dgesvx_'s WORK(1) = reciprocal pivot growth):
 * cxxlapack/netlib/interface/lapack.in.h:8145-8151 (dlange), 6155-6166
 * (dgeequ), 6143-6153 (dgecon), 6535-6553 (dgerfs), 6676-6699 (dgesvx). */
double dlange_(const char *NORM, const int *M, const int *N, const double *A,
               const int *LDA, double *WORK);
void dgeequ_(const int *M, const int *N, const double *A, const int *LDA,
             double *R, double *C, double *ROWCND, double *COLCND,
             double *AMAX, int *INFO);
void dgecon_(const char *NORM, const int *N, const double *A, const int *LDA,
             const double *ANORM, double *RCOND, double *WORK, int *IWORK,
             int *INFO);
void dgerfs_(const char *TRANS, const int *N, const int *NRHS, const double *A,
             const int *LDA, const double *AF, const int *LDAF,
             const int *IPIV, const double *B, const int *LDB, double *X,
             const int *LDX, double *FERR, double *BERR, double *WORK,
             int *IWORK, int *INFO);
void dgesvx_(const char *FACT, const char *TRANS, const int *N,
             const int *NRHS, double *A, const int *LDA, double *AF,
             const int *LDAF, int *IPIV, char *EQUED, double *R, double *C,
             double *B, const int *LDB, double *X, const int *LDX,
             double *RCOND, double *FERR, double *BERR, double *WORK,
             int *IWORK, int *INFO);

/* lapack.in.h:6758-6766.  LWORK = -1 is a workspace query (WORK(1) =
 * 64 n, the reference's n * ilaenv(GETRI)); WORK is otherwise unused. */
void dgetri_(const int *N, double *A, const int *LDA, const int *IPIV,
             double *WORK, const int *LWORK, int *INFO);

/* Error hook called before a negative INFO is returned, as the reference's
 * LAPACK_ERROR (flens/lapack/interface/include/config.h:40-45).  libflx
 * provides a weak default that prints the LAPACK message to stderr. */
void xerbla_(const char *SRNAME, const int *INFO, int SRNAME_LEN);

#ifdef __cplusplus
}
#endif

#endif /* FLX_H */
