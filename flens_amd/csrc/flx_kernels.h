// flx_kernels.h -- internal launchers shared by the flx translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flx {

// ---- launch profiler -------------------------------------------------------
// When enabled, every launch of a profiled class is bracketed by HIP events
// on the stream it runs on; flops/bytes are the algorithmic amounts.
enum ProfClass {
    PROF_GEMM = 0,      // trailing-update / general GEMM
    PROF_PANEL = 1,     // cooperative panel factorisation
    PROF_ROWPERM = 2,   // row interchanges (laswp)
    PROF_TRSM_RM = 3,   // block-row triangular solve inside getrf
    PROF_TRANSPOSE = 4, // layout conversion at entry / exit
    PROF_TRS = 5,       // getrs triangular-solve blocks
    PROF_OTHER = 6,
    PROF_NCLS = 7
};

struct Profiler;
extern thread_local Profiler *tl_prof;   // profiler of the calling handle
Profiler *prof_new();
void prof_delete(Profiler *p);
void prof_reset(Profiler *p);
void prof_collect(Profiler *p, double *out);
// opens a record (returns nullptr when profiling is off)
void *prof_begin(Profiler *p, hipStream_t s, int cls, double flops, double bytes);
void prof_end(Profiler *p, hipStream_t s, void *rec);

struct ProfScope {
    Profiler *p;
    hipStream_t s;
    void *rec;
    ProfScope(Profiler *p_, hipStream_t s_, int cls, double flops, double bytes)
        : p(p_), s(s_), rec(prof_begin(p_, s_, cls, flops, bytes)) {}
    ~ProfScope() { if (rec) prof_end(p, s, rec); }
};

// ---- GEMM (gemm.hip) -------------------------------------------------------
// column-major C <- beta*C + alpha*op(A)*op(B); blas_beta0 selects the BLAS
// convention that beta == 0 never reads C (else the FLENS generic 0*C).
hipError_t gemm_f64(hipStream_t s, int ta, int tb, int M, int N, int K,
                    double alpha, const double *A, long lda, const double *B,
                    long ldb, double beta, double *C, long ldc, int blas_beta0);
hipError_t gemm_f32(hipStream_t s, int ta, int tb, int M, int N, int K,
                    float alpha, const float *A, long lda, const float *B,
                    long ldb, float beta, float *C, long ldc, int blas_beta0);

// ---- LU (getrf.hip) --------------------------------------------------------
// Device scratch owned by a handle and reused by every factorisation.
struct LuWork {
    uint64_t *exch = nullptr;   // panel all-gather slots
    size_t exch_words = 0;
    int *list = nullptr;        // row-permutation list of one sub-panel
    int *perm = nullptr;        // composed row permutation (length m)
    int perm_len = 0;
    int *flags = nullptr;       // [0] = sync failure flag
    unsigned long long *moved = nullptr; // [0] = row-interchange elements moved
    uint32_t gen = 0;           // all-gather generation counter
    Profiler *prof = nullptr;
};

struct LuOptions {
    int nb = 256;   // outer block (trailing update K)
    int ib = 32;    // panel width of the cooperative panel kernel
};

size_t lu_exch_words();
// Panel kernel geometry: 512 rows per workgroup, at most 32 workgroups; the
// interchange list has one slice of 1 + 2*64 ints per workgroup:
// [count, dst_0, src_0, dst_1, src_1, ...] in absolute rows.
constexpr int PANEL_ROWS_PER_WG = 512;
constexpr int PANEL_LIST_SLICE = 1 + 2 * 64;
constexpr int PANEL_GROUPS_MAX = 32;
inline int panel_groups(int mp) { return (mp + PANEL_ROWS_PER_WG - 1) / PANEL_ROWS_PER_WG; }
// Panel factorisation of the mp x w row-major panel P (w <= 32); piv gets
// 0-based absolute pivot rows, the composed interchanges go to wk.list.
hipError_t launch_panel(hipStream_t s, LuWork &wk, double *P, long ldw, int mp, int w,
                        int *piv, int piv_base, int *info, int ib);
// Factor the m x n row-major matrix W (leading dimension ldw) in place.
// piv0 receives 0-based absolute pivot rows (length min(m,n)); info receives
// the 1-based first zero pivot column (INT_MAX if none; caller initialises).
hipError_t getrf_rm_f64(hipStream_t s, LuWork &wk, double *W, long ldw, int m,
                        int n, int *piv0, int *info, const LuOptions &opt);
hipError_t transpose_f64(hipStream_t s, const double *in, long ldi, double *out,
                         long ldo, int rows_in, int cols_in);
hipError_t fill_int(hipStream_t s, int *p, int n, int v);
hipError_t piv0_to_ipiv(hipStream_t s, const int *piv0, int *ipiv, int n,
                        int *info);

// ---- solve (getrs.hip) -----------------------------------------------------
hipError_t ipiv_to_perm(hipStream_t s, const int *ipiv, int *perm, int n,
                        int reverse);
hipError_t gather_rows_cm(hipStream_t s, const double *B, long ldb,
                          double *out, long ldo, const int *perm, int n,
                          int nrhs);
hipError_t copy_cm(hipStream_t s, const double *in, long ldi, double *out,
                   long ldo, int m, int n);
// Triangular solve with the factors in LU; B (n x nrhs) is consumed (its
// rows are updated in place), the solution is written to X.  inv is device
// scratch of trs_inv_doubles(n) doubles (inverted diagonal blocks).
size_t trs_inv_doubles(int n);
hipError_t trsm_cm_f64(hipStream_t s, int uplo_upper, int trans, int unit,
                       int n, int nrhs, const double *LU, long lda, double *B,
                       long ldb, double *X, long ldx, double *inv);
hipError_t laswp_cm_f64(hipStream_t s, double *A, long lda, int ncols, int k1,
                        int k2, const int *ipiv, int incx);

// ---- utilities (util.hip) --------------------------------------------------
hipError_t fill_uniform_f64(hipStream_t s, double *A, int m, int n, long lda,
                            uint64_t seed, uint64_t offset);

} // namespace flx
