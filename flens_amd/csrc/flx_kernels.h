// flx_kernels.h -- internal launchers shared by the flx translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FLX_CHECK(x)                                                          \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) return e_;                                      \
    } while (0)

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
void prof_set_mask(Profiler *p, unsigned mask);
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
// max_groups > 0 caps the workgroups (each then loops over several tiles).
hipError_t gemm_f64(hipStream_t s, int ta, int tb, int M, int N, int K,
                    double alpha, const double *A, long lda, const double *B,
                    long ldb, double beta, double *C, long ldc, int blas_beta0,
                    int max_groups = 0);
hipError_t gemm_f32(hipStream_t s, int ta, int tb, int M, int N, int K,
                    float alpha, const float *A, long lda, const float *B,
                    long ldb, float beta, float *C, long ldc, int blas_beta0,
                    int max_groups = 0);

// ---- LU (getrf.hip) --------------------------------------------------------
// Device scratch owned by a handle and reused by every factorisation.
struct LuWork {
    uint64_t *exch = nullptr;   // panel all-gather slots
    size_t exch_words = 0;
    double *lsc = nullptr;      // panel multipliers (lu_lsc_doubles of the tallest panel)
    int *list = nullptr;        // row-permutation list of one sub-panel
    int *perm = nullptr;        // composed row permutation (length m)
    int perm_len = 0;
    int *flags = nullptr;       // [0] = sync failure flag
    unsigned long long *moved = nullptr; // [0] = row-interchange elements moved
    uint32_t gen = 0;           // all-gather generation counter
    Profiler *prof = nullptr;
    // lookahead (the rest of each trailing update on a second stream)
    int *inv = nullptr;         // inverse of perm before the current block (length m)
    int *mv = nullptr;          // block moves [count, dst, src, ...], 1 + 2 * LU_BLOCK_MOVES
    hipStream_t side = nullptr; // default priority, ordered by the events
    hipEvent_t ev_a = nullptr, ev_r = nullptr;
    hipEvent_t ev_n = nullptr;  // optional: the next block's columns done on the side stream
    // set by the caller when ev_r already marks side-stream work the first
    // block end must wait for (the working copy's columns right of the first
    // block, copied in there); cleared by getrf_rm_f64
    bool side_pending = false;
    // workgroups the side stream's launches may use at once, so that the
    // panels and the small launches of the next block always find free
    // compute units (a trailing update left uncapped fills the device)
    int side_groups = 0;
    // the same for the first blocks, whose trailing updates outweigh their
    // panels (more rows left than early_rows_frac of the matrix)
    int side_groups_early = 0;
    int cus = 0;                // compute units of the device
    // fused panel step (lookahead): staged rows (LU_STAGE_ROWS x LU_STAGE_LD
    // doubles), their destinations (+ the count at [LU_STAGE_TOP]) and perm
    // entries; null: separate interchange / solve / update launches
    double *stage = nullptr;
    int *sdst = nullptr;
    int *sperm = nullptr;       // 2 * LU_STAGE_TOP: moved rows' perm entries, then their sources
};
constexpr int LU_BLOCK_MOVES = 2 * 256;   // rows one outer block can move
constexpr int LU_STAGE_TOP = 128;         // stage rows: list sources, then the top rows
constexpr int LU_STAGE_ROWS = LU_STAGE_TOP + 32, LU_STAGE_LD = 256;
// after the stage rows: inv(L11) of the panel step (32 x 32) and its flag
constexpr int LU_STAGE_INV = LU_STAGE_ROWS * LU_STAGE_LD;
constexpr int LU_STAGE_DOUBLES = LU_STAGE_INV + 32 * 32 + 1;

struct LuOptions {
    int nb = 256;   // outer block (trailing update K)
    int ib = 32;    // panel width of the cooperative panel kernel
    int lookahead = 1;   // overlap the next block's panels with the rest of the update
};

size_t lu_exch_words();
// the outer block width getrf_rm_f64 uses for opt (nb clamped and rounded
// to a multiple of the panel width)
int lu_block_width(const LuOptions &opt);

// Panel kernel geometry: 512 rows per workgroup, at most 64 workgroups; the
// interchange list has one slice of 1 + 2*64 ints per workgroup:
// [count, dst_0, src_0, dst_1, src_1, ...] in absolute rows.
constexpr int PANEL_ROWS_PER_WG = 512;
constexpr int PANEL_LIST_SLICE = 1 + 2 * 64;
constexpr int PANEL_GROUPS_MAX = 64;
// largest matrix order of the FP64 LU and the paths built on it (panel
// heights up to PANEL_GROUPS_MAX x PANEL_ROWS_PER_WG rows)
constexpr int LU_N_MAX = 32768;
// doubles of the panel kernel's multiplier scratch (LuWork::lsc) for panels of
// up to m rows
inline size_t lu_lsc_doubles(int m) { return (size_t)(m > 0 ? m : 1) * 32; }
inline int panel_groups(int mp) { return (mp + PANEL_ROWS_PER_WG - 1) / PANEL_ROWS_PER_WG; }
// Panel factorisation of the mp x w row-major panel P (w <= 32); piv gets
// 0-based absolute pivot rows, the composed interchanges go to wk.list.
// defer_l: write only the pivot rows; the other rows' multipliers stay in
// wk.lsc for the fused panel step (lu_apply_kernel) to place
hipError_t launch_panel(hipStream_t s, LuWork &wk, double *P, long ldw, int mp, int w,
                        int *piv, int piv_base, int *info, int ib, int defer_l = 0);
// Factor the m x n row-major matrix W (leading dimension ldw) in place.
// piv0 receives 0-based absolute pivot rows (length min(m,n)); info receives
// the 1-based first zero pivot column (INT_MAX if none; caller initialises).
// Deferred interchanges of one outer block [J, J+jb): inv = inverse of perm
// before the block (inv_perm, which also clears the move count mv[0]);
// afterwards defer_moves moves the rows the block's sub-panels moved in its
// own columns in every other column too (mv holds 1 + 2 * LU_BLOCK_MOVES
// ints, mv[0] zero on entry).
hipError_t inv_perm(hipStream_t s, const int *perm, int *inv, int m, int *reset = nullptr);
template <typename T>
hipError_t defer_moves(hipStream_t s, Profiler *prof, T *W, long ldw, int m, int n, int J,
                       int jb, const int *perm, const int *inv, int *mv,
                       unsigned long long *moved);
hipError_t getrf_rm_f64(hipStream_t s, LuWork &wk, double *W, long ldw, int m,
                        int n, int *piv0, int *info, const LuOptions &opt);
// Row-major building blocks shared by the FP64 and the mixed-precision LU.
// X <- L^{-1} X with L the k x k unit lower triangle at L (k <= 256).
template <typename T>
hipError_t trsm_rm_llu(hipStream_t s, Profiler *prof, const T *L, long ldl, T *X, long ldx,
                       int k, int ncols, int max_groups = 0);
// Applies an interchange list (per-workgroup slices of PANEL_LIST_SLICE ints,
// G slices) to the columns [0, c1) and [c2, n) of the row-major W, and
// composes it into perm (when not null).
template <typename T>
hipError_t rowperm_rm(hipStream_t s, Profiler *prof, T *W, long ldw, int m, int c1, int c2,
                      int n, const int *list, int G, int *perm, unsigned long long *moved);
hipError_t iota(hipStream_t s, int *p, int n);
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
// Single-launch-per-triangle solve (see getrs.hip): X = U^{-1} L^{-1} Bp with
// Bp already row-permuted (trans = 0), or X = L^{-T} U^{-T} Bp (trans = 1,
// the row order applied afterwards); T is n x 16 scratch; xt holds
// trs_persist_words(n) tagged words (zeroed once), epoch is their host tag.
size_t trs_persist_words(int n);
// NoTrans solve U^{-1} L^{-1} Bp by the chain-and-helpers kernel: inv holds
// trs_inv_doubles(n) doubles, w trs_chain_words(n) tagged words (zeroed
// once), at most cus workgroups; lower = 0: Bp is L^{-1} P B already, only
// the U solve runs; rowmajor = 1 (with lower = 0): LU and Bp are row-major
// (the factorisation's working copy), X column-major.  inv_done (optional)
// is recorded on s once the U solve's diagonal inverses are launched.
size_t trs_chain_words(int n);
hipError_t trs_chain_f64(hipStream_t s, int n, int nrhs, const double *LU, long lda,
                         const double *Bp, long ldbp, double *T, double *X, long ldx,
                         double *inv, uint64_t *w, unsigned *epoch, int *err, int cus,
                         int lower = 1, int rowmajor = 0, hipEvent_t inv_done = nullptr);
hipError_t trs_persist_f64(hipStream_t s, int n, int nrhs, const double *LU, long lda,
                           const double *Bp, long ldbp, double *T, double *X, long ldx,
                           uint64_t *xt, unsigned *epoch, int *err, int trans = 0);
hipError_t laswp_cm_f64(hipStream_t s, double *A, long lda, int ncols, int k1,
                        int k2, const int *ipiv, int incx);

// ---- mixed-precision LU (mixed.hip) ----------------------------------------
enum MixPrec { MIX_FP32 = 0, MIX_BF16 = 1, MIX_FP16 = 2 };
// Device scratch of the mixed-precision factorisation (owned by a handle).
struct MixWork {
    int *perm = nullptr;          // n: original row at each position
    int *cand = nullptr;          // tslu_nodes(n) x 32 candidate positions
    int *ccount = nullptr;        // tslu_nodes(n)
    unsigned *counters = nullptr; // tslu_nodes(n), zero between launches
    float *u11 = nullptr;         // MIX_U11_FLOATS
    int *win = nullptr;           // 32
    int *list = nullptr;          // PANEL_LIST_SLICE
    int *zero = nullptr;          // first zero pivot column (INT_MAX: none)
    unsigned *lpmax = nullptr;    // 2 (FP16 operand maxima)
    void *la = nullptr, *ub = nullptr;   // lp_pack_elems(n, nb) 16-bit each
    Profiler *prof = nullptr;
    // lookahead, as in LuWork (all null: no lookahead)
    int *inv = nullptr, *mv = nullptr;
    hipStream_t side = nullptr;
    hipEvent_t ev_a = nullptr, ev_r = nullptr;
    int side_groups = 0;
    // with lookahead: the rows a panel step moves (MIX_STAGE_ROWS x
    // MIX_STAGE_LD floats) and their perm entries (MIX_STAGE_ROWS ints);
    // null: the panel step runs as separate finish / interchange / update
    float *stage = nullptr;
    int *stage_perm = nullptr;
};
constexpr int MIX_STAGE_ROWS = 64, MIX_STAGE_LD = 256;
// u11: L11\U11 (32 x 32), the strict triangles of U11 and L11 shifted
// (32 x 32 each), 1 / diag(U11)
constexpr int MIX_U11_FLOATS = 3 * 32 * 32 + 32;
size_t tslu_nodes(int mp);
size_t lp_pack_elems(int n, int nb);
// Factor the n x n row-major FP32 W in place (tournament-pivoted panels,
// trailing update in prec); mw.perm receives the row order.
hipError_t getrf_rm_mixed(hipStream_t s, const MixWork &mw, float *W, long ldw, int n,
                          int prec, int nb);
// W[i][j] = float(A(i,j) dr[i] dc[j]) (dr, dc optional); *flag |= 1 on overflow
hipError_t conv_d2s_rm(hipStream_t s, const double *A, long lda, int n, const double *dr,
                       const double *dc, float *W, long ldw, int *flag);
// Factor the n x n row-major FP32 W in place without pivoting (for a matrix
// made safe for it by random butterflies, rbt.hip); *mw.zero gets the first
// exact zero pivot column.
hipError_t getrf_rm_genp(hipStream_t s, const MixWork &mw, float *W, long ldw, int n,
                         int prec, int nb);

// ---- random butterfly transforms (rbt.hip) ----------------------------------
// Two depth-2 recursive butterflies U, V of order n (n % 4 == 0), each as its
// two level arrays of n doubles (1/sqrt(2) folded in).
struct Rbt {
    const double *u1 = nullptr, *u2 = nullptr, *v1 = nullptr, *v2 = nullptr;
};
bool rbt_supported(int n);
// n random butterfly values exp(r / 10) / sqrt(2), r uniform in [-1/2, 1/2)
void rbt_values(int n, unsigned long long seed, double *out);
// W = float(U^T D_r A D_c V) row-major (dr, dc optional); *flag |= 1 on overflow
hipError_t rbt_conv(hipStream_t s, const double *A, long lda, int n, const double *dr,
                    const double *dc, const Rbt &rb, float *W, long ldw, int *flag);
// y = U^T D_r v: to y32 scaled by 2^-e of *nrm (when nrm), else to y64
hipError_t rbt_in(hipStream_t s, const double *v, const double *dr,
                  const unsigned long long *nrm, const Rbt &rb, float *y32, double *y64, int n);
// out = (add ? out : 0) + D_c V y (y32 scaled back by 2^e of *nrm, or y64);
// ||out||_inf -> onrm when given
hipError_t rbt_out(hipStream_t s, const float *y32, const double *y64, const double *dc,
                   const unsigned long long *nrm, const Rbt &rb, double *out, int add, int n,
                   unsigned long long *onrm);

// ---- refinement (ir.hip) -------------------------------------------------------
// The FP32 factor as a preconditioner: M^{-1} v = D_c U^{-1} L^{-1} P D_r v,
// or with random butterflies M^{-1} v = D_c V U^{-1} L^{-1} U^T D_r v.
struct PrecFactor {
    int n = 0;
    const float *W = nullptr;  // row-major L\U
    long ldw = 0;
    const int *perm = nullptr;
    const double *dr = nullptr, *dc = nullptr;   // optional power-of-two scales
    bool rbt = false;          // factor of U^T D_r A D_c V (rb): no perm
    Rbt rb;
    float *y32 = nullptr;      // n
    double *y64 = nullptr;     // n (FP64 application)
    const double *dinv = nullptr;   // trsv_inv_doubles(n): inverted diagonal blocks of W
    uint64_t *xt = nullptr;    // trsv_words(n) tagged words, zero-initialised once
    unsigned *epoch = nullptr; // host counter paired with xt
    int *err = nullptr;        // device error flag
};
size_t ir_part_doubles(int n);
size_t trsv_words(int n);
size_t mdot_part_doubles(int n, int kmax);
hipError_t gemv_resid(hipStream_t s, const double *A, long lda, int n, const double *x,
                      const double *b, double *r, double *part, unsigned long long *nrm);
hipError_t norm_inf(hipStream_t s, const double *A, long lda, int n, double *part,
                    unsigned long long *nrm);
hipError_t equilibrate(hipStream_t s, const double *A, long lda, int n, double *part,
                       double *dr, double *dc);
size_t trsv_inv_doubles(int n);
// the inverted 64 x 64 diagonal blocks of both triangles of W (after every
// factorisation, before any trsv_rm on it)
hipError_t trsv_diag_inv(hipStream_t s, const float *W, long ldw, int n, double *dinv);
template <typename TY>
hipError_t trsv_rm(hipStream_t s, int upper, const float *W, long ldw, int n, const double *dinv,
                   TY *y, uint64_t *xt, unsigned epoch, int *err);
hipError_t zero_slots(hipStream_t s, unsigned long long *p, int k);
hipError_t vmax(hipStream_t s, const double *v, int n, unsigned long long *nrm);
hipError_t prec_apply(hipStream_t s, const PrecFactor &f, const double *v, double *z, int add,
                      unsigned long long *vnrm, unsigned long long *znrm);
hipError_t prec_apply64(hipStream_t s, const PrecFactor &f, const double *v, double *z, int add);
hipError_t mdot(hipStream_t s, const double *V, long ldv, int n, int k, const double *w,
                double *part, double *out);
hipError_t maxpy(hipStream_t s, const double *V, long ldv, int n, int k, const double *h, double *w);
hipError_t vscale_copy(hipStream_t s, const double *src, double alpha, double *dst, int n);
hipError_t fill_const(hipStream_t s, double *v, int n, double c);
hipError_t sign_of(hipStream_t s, const double *x, double *y, int n);

// ---- precision selector network (mlp.hip) ----------------------------------
constexpr int MLP_IN = 5, MLP_HIDDEN = 64, MLP_OUT = 8;
constexpr int MLP_PARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN +
                           MLP_HIDDEN + MLP_OUT * MLP_HIDDEN + MLP_OUT;
size_t mlp_part_floats(int nb);
hipError_t mlp_forward(hipStream_t s, const float *P, const float *X, int nb, float *logits);
hipError_t mlp_grad(hipStream_t s, const float *P, const float *X, const int *label,
                    const float *weight, int nb, float *part, float *grad, float *loss);
hipError_t mlp_adam(hipStream_t s, float *P, const float *grad, float *m, float *v, int step,
                    float lr, float wd);

// ---- batched solves and GEMMs (batched.hip) ----------------------------------
// Problem tables are built by the host in the layouts of batched.hip.
size_t batched_prob_bytes();
size_t batched_gemm_prob_bytes();
// Optional outputs: feat (count x MLP_IN selector features), tsys (count,
// in-kernel time per system in 100 MHz ticks), choice (count, the
// selector's class), inner (GMRES steps), used (the ladder | tile << 2 the
// system ran); sel = selector parameters (policy 1).
struct BatchOut {
    float *feat = nullptr;
    long long *tsys = nullptr;
    int *choice = nullptr;
    int *inner = nullptr;
    int *used = nullptr;
    const float *sel = nullptr;
    double *stats = nullptr;   // 3 per system: row statistics from a pre-pass kernel
    float *selT = nullptr;     // MLP_PARAMS floats: the selector transposed for the kernel
    int *cls = nullptr;        // per system: the selector's pre-factor class
};
// policy 0: each system's table ladder and tile; 1: the selector network;
// 2: the rule-based ladder.  G: per-workgroup scratch of the low-precision
// ladders (batched_scratch_doubles(grid) doubles; null: FP64 only).
hipError_t batched_gesv(hipStream_t s, int policy, const void *probs, int nprob, int *counter,
                        double *A, double *B, float *W32, double *R, double *G, int *ipiv,
                        int *info, int *iters, int itmax, int grid, const BatchOut &out);
size_t batched_scratch_doubles(int grid);
// GEMMs: each problem's tile shape in its table entry (cfg); tticks (may be
// null) accumulates every problem's in-kernel time by its id.
hipError_t batched_gemm(hipStream_t s, const void *probs, int nprob, int ntiles, int *counter,
                        const double *A, const double *B, double *C, int grid,
                        unsigned long long *tticks);
int batched_gemm_tile(int cfg, int *tm, int *tn);

// ---- utilities (util.hip) --------------------------------------------------
hipError_t fill_uniform_f64(hipStream_t s, double *A, int m, int n, long lda,
                            uint64_t seed, uint64_t offset);

} // namespace flx
