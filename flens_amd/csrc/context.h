// context.h -- the handle behind flx_handle and its device buffers (internal).
#pragma once
#include <climits>
#include "../../include/flx.h"
#include "flx_kernels.h"

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

constexpr int MIX_GKEY = 23;   // values keying the captured factorisation

// buffers of the mixed-precision solver
struct MixBufs {
    DevBuf W32, perm, cand, ccount, counters, u11, win, list, zero, lpmax, la, ub;
    DevBuf inv, mv;                         // lookahead of the factor
    DevBuf stage, stage_perm;               // rows that move in a panel step
    hipStream_t side = nullptr;             // default priority, ordered by events
    hipEvent_t ev_a = nullptr, ev_r = nullptr;
    DevBuf dinv, y32, y64, flags, err, dr, dc, part, scal, R, V, w, t, zeros, mpart, hbuf;
    DevBuf F;                               // FP64 fallback factors (A kept for the residual)
    DevBuf rbt;                             // random butterflies of order rbt_n (4 x n doubles)
    int rbt_n = 0;
    unsigned epoch = 0;
    // captured factorisation: valid while the key matches
    hipGraphExec_t graph = nullptr;
    hipStream_t gstream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    size_t gkey[MIX_GKEY] = {};
};

// buffers of the norm / equilibration / condition / refinement routines
struct AuxBufs {
    DevBuf slot, v, x, r, w, lines, isgn, part, ipiv1, perm0, pf, pb;
    int ident_n = 0;   // length of the identity ipiv1 / perm0 held
};

struct flx_context {
    hipStream_t stream = nullptr;
    int device = 0;
    flx::LuWork wk;
    flx::LuOptions opt;
    DevBuf W, A, B, C, X, piv0, ipiv, info, perm, tinv;
    DevBuf sb, sx;   // staging of the dsgesv_ seam (and AF, X of dgerfs_ / dgesvx_)
    DevBuf sr, sc;   // staging of the equilibration scales (dgeequ_, dgesvx_)
    DevBuf mlp_part; // per-block gradients of flx_mlp_grad
    DevBuf trsw;     // tagged solution words of the single-launch getrs
    DevBuf lu_inv, lu_mv;   // lookahead: inverse permutation, block moves
    DevBuf lu_stage, lu_sdst, lu_sperm;   // fused panel step
    unsigned trs_epoch = 0;
    int mix_side_groups = 0;   // mixed factor's second-stream cap (0: none)
    MixBufs mix;
    AuxBufs aux;
    int last_error = 0;
    flx::Profiler *prof = nullptr;   // non-null while profiling
};

// binds the calling handle's profiler for the duration of one entry point
struct ProfBind {
    ProfBind(flx_context *h) { flx::tl_prof = h ? h->prof : nullptr; if (h) h->wk.prof = h->prof; }
    ~ProfBind() { flx::tl_prof = nullptr; }
};

int hip_status(hipError_t e);

#define FLX_RET(x)                                                            \
    do {                                                                      \
        int r_ = (x);                                                         \
        if (r_) return r_;                                                    \
    } while (0)

#define FLX_TRY(x)                                                            \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) return hip_status(e_);                          \
    } while (0)

// grow-only device buffer; waits for queued work before releasing memory
hipError_t ensure(flx_context *h, DevBuf &b, size_t bytes);
// as ensure, and zero-filled whenever it is (re)allocated
hipError_t ensure_zeroed(flx_context *h, DevBuf &b, size_t bytes);
hipError_t init_work(flx_context *h);
void free_buf(DevBuf &b);
inline long round_ld(long n) { return ((n + 31) / 32) * 32; }

// FP64 LU / solve on device column-major data (api.hip)
// out (ldo): where the factors go (default: over A; else A is left as is).
// nrhs > 0 (m = n only): B (ldb) rides along as nrhs extra columns of the
// working copy, so the factorisation's interchanges and updates leave
// L^{-1} P B in them, copied to Y (n x nrhs, ld n); B itself is not written.
// With Y null instead, the caller solves from the row-major working copy
// (h->W, row stride *w_ld, L^{-1} P B at column n) and copies the factors
// back to A itself (nothing is copied out).
int dgetrf_dev(flx_context *h, int m, int n, double *A, int lda, int *ipiv, int *info,
               bool keep_perm, double *out = nullptr, int ldo = 0, int nrhs = 0,
               const double *B = nullptr, int ldb = 0, double *Y = nullptr,
               long *w_ld = nullptr);
// perm_fwd / perm_bwd: the row orders of ipiv applied forward / backward
// (ipiv_to_perm), when the caller already holds them; else derived from ipiv
int dgetrs_dev(flx_context *h, int trans, int n, int nrhs, const double *A, int lda,
               const int *ipiv, const int *perm_fwd, double *B, int ldb,
               const int *perm_bwd = nullptr);
