// api.hip -- the C ABI of libflx (include/flx.h): handles, device-pointer
// entry points and the FLENS plugin seams (CBLAS + Fortran LAPACK symbols).
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include "context.h"

using namespace flx;

int hip_status(hipError_t e)
{
    if (e == hipSuccess) return FLX_SUCCESS;
    if (e == hipErrorOutOfMemory) return FLX_ALLOC_FAILED;
    if (e == hipErrorInvalidValue) return FLX_UNSUPPORTED;
    return FLX_HIP_ERROR;
}

hipError_t ensure(flx_context *h, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return hipSuccess;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return e;
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t sz = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, sz);
    if (e != hipSuccess) { b.p = nullptr; return e; }
    b.cap = sz;
    return hipSuccess;
}

hipError_t ensure_zeroed(flx_context *h, DevBuf &b, size_t bytes)
{
    if (b.cap >= bytes) return hipSuccess;
    hipError_t e = ensure(h, b, bytes);
    if (e != hipSuccess) return e;
    return hipMemsetAsync(b.p, 0, b.cap, h->stream);
}

hipError_t init_work(flx_context *h)
{
    if (h->wk.exch) return hipSuccess;
    const size_t words = lu_exch_words();
    hipError_t e = hipMalloc(&h->wk.exch, words * sizeof(uint64_t));
    if (e != hipSuccess) return e;
    h->wk.exch_words = words;
    e = hipMemsetAsync(h->wk.exch, 0, words * sizeof(uint64_t), h->stream);
    if (e != hipSuccess) return e;
    e = hipMalloc(&h->wk.list, sizeof(int) * PANEL_LIST_SLICE * PANEL_GROUPS_MAX);
    if (e != hipSuccess) return e;
    e = hipMalloc(&h->wk.lsc, sizeof(double) * lu_lsc_doubles(LU_N_MAX));
    if (e != hipSuccess) return e;
    e = hipMalloc(&h->wk.flags, sizeof(int) * 4);
    if (e != hipSuccess) return e;
    e = hipMalloc(&h->wk.moved, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(h->wk.moved, 0, sizeof(unsigned long long), h->stream);
    if (e != hipSuccess) return e;
    // side-stream launches keep a quarter of the compute units free, an
    // eighth while the trailing matrix is large
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device);
    if (e != hipSuccess) return e;
    h->wk.side_groups = cus > 8 ? cus - cus / 4 : cus;
    h->wk.side_groups_early = cus > 8 ? cus - cus / 8 : cus;
    h->wk.cus = cus;
    return hipMemsetAsync(h->wk.flags, 0, sizeof(int) * 4, h->stream);
}

extern "C" int flx_create(flx_handle *out)
{
    if (!out) return FLX_INVALID_ARG;
    flx_context *h = new (std::nothrow) flx_context();
    if (!h) return FLX_ALLOC_FAILED;
    if (hipGetDevice(&h->device) != hipSuccess) { delete h; return FLX_HIP_ERROR; }
    *out = h;
    return FLX_SUCCESS;
}

void free_buf(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

extern "C" int flx_destroy(flx_handle h)
{
    if (!h) return FLX_INVALID_ARG;
    (void)hipStreamSynchronize(h->stream);
    free_buf(h->W); free_buf(h->A); free_buf(h->B); free_buf(h->C); free_buf(h->X);
    free_buf(h->piv0); free_buf(h->ipiv); free_buf(h->info); free_buf(h->perm);
    free_buf(h->tinv); free_buf(h->sb); free_buf(h->sx); free_buf(h->mlp_part);
    free_buf(h->trsw);
    MixBufs &x = h->mix;
    for (DevBuf *b : {&x.W32, &x.perm, &x.cand, &x.ccount, &x.counters, &x.u11, &x.win, &x.list,
                      &x.zero, &x.lpmax, &x.la, &x.ub, &x.dinv, &x.y32, &x.y64, &x.flags,
                      &x.err, &x.dr, &x.dc, &x.part, &x.scal, &x.R, &x.V, &x.w, &x.t, &x.zeros,
                      &x.mpart, &x.hbuf, &x.inv, &x.mv, &x.stage, &x.stage_perm, &x.F, &x.rbt})
        free_buf(*b);
    {
        AuxBufs &y = h->aux;
        for (DevBuf *b : {&y.slot, &y.v, &y.x, &y.r, &y.w, &y.lines, &y.isgn, &y.part, &y.ipiv1,
                          &y.perm0, &y.pf, &y.pb})
            free_buf(*b);
    }
    if (x.side) (void)hipStreamDestroy(x.side);
    if (x.ev_a) (void)hipEventDestroy(x.ev_a);
    if (x.ev_r) (void)hipEventDestroy(x.ev_r);
    if (x.graph) (void)hipGraphExecDestroy(x.graph);
    if (x.gstream) (void)hipStreamDestroy(x.gstream);
    if (x.ev0) (void)hipEventDestroy(x.ev0);
    if (x.ev1) (void)hipEventDestroy(x.ev1);
    if (h->wk.exch) (void)hipFree(h->wk.exch);
    if (h->wk.list) (void)hipFree(h->wk.list);
    if (h->wk.lsc) (void)hipFree(h->wk.lsc);
    if (h->wk.flags) (void)hipFree(h->wk.flags);
    if (h->wk.moved) (void)hipFree(h->wk.moved);
    if (h->wk.side) (void)hipStreamDestroy(h->wk.side);
    if (h->wk.ev_a) (void)hipEventDestroy(h->wk.ev_a);
    if (h->wk.ev_r) (void)hipEventDestroy(h->wk.ev_r);
    if (h->wk.ev_n) (void)hipEventDestroy(h->wk.ev_n);
    free_buf(h->lu_inv); free_buf(h->lu_mv);
    free_buf(h->lu_stage); free_buf(h->lu_sdst); free_buf(h->lu_sperm);
    free_buf(h->sr); free_buf(h->sc);
    prof_delete(h->prof);
    delete h;
    return FLX_SUCCESS;
}

extern "C" int flx_set_stream(flx_handle h, void *stream)
{
    if (!h) return FLX_INVALID_ARG;
    h->stream = (hipStream_t)stream;
    return FLX_SUCCESS;
}

extern "C" int flx_set_lu_blocking(flx_handle h, int nb, int ib)
{
    if (!h || nb < 1 || ib < 1 || ib > 64 || nb > 256) return FLX_INVALID_ARG;
    h->opt.nb = nb;
    h->opt.ib = ib;
    return FLX_SUCCESS;
}

extern "C" int flx_set_lu_lookahead(flx_handle h, int depth)
{
    if (!h || depth < 0 || depth > 1) return FLX_INVALID_ARG;
    h->opt.lookahead = depth;
    return FLX_SUCCESS;
}

extern "C" int flx_set_lu_side_groups(flx_handle h, int which, int groups)
{
    if (!h || groups < 0 || (which != 0 && which != 1)) return FLX_INVALID_ARG;
    FLX_TRY(init_work(h));
    if (which == 1) {
        h->mix_side_groups = groups;
        return FLX_SUCCESS;
    }
    int cus = 0;
    FLX_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    h->wk.side_groups = (groups == 0) ? (cus > 8 ? cus - cus / 4 : cus) : groups;
    h->wk.side_groups_early = (groups == 0) ? (cus > 8 ? cus - cus / 8 : cus) : groups;
    return FLX_SUCCESS;
}

extern "C" int flx_sync(flx_handle h)
{
    if (!h) return FLX_INVALID_ARG;
    FLX_TRY(hipStreamSynchronize(h->stream));
    if (h->wk.flags) {
        int f = 0;
        FLX_TRY(hipMemcpy(&f, h->wk.flags, sizeof(int), hipMemcpyDeviceToHost));
        if (f) {
            int z = 0;   // re-armed for the next call; the timeout is what is reported
            (void)hipMemcpy(h->wk.flags, &z, sizeof(int), hipMemcpyHostToDevice);
            return FLX_SYNC_TIMEOUT;
        }
    }
    return FLX_SUCCESS;
}

extern "C" int flx_test_raise_sync_flag(flx_handle h)
{
    if (!h) return FLX_INVALID_ARG;
    FLX_TRY(init_work(h));
    FLX_TRY(fill_int(h->stream, h->wk.flags, 1, 1));
    return FLX_SUCCESS;
}

extern "C" int flx_profile_start(flx_handle h)
{
    if (!h) return FLX_INVALID_ARG;
    FLX_TRY(init_work(h));
    if (!h->prof) h->prof = prof_new();
    FLX_TRY(hipStreamSynchronize(h->stream));
    prof_reset(h->prof);
    FLX_TRY(hipMemsetAsync(h->wk.moved, 0, sizeof(unsigned long long), h->stream));
    return FLX_SUCCESS;
}

extern "C" int flx_profile_classes(flx_handle h, unsigned mask)
{
    if (!h || !h->prof) return FLX_INVALID_ARG;
    prof_set_mask(h->prof, mask);
    return FLX_SUCCESS;
}

extern "C" int flx_profile_stop(flx_handle h, double *out)
{
    if (!h || !h->prof || !out) return FLX_INVALID_ARG;
    FLX_TRY(hipStreamSynchronize(h->stream));
    prof_collect(h->prof, out);
    unsigned long long moved = 0;
    FLX_TRY(hipMemcpy(&moved, h->wk.moved, sizeof(moved), hipMemcpyDeviceToHost));
    out[PROF_ROWPERM * 4 + 3] = 16.0 * (double)moved;   // each element read + written
    h->wk.prof = nullptr;
    Profiler *p = h->prof;
    h->prof = nullptr;
    prof_delete(p);
    return FLX_SUCCESS;
}

extern "C" const char *flx_status_string(int s)
{
    switch (s) {
    case FLX_SUCCESS: return "success";
    case FLX_INVALID_ARG: return "invalid argument";
    case FLX_HIP_ERROR: return "HIP runtime error";
    case FLX_ALLOC_FAILED: return "device allocation failed";
    case FLX_SYNC_TIMEOUT: return "device-side wait timed out";
    case FLX_UNSUPPORTED: return "unsupported shape or option";
    case FLX_SCRATCH_LIMIT: return "problem larger than the no-allocation scratch";
    default: return "unknown status";
    }
}

// ---------------------------------------------------------------------------
// device-pointer entries
// ---------------------------------------------------------------------------
extern "C" int flx_dgemm(flx_handle h, int transa, int transb, int m, int n, int k,
                         double alpha, const double *A, int lda, const double *B, int ldb,
                         double beta, double *C, int ldc)
{
    ProfBind pb_(h);
    if (!h || m < 0 || n < 0 || k < 0) return FLX_INVALID_ARG;
    const int arows = transa ? k : m, brows = transb ? n : k;
    if (lda < (arows > 1 ? arows : 1) || ldb < (brows > 1 ? brows : 1) ||
        ldc < (m > 1 ? m : 1))
        return FLX_INVALID_ARG;
    if (m == 0 || n == 0) return FLX_SUCCESS;
    ProfScope ps(h->prof, h->stream, PROF_GEMM, 2.0 * m * n * k,
                 8.0 * ((double)m * k + (double)k * n + 2.0 * m * n));
    FLX_TRY(gemm_f64(h->stream, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                     0));
    return FLX_SUCCESS;
}

// LU of device column-major A via the row-major working copy
int dgetrf_dev(flx_context *h, int m, int n, double *A, int lda, int *ipiv, int *info,
               bool keep_perm, double *out, int ldo, int nrhs, const double *B, int ldb,
               double *Y, long *w_ld)
{
    const int mn = m < n ? m : n;
    if (nrhs > 0 && (m != n || !B || (!Y && !w_ld))) return FLX_INVALID_ARG;
    const int nc = n + (nrhs > 0 ? nrhs : 0);   // columns of the working copy
    FLX_TRY(init_work(h));
    // with right-hand sides along, rows padded to a multiple of 256 doubles:
    // at n = 8192 + 16 a stride of 8448 measured 0.2 ms faster per solve than
    // 8224 (the trailing updates' tiles and the panels' rows spread over
    // the memory channels as with 8192)
    const long ldw = nrhs > 0 ? ((long)nc + 255) / 256 * 256 : round_ld(nc);
    FLX_TRY(ensure(h, h->W, sizeof(double) * (size_t)ldw * (size_t)(m > 0 ? m : 1)));
    FLX_TRY(ensure(h, h->piv0, sizeof(int) * (size_t)(mn > 0 ? mn : 1)));
    (void)keep_perm;   // the composed permutation is always kept (lookahead needs it)
    FLX_TRY(ensure(h, h->perm, sizeof(int) * (size_t)(m > 0 ? m : 1)));
    h->wk.perm = (int *)h->perm.p;
    FLX_TRY(ensure(h, h->lu_inv, sizeof(int) * (size_t)(m > 0 ? m : 1)));
    FLX_TRY(ensure(h, h->lu_mv, sizeof(int) * 2 * (1 + 2 * LU_BLOCK_MOVES)));   // two blocks' lists
    h->wk.inv = (int *)h->lu_inv.p;
    h->wk.mv = (int *)h->lu_mv.p;
    FLX_TRY(ensure(h, h->lu_stage, sizeof(double) * LU_STAGE_DOUBLES));
    FLX_TRY(ensure(h, h->lu_sdst, sizeof(int) * (LU_STAGE_TOP + 1)));
    FLX_TRY(ensure(h, h->lu_sperm, sizeof(int) * 2 * LU_STAGE_TOP));
    h->wk.stage = (double *)h->lu_stage.p;
    h->wk.sdst = (int *)h->lu_sdst.p;
    h->wk.sperm = (int *)h->lu_sperm.p;
    if (!h->wk.side) {
        FLX_TRY(hipStreamCreateWithFlags(&h->wk.side, hipStreamNonBlocking));
        FLX_TRY(hipEventCreateWithFlags(&h->wk.ev_a, hipEventDisableTiming));
        FLX_TRY(hipEventCreateWithFlags(&h->wk.ev_r, hipEventDisableTiming));
        FLX_TRY(hipEventCreateWithFlags(&h->wk.ev_n, hipEventDisableTiming));
    }
    double *W = (double *)h->W.p;
    FLX_TRY(fill_int(h->stream, info, 1, INT_MAX));
    // Into the working copy: the first block's columns on the stream, the
    // rest (and B) on the side stream beside the first block's panels; the
    // first block end waits for them (ev_r).
    const int bw = lu_block_width(h->opt);
    const int c1 = h->opt.lookahead && n > bw ? bw : n;
    FLX_TRY(transpose_f64(h->stream, A, lda, W, ldw, m, c1));
    if (c1 < n || nrhs > 0) {
        FLX_TRY(hipEventRecord(h->wk.ev_a, h->stream));
        FLX_TRY(hipStreamWaitEvent(h->wk.side, h->wk.ev_a, 0));
        if (c1 < n)
            FLX_TRY(transpose_f64(h->wk.side, A + (size_t)c1 * lda, lda, W + c1, ldw, m, n - c1));
        if (nrhs > 0) FLX_TRY(transpose_f64(h->wk.side, B, ldb, W + n, ldw, n, nrhs));
        FLX_TRY(hipEventRecord(h->wk.ev_r, h->wk.side));
        h->wk.side_pending = true;
    }
    FLX_TRY(getrf_rm_f64(h->stream, h->wk, W, ldw, m, nc, (int *)h->piv0.p, info, h->opt));
    if (nrhs > 0 && !Y) {
        *w_ld = ldw;   // the caller solves from W and copies the factors out
    } else {
        if (out) FLX_TRY(transpose_f64(h->stream, W, ldw, out, ldo, n, m));
        else FLX_TRY(transpose_f64(h->stream, W, ldw, A, lda, n, m));
        if (nrhs > 0) FLX_TRY(transpose_f64(h->stream, W + n, ldw, Y, n, nrhs, n));
    }
    FLX_TRY(piv0_to_ipiv(h->stream, (int *)h->piv0.p, ipiv, mn, info));
    return FLX_SUCCESS;
}

extern "C" int flx_dgetrf(flx_handle h, int m, int n, double *A, int lda, int *ipiv, int *info)
{
    ProfBind pb_(h);
    if (!h || m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLX_INVALID_ARG;
    if (m == 0 || n == 0) {
        FLX_TRY(fill_int(h->stream, info, 1, 0));
        return FLX_SUCCESS;
    }
    if (m > LU_N_MAX) return FLX_UNSUPPORTED;
    return dgetrf_dev(h, m, n, A, lda, ipiv, info, false);
}

// NoTrans: X = U^{-1} L^{-1} P B ; Trans: X = P^T L^{-T} U^{-T} B
int dgetrs_dev(flx_context *h, int trans, int n, int nrhs, const double *A, int lda,
                      const int *ipiv, const int *perm_fwd, double *B, int ldb,
                      const int *perm_bwd)
{
    const size_t bytes = sizeof(double) * (size_t)n * (size_t)nrhs;
    FLX_TRY(ensure(h, h->B, bytes));
    FLX_TRY(ensure(h, h->X, bytes));
    FLX_TRY(ensure(h, h->tinv, sizeof(double) * trs_inv_doubles(n)));
    double *T1 = (double *)h->B.p, *T2 = (double *)h->X.p;
    double *ti = (double *)h->tinv.p;
    if (!trans) {
        const int *perm = perm_fwd;
        if (!perm) {
            FLX_TRY(ensure(h, h->perm, sizeof(int) * (size_t)n));
            FLX_TRY(ipiv_to_perm(h->stream, ipiv, (int *)h->perm.p, n, 0));
            perm = (const int *)h->perm.p;
        }
        FLX_TRY(gather_rows_cm(h->stream, B, ldb, T1, n, perm, n, nrhs));
        FLX_TRY(init_work(h));
        FLX_TRY(ensure_zeroed(h, h->trsw, sizeof(uint64_t) * trs_chain_words(n)));
        FLX_TRY(trs_chain_f64(h->stream, n, nrhs, A, lda, T1, n, T2, B, ldb, ti,
                              (uint64_t *)h->trsw.p, &h->trs_epoch, h->wk.flags, h->wk.cus));
    } else {
        // Y = L^{-T} U^{-T} B into T1, one launch per triangle
        FLX_TRY(init_work(h));
        FLX_TRY(ensure_zeroed(h, h->trsw, sizeof(uint64_t) * trs_chain_words(n)));
        FLX_TRY(trs_persist_f64(h->stream, n, nrhs, A, lda, B, ldb, T2, T1, n,
                                (uint64_t *)h->trsw.p, &h->trs_epoch, h->wk.flags, 1));
        (void)ti;
        // apply the interchanges last to first: X = P^T Y
        const int *pb = perm_bwd;
        if (!pb) {
            FLX_TRY(ensure(h, h->perm, sizeof(int) * (size_t)n));
            FLX_TRY(ipiv_to_perm(h->stream, ipiv, (int *)h->perm.p, n, 1));
            pb = (const int *)h->perm.p;
        }
        FLX_TRY(gather_rows_cm(h->stream, T1, n, B, ldb, pb, n, nrhs));
    }
    return FLX_SUCCESS;
}

extern "C" int flx_dgetrs(flx_handle h, int trans, int n, int nrhs, const double *A, int lda,
                          const int *ipiv, double *B, int ldb)
{
    ProfBind pb_(h);
    if (!h || n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1))
        return FLX_INVALID_ARG;
    if (n == 0 || nrhs == 0) return FLX_SUCCESS;
    return dgetrs_dev(h, trans ? 1 : 0, n, nrhs, A, lda, ipiv, nullptr, B, ldb);
}

extern "C" int flx_dgesv(flx_handle h, int n, int nrhs, double *A, int lda, int *ipiv,
                         double *B, int ldb, int *info_host)
{
    ProfBind pb_(h);
    if (!h || n < 0 || nrhs < 0 || lda < (n > 1 ? n : 1) || ldb < (n > 1 ? n : 1))
        return FLX_INVALID_ARG;
    if (n == 0) { if (info_host) *info_host = 0; return FLX_SUCCESS; }
    if (n > LU_N_MAX) return FLX_UNSUPPORTED;
    FLX_TRY(ensure(h, h->info, sizeof(int)));
    int *dinfo = (int *)h->info.p;
    // B rides along as extra columns of the factorisation (the forward
    // solve L^{-1} P B is then done by its interchanges and block updates),
    // so only the U solve remains, read straight from the row-major working
    // copy while the factors go back to A on the side stream; B is written
    // only by that solve
    if (nrhs == 0) {
        int st = dgetrf_dev(h, n, n, A, lda, ipiv, dinfo, true);
        if (st) return st;
        int info = 0;
        FLX_TRY(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        FLX_RET(flx_sync(h));
        if (info_host) *info_host = info;
        return FLX_SUCCESS;
    }
    FLX_TRY(ensure(h, h->tinv, sizeof(double) * trs_inv_doubles(n)));
    FLX_TRY(ensure_zeroed(h, h->trsw, sizeof(uint64_t) * trs_chain_words(n)));
    long wld = 0;
    int st = dgetrf_dev(h, n, n, A, lda, ipiv, dinfo, true, nullptr, 0, nrhs, B, ldb, nullptr,
                        &wld);
    if (st) return st;
    int info = 0;
    FLX_TRY(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    // INFO is only meaningful when every panel wait succeeded
    st = flx_sync(h);
    const double *W = (const double *)h->W.p;
    if (st == FLX_SUCCESS && info_host) *info_host = info;
    if (st == FLX_SUCCESS && info == 0) {
        // the solve's diagonal inverses first; the factors go back to A on
        // the side stream beside the solve itself
        FLX_TRY(trs_chain_f64(h->stream, n, nrhs, W, wld, W + n, wld, nullptr, B, ldb,
                              (double *)h->tinv.p, (uint64_t *)h->trsw.p, &h->trs_epoch,
                              h->wk.flags, h->wk.cus, 0, 1, h->wk.ev_a));
    } else {
        FLX_TRY(hipEventRecord(h->wk.ev_a, h->stream));
    }
    FLX_TRY(hipStreamWaitEvent(h->wk.side, h->wk.ev_a, 0));
    FLX_TRY(transpose_f64(h->wk.side, W, wld, A, lda, n, n));
    FLX_TRY(hipEventRecord(h->wk.ev_r, h->wk.side));
    // A holds the factors once the side stream's copy is done
    FLX_TRY(hipStreamWaitEvent(h->stream, h->wk.ev_r, 0));
    return st;
}

extern "C" int flx_dlaswp(flx_handle h, int n, double *A, int lda, int k1, int k2,
                          const int *ipiv, int incx)
{
    if (!h || n < 0 || k1 < 1 || k2 < k1 - 1 || (incx != 1 && incx != -1))
        return FLX_INVALID_ARG;
    FLX_TRY(laswp_cm_f64(h->stream, A, lda, n, k1, k2, ipiv, incx));
    return FLX_SUCCESS;
}

extern "C" int flx_fill_uniform(flx_handle h, double *A, int m, int n, int lda, uint64_t seed,
                                uint64_t offset)
{
    if (!h || m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return FLX_INVALID_ARG;
    FLX_TRY(fill_uniform_f64(h->stream, A, m, n, lda, seed, offset));
    return FLX_SUCCESS;
}

// ---------------------------------------------------------------------------
// precision selector network
// ---------------------------------------------------------------------------
static_assert(FLX_MLP_PARAMS == MLP_PARAMS && FLX_MLP_IN == MLP_IN && FLX_MLP_OUT == MLP_OUT,
              "selector layout");

extern "C" int flx_mlp_forward(flx_handle h, const float *params, const float *X, int nb,
                               float *logits)
{
    if (!h || nb < 0 || (nb > 0 && (!params || !X || !logits))) return FLX_INVALID_ARG;
    FLX_TRY(mlp_forward(h->stream, params, X, nb, logits));
    return FLX_SUCCESS;
}

extern "C" int flx_mlp_grad(flx_handle h, const float *params, const float *X, const int *label,
                            const float *weight, int nb, float *grad, float *loss)
{
    if (!h || nb <= 0 || !params || !X || !label || !grad) return FLX_INVALID_ARG;
    FLX_TRY(ensure(h, h->mlp_part, sizeof(float) * mlp_part_floats(nb)));
    FLX_TRY(mlp_grad(h->stream, params, X, label, weight, nb, (float *)h->mlp_part.p, grad, loss));
    return FLX_SUCCESS;
}

extern "C" int flx_mlp_adam(flx_handle h, float *params, const float *grad, float *m, float *v,
                            int step, float lr, float weight_decay)
{
    if (!h || !params || !grad || !m || !v || step < 1) return FLX_INVALID_ARG;
    FLX_TRY(mlp_adam(h->stream, params, grad, m, v, step, lr, weight_decay));
    return FLX_SUCCESS;
}

// ---------------------------------------------------------------------------
// FLENS plugin seams (host pointers, synchronous)
// ---------------------------------------------------------------------------
extern "C" __attribute__((weak)) void xerbla_(const char *srname, const int *info, int len)
{
    fprintf(stderr, " ** On entry to %.*s parameter number %d had an illegal value\n",
            len, srname, *info);
}

static flx_context *default_handle()
{
    // one handle per host thread, created on first use
    static thread_local flx_context *h = nullptr;
    if (!h) {
        if (flx_create(&h) != FLX_SUCCESS) h = nullptr;
    }
    return h;
}

// The Fortran/CBLAS seams have no error channel for device failures (INFO
// > 0 means a singular factor); returning with INFO = 0 would hand the caller
// unfactored data.  A device failure therefore ends the program the way the
// reference's failed ASSERT does (flens/auxiliary/macros.h:63-66).
static void report(const char *who, int status)
{
    if (status != FLX_SUCCESS) {
        fprintf(stderr, "libflx: %s failed: %s\n", who, flx_status_string(status));
        abort();
    }
}

// host (ld) <-> device (compact, ld = rows) copies
static hipError_t h2d(void *dst, const double *src, int rows, int cols, int ld)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    return hipMemcpy2D(dst, sizeof(double) * rows, src, sizeof(double) * ld,
                       sizeof(double) * rows, cols, hipMemcpyHostToDevice);
}
static hipError_t d2h(double *dst, const void *src, int rows, int cols, int ld)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    return hipMemcpy2D(dst, sizeof(double) * ld, src, sizeof(double) * rows,
                       sizeof(double) * rows, cols, hipMemcpyDeviceToHost);
}

// blas_beta0 = 1: BLAS convention (beta = 0 never reads C); 0: the FLENS
// generic order, C scaled by beta first (0 * C propagates NaN / Inf)
static int host_gemm(int ta, int tb, int M, int N, int K, double alpha, const double *A, int lda,
                     const double *B, int ldb, double beta, double *C, int ldc,
                     int blas_beta0 = 1)
{
    flx_context *h = default_handle();
    if (!h) return FLX_HIP_ERROR;
    const int ar = ta ? K : M, ac = ta ? M : K;
    const int br = tb ? N : K, bc = tb ? K : N;
    FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)(ar > 0 ? ar : 1) * (ac > 0 ? ac : 1)));
    FLX_TRY(ensure(h, h->B, sizeof(double) * (size_t)(br > 0 ? br : 1) * (bc > 0 ? bc : 1)));
    FLX_TRY(ensure(h, h->C, sizeof(double) * (size_t)M * N));
    const bool need_ab = alpha != 0.0 && K > 0;
    if (need_ab) {
        FLX_TRY(h2d(h->A.p, A, ar, ac, lda));
        FLX_TRY(h2d(h->B.p, B, br, bc, ldb));
    }
    if (beta != 0.0 || !blas_beta0) FLX_TRY(h2d(h->C.p, C, M, N, ldc));
    FLX_TRY(gemm_f64(h->stream, ta, tb, M, N, need_ab ? K : 0, alpha, (const double *)h->A.p,
                     ar > 1 ? ar : 1, (const double *)h->B.p, br > 1 ? br : 1, beta,
                     (double *)h->C.p, M > 1 ? M : 1, blas_beta0));
    FLX_RET(flx_sync(h));
    FLX_TRY(d2h(C, h->C.p, M, N, ldc));
    return FLX_SUCCESS;
}

extern "C" void cblas_dgemm(int order, int transA, int transB, int M, int N, int K,
                            double alpha, const double *A, int lda, const double *B, int ldb,
                            double beta, double *C, int ldc)
{
    // CBLAS enums (cxxblas/drivers/cblas.h:22-27): 101 RowMajor, 102 ColMajor,
    // 111 NoTrans, 112 Trans, 113 ConjTrans (= Trans for real data).
    if (order != 101 && order != 102) { fprintf(stderr, "cblas_dgemm: parameter 1 illegal\n"); return; }
    int ta = (transA == 112 || transA == 113) ? 1 : (transA == 111 ? 0 : -1);
    int tb = (transB == 112 || transB == 113) ? 1 : (transB == 111 ? 0 : -1);
    if (ta < 0) { fprintf(stderr, "cblas_dgemm: parameter 2 illegal\n"); return; }
    if (tb < 0) { fprintf(stderr, "cblas_dgemm: parameter 3 illegal\n"); return; }
    if (order == 101) {
        // row-major C = op(A) op(B)  <=>  column-major C^T = op(B)^T op(A)^T
        std::swap(ta, tb);
        std::swap(M, N);
        std::swap(A, B);
        std::swap(lda, ldb);
    }
    const int arows = ta ? K : M, brows = tb ? N : K;
    int bad = 0;
    if (M < 0) bad = 4;
    else if (N < 0) bad = 5;
    else if (K < 0) bad = 6;
    else if (lda < (arows > 1 ? arows : 1)) bad = 9;
    else if (ldb < (brows > 1 ? brows : 1)) bad = 11;
    else if (ldc < (M > 1 ? M : 1)) bad = 14;
    if (bad) { fprintf(stderr, "cblas_dgemm: parameter %d illegal\n", bad); return; }
    if (M == 0 || N == 0) return;
    report("cblas_dgemm", host_gemm(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc));
}

static int host_trsm(int right, int upper, int trans, int unit, int M, int N, double alpha,
                     const double *A, int lda, double *B, int ldb)
{
    flx_context *h = default_handle();
    if (!h) return FLX_HIP_ERROR;
    const int na = right ? N : M;
    FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)na * na));
    FLX_TRY(ensure(h, h->C, sizeof(double) * (size_t)M * N));
    FLX_TRY(ensure(h, h->W, sizeof(double) * (size_t)M * N));
    FLX_TRY(ensure(h, h->tinv, sizeof(double) * trs_inv_doubles(na)));
    double *ti = (double *)h->tinv.p;
    FLX_TRY(h2d(h->A.p, A, na, na, lda));
    FLX_TRY(h2d(h->C.p, B, M, N, ldb));
    double *dB = (double *)h->C.p, *dT = (double *)h->W.p;
    if (!right) {
        // X = op(A)^{-1} B : solve consumes dT copy, writes dB
        FLX_TRY(copy_cm(h->stream, dB, M, dT, M, M, N));
        FLX_TRY(trsm_cm_f64(h->stream, upper, trans, unit, M, N, (const double *)h->A.p, na,
                            dT, M, dB, M, ti));
    } else {
        // X op(A) = B  <=>  op(A)^T X^T = B^T
        FLX_TRY(transpose_f64(h->stream, dB, M, dT, N, M, N));       // dT = B^T (N x M)
        FLX_TRY(trsm_cm_f64(h->stream, upper, !trans, unit, N, M, (const double *)h->A.p, na,
                            dT, N, dB, N, ti));                       // dB = X^T (N x M)
        FLX_TRY(transpose_f64(h->stream, dB, N, dT, M, N, M));       // dT = X (M x N)
        FLX_TRY(copy_cm(h->stream, dT, M, dB, M, M, N));
    }
    // alpha after the solve, as trsm_generic's closing gescal
    // (cxxblas/level3/trsm.tcc:67): B <- alpha * X, also for alpha = 0
    if (alpha != 1.0)
        FLX_TRY(gemm_f64(h->stream, 0, 0, M, N, 0, 0.0, nullptr, 1, nullptr, 1, alpha, dB, M, 0));
    FLX_RET(flx_sync(h));
    FLX_TRY(d2h(B, h->C.p, M, N, ldb));
    return FLX_SUCCESS;
}

extern "C" void cblas_dtrsm(int order, int side, int uplo, int transA, int diag, int M, int N,
                            double alpha, const double *A, int lda, double *B, int ldb)
{
    if (order != 101 && order != 102) { fprintf(stderr, "cblas_dtrsm: parameter 1 illegal\n"); return; }
    int right = side == 142 ? 1 : (side == 141 ? 0 : -1);
    int upper = uplo == 121 ? 1 : (uplo == 122 ? 0 : -1);
    int trans = (transA == 112 || transA == 113) ? 1 : (transA == 111 ? 0 : -1);
    int unit = diag == 132 ? 1 : (diag == 131 ? 0 : -1);
    if (right < 0 || upper < 0 || trans < 0 || unit < 0) {
        fprintf(stderr, "cblas_dtrsm: illegal enum parameter\n");
        return;
    }
    if (order == 101) {
        // row-major B (M x N) is column-major B^T (N x M): swap side and uplo
        right = !right;
        upper = !upper;
        std::swap(M, N);
    }
    const int na = right ? N : M;
    if (M < 0 || N < 0 || lda < (na > 1 ? na : 1) || ldb < (M > 1 ? M : 1)) {
        fprintf(stderr, "cblas_dtrsm: illegal size parameter\n");
        return;
    }
    if (M == 0 || N == 0) return;
    report("cblas_dtrsm", host_trsm(right, upper, trans, unit, M, N, alpha, A, lda, B, ldb));
}

extern "C" void dgetrf_(const int *M, const int *N, double *A, const int *LDA, int *IPIV,
                        int *INFO)
{
    // argument checks of flens/lapack/interface/src/getrf.cc:22-35
    *INFO = 0;
    if (*M < 0) *INFO = -1;
    else if (*N < 0) *INFO = -2;
    else if (*LDA < (*M > 1 ? *M : 1)) *INFO = -4;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGETRF", &i, 6);
        return;
    }
    const int m = *M, n = *N, mn = m < n ? m : n;
    if (m == 0 || n == 0) return;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && m > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)m * n));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)mn));
            FLX_TRY(ensure(h, h->info, sizeof(int)));
            FLX_TRY(h2d(h->A.p, A, m, n, *LDA));
            int s = dgetrf_dev(h, m, n, (double *)h->A.p, m, (int *)h->ipiv.p, (int *)h->info.p,
                               false);
            if (s) return s;
            s = flx_sync(h);
            if (s) return s;
            FLX_TRY(d2h(A, h->A.p, m, n, *LDA));
            FLX_TRY(hipMemcpy(IPIV, h->ipiv.p, sizeof(int) * mn, hipMemcpyDeviceToHost));
            FLX_TRY(hipMemcpy(INFO, h->info.p, sizeof(int), hipMemcpyDeviceToHost));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgetrf_", st);
}

extern "C" void dgetrs_(const char *TRANS, const int *N, const int *NRHS, const double *A,
                        const int *LDA, const int *IPIV, double *B, const int *LDB, int *INFO)
{
    const char t = TRANS ? (char)(TRANS[0] & ~0x20) : 'N';
    *INFO = 0;
    if (t != 'N' && t != 'T' && t != 'C') *INFO = -1;
    else if (*N < 0) *INFO = -2;
    else if (*NRHS < 0) *INFO = -3;
    else if (*LDA < (*N > 1 ? *N : 1)) *INFO = -5;
    else if (*LDB < (*N > 1 ? *N : 1)) *INFO = -8;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGETRS", &i, 6);
        return;
    }
    const int n = *N, nrhs = *NRHS;
    if (n == 0 || nrhs == 0) return;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)n * n));
            FLX_TRY(ensure(h, h->C, sizeof(double) * (size_t)n * nrhs));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)n));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(h2d(h->C.p, B, n, nrhs, *LDB));
            FLX_TRY(hipMemcpy(h->ipiv.p, IPIV, sizeof(int) * n, hipMemcpyHostToDevice));
            int s = dgetrs_dev(h, t != 'N', n, nrhs, (const double *)h->A.p, n,
                               (const int *)h->ipiv.p, nullptr, (double *)h->C.p, n);
            if (s) return s;
            FLX_RET(flx_sync(h));
            FLX_TRY(d2h(B, h->C.p, n, nrhs, *LDB));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgetrs_", st);
}

extern "C" void dgesv_(const int *N, const int *NRHS, double *A, const int *LDA, int *IPIV,
                       double *B, const int *LDB, int *INFO)
{
    *INFO = 0;
    if (*N < 0) *INFO = -1;
    else if (*NRHS < 0) *INFO = -2;
    else if (*LDA < (*N > 1 ? *N : 1)) *INFO = -4;
    else if (*LDB < (*N > 1 ? *N : 1)) *INFO = -7;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGESV ", &i, 6);
        return;
    }
    const int n = *N, nrhs = *NRHS;
    if (n == 0) return;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)n * n));
            FLX_TRY(ensure(h, h->C, sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1)));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)n));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(h2d(h->C.p, B, n, nrhs, *LDB));
            int info = 0;
            int s = flx_dgesv(h, n, nrhs, (double *)h->A.p, n, (int *)h->ipiv.p,
                              (double *)h->C.p, n, &info);
            if (s) return s;
            s = flx_sync(h);
            if (s) return s;
            *INFO = info;
            FLX_TRY(d2h(A, h->A.p, n, n, *LDA));
            FLX_TRY(hipMemcpy(IPIV, h->ipiv.p, sizeof(int) * n, hipMemcpyDeviceToHost));
            if (info == 0) FLX_TRY(d2h(B, h->C.p, n, nrhs, *LDB));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgesv_", st);
}

extern "C" void dsgesv_(const int *N, const int *NRHS, double *A, const int *LDA, int *IPIV,
                        double *B, const int *LDB, double *X, const int *LDX, double *WORK,
                        float *SWORK, int *ITER, int *INFO)
{
    (void)WORK;
    (void)SWORK;
    *INFO = 0;
    *ITER = 0;
    if (*N < 0) *INFO = -1;
    else if (*NRHS < 0) *INFO = -2;
    else if (*LDA < (*N > 1 ? *N : 1)) *INFO = -4;
    else if (*LDB < (*N > 1 ? *N : 1)) *INFO = -7;
    else if (*LDX < (*N > 1 ? *N : 1)) *INFO = -9;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DSGESV", &i, 6);
        return;
    }
    const int n = *N, nrhs = *NRHS;
    if (n == 0) return;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            const size_t rb = sizeof(double) * (size_t)n * (nrhs > 0 ? nrhs : 1);
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)n * n));
            FLX_TRY(ensure(h, h->sb, rb));
            FLX_TRY(ensure(h, h->sx, rb));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)n));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(h2d(h->sb.p, B, n, nrhs, *LDB));
            int iter = 0, info = 0;
            int s = flx_dsgesv(h, n, nrhs, (double *)h->A.p, n, (int *)h->ipiv.p,
                               (const double *)h->sb.p, n, (double *)h->sx.p, n, &iter, &info);
            if (s) return s;
            s = flx_sync(h);
            if (s) return s;
            *ITER = iter;
            *INFO = info;
            // A is overwritten only when the FP64 fallback factorised it
            if (iter < 0) FLX_TRY(d2h(A, h->A.p, n, n, *LDA));
            FLX_TRY(hipMemcpy(IPIV, h->ipiv.p, sizeof(int) * n, hipMemcpyDeviceToHost));
            if (info == 0) FLX_TRY(d2h(X, h->sx.p, n, nrhs, *LDX));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dsgesv_", st);
}

extern "C" void dlaswp_(const int *N, double *A, const int *LDA, const int *K1, const int *K2,
                        const int *IPIV, const int *INCX)
{
    const int n = *N, k1 = *K1, k2 = *K2, incx = *INCX;
    if (n <= 0 || incx == 0 || k2 < k1) return;
    // rows touched: 1..max(k2, max ipiv)
    int rows = k2;
    for (int i = k1 - 1; i < k2; ++i) rows = IPIV[i] > rows ? IPIV[i] : rows;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)rows * n));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)k2));
            FLX_TRY(h2d(h->A.p, A, rows, n, *LDA));
            FLX_TRY(hipMemcpy(h->ipiv.p, IPIV, sizeof(int) * k2, hipMemcpyHostToDevice));
            FLX_TRY(laswp_cm_f64(h->stream, (double *)h->A.p, rows, n, k1, k2,
                                 (const int *)h->ipiv.p, incx > 0 ? 1 : -1));
            FLX_RET(flx_sync(h));
            FLX_TRY(d2h(A, h->A.p, rows, n, *LDA));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dlaswp_", st);
}

// ---- norms, equilibration, condition estimate, refinement, expert driver --
static char up(const char *c) { return c ? (char)(c[0] & ~0x20) : ' '; }

extern "C" double dlange_(const char *NORM, const int *M, const int *N, const double *A,
                          const int *LDA, double *WORK)
{
    (void)WORK;
    const int m = *M, n = *N;
    if ((m < n ? m : n) <= 0) return 0.0;
    char nm = up(NORM);
    if (nm == 'E') nm = 'F';
    if (NORM && NORM[0] == '1') nm = 'O';
    if (nm != 'M' && nm != 'O' && nm != 'I' && nm != 'F') return 0.0;
    flx_context *h = default_handle();
    double v = 0.0;
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)m * n));
            FLX_TRY(h2d(h->A.p, A, m, n, *LDA));
            return flx_dlange(h, nm, m, n, (const double *)h->A.p, m, &v);
        };
        st = run();
    }
    report("dlange_", st);
    return v;
}

extern "C" void dgeequ_(const int *M, const int *N, const double *A, const int *LDA, double *R,
                        double *C, double *ROWCND, double *COLCND, double *AMAX, int *INFO)
{
    *INFO = 0;
    if (*M < 0) *INFO = -1;
    else if (*N < 0) *INFO = -2;
    else if (*LDA < (*M > 1 ? *M : 1)) *INFO = -4;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGEEQU", &i, 6);
        return;
    }
    const int m = *M, n = *N;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)(m > 0 ? m : 1) * (n > 0 ? n : 1)));
            FLX_TRY(ensure(h, h->sr, sizeof(double) * (size_t)(m > 0 ? m : 1)));
            FLX_TRY(ensure(h, h->sc, sizeof(double) * (size_t)(n > 0 ? n : 1)));
            FLX_TRY(h2d(h->A.p, A, m, n, *LDA));
            int info = 0;
            int s = flx_dgeequ(h, m, n, (const double *)h->A.p, m > 1 ? m : 1, (double *)h->sr.p,
                               (double *)h->sc.p, ROWCND, COLCND, AMAX, &info);
            if (s) return s;
            FLX_RET(flx_sync(h));
            *INFO = info;
            if (m > 0 && n > 0) {
                // R holds the row maxima even when a zero row stops the routine
                FLX_TRY(hipMemcpy(R, h->sr.p, sizeof(double) * m, hipMemcpyDeviceToHost));
                if (info == 0 || info > m)
                    FLX_TRY(hipMemcpy(C, h->sc.p, sizeof(double) * n, hipMemcpyDeviceToHost));
            }
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgeequ_", st);
}

extern "C" void dgecon_(const char *NORM, const int *N, const double *A, const int *LDA,
                        const double *ANORM, double *RCOND, double *WORK, int *IWORK, int *INFO)
{
    (void)WORK;
    (void)IWORK;
    const char nm = (NORM && NORM[0] == '1') ? 'O' : up(NORM);
    *INFO = 0;
    if (nm != 'O' && nm != 'I') *INFO = -1;
    else if (*N < 0) *INFO = -2;
    else if (*LDA < (*N > 1 ? *N : 1)) *INFO = -4;
    else if (*ANORM < 0.0) *INFO = -5;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGECON", &i, 6);
        return;
    }
    const int n = *N;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)(n > 0 ? n : 1) * (n > 0 ? n : 1)));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            return flx_dgecon(h, nm, n, (const double *)h->A.p, n > 1 ? n : 1, *ANORM, RCOND);
        };
        st = run();
    }
    report("dgecon_", st);
}

extern "C" void dgerfs_(const char *TRANS, const int *N, const int *NRHS, const double *A,
                        const int *LDA, const double *AF, const int *LDAF, const int *IPIV,
                        const double *B, const int *LDB, double *X, const int *LDX, double *FERR,
                        double *BERR, double *WORK, int *IWORK, int *INFO)
{
    (void)WORK;
    (void)IWORK;
    const char t = up(TRANS);
    const int l = *N > 1 ? *N : 1;
    *INFO = 0;
    if (t != 'N' && t != 'T' && t != 'C') *INFO = -1;
    else if (*N < 0) *INFO = -2;
    else if (*NRHS < 0) *INFO = -3;
    else if (*LDA < l) *INFO = -5;
    else if (*LDAF < l) *INFO = -7;
    else if (*LDB < l) *INFO = -10;
    else if (*LDX < l) *INFO = -12;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGERFS", &i, 6);
        return;
    }
    const int n = *N, nrhs = *NRHS;
    if (n == 0 || nrhs == 0) {
        for (int j = 0; j < nrhs; ++j) { FERR[j] = 0.0; BERR[j] = 0.0; }
        return;
    }
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            const size_t nn = sizeof(double) * (size_t)n * n, nb = sizeof(double) * (size_t)n * nrhs;
            FLX_TRY(ensure(h, h->A, nn));
            FLX_TRY(ensure(h, h->sb, nn));
            FLX_TRY(ensure(h, h->C, nb));
            FLX_TRY(ensure(h, h->sx, nb));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)n));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(h2d(h->sb.p, AF, n, n, *LDAF));
            FLX_TRY(h2d(h->C.p, B, n, nrhs, *LDB));
            FLX_TRY(h2d(h->sx.p, X, n, nrhs, *LDX));
            FLX_TRY(hipMemcpy(h->ipiv.p, IPIV, sizeof(int) * n, hipMemcpyHostToDevice));
            int s = flx_dgerfs(h, t != 'N', n, nrhs, (const double *)h->A.p, n,
                               (const double *)h->sb.p, n, (const int *)h->ipiv.p,
                               (const double *)h->C.p, n, (double *)h->sx.p, n, FERR, BERR);
            if (s) return s;
            FLX_RET(flx_sync(h));
            FLX_TRY(d2h(X, h->sx.p, n, nrhs, *LDX));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgerfs_", st);
}

extern "C" void dgesvx_(const char *FACT, const char *TRANS, const int *N, const int *NRHS,
                        double *A, const int *LDA, double *AF, const int *LDAF, int *IPIV,
                        char *EQUED, double *R, double *C, double *B, const int *LDB, double *X,
                        const int *LDX, double *RCOND, double *FERR, double *BERR, double *WORK,
                        int *IWORK, int *INFO)
{
    (void)IWORK;
    const char f = up(FACT), t = up(TRANS);
    const int n = *N, nrhs = *NRHS, l = n > 1 ? n : 1;
    char eq = (f == 'F') ? up(EQUED) : 'N';
    const bool rowequ = f == 'F' && (eq == 'R' || eq == 'B');
    const bool colequ = f == 'F' && (eq == 'C' || eq == 'B');
    *INFO = 0;
    if (f != 'N' && f != 'E' && f != 'F') *INFO = -1;
    else if (t != 'N' && t != 'T' && t != 'C') *INFO = -2;
    else if (n < 0) *INFO = -3;
    else if (nrhs < 0) *INFO = -4;
    else if (*LDA < l) *INFO = -6;
    else if (*LDAF < l) *INFO = -8;
    else if (f == 'F' && eq != 'N' && eq != 'R' && eq != 'C' && eq != 'B') *INFO = -10;
    else if (*LDB < l) *INFO = -14;
    else if (*LDX < l) *INFO = -16;
    if (*INFO == 0 && rowequ)
        for (int j = 0; j < n; ++j)
            if (!(R[j] > 0.0)) { *INFO = -11; break; }
    if (*INFO == 0 && colequ)
        for (int j = 0; j < n; ++j)
            if (!(C[j] > 0.0)) { *INFO = -12; break; }
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGESVX", &i, 6);
        return;
    }
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            const size_t nn = sizeof(double) * (size_t)l * l;
            const size_t nb = sizeof(double) * (size_t)l * (nrhs > 0 ? nrhs : 1);
            FLX_TRY(ensure(h, h->A, nn));
            FLX_TRY(ensure(h, h->sb, nn));
            FLX_TRY(ensure(h, h->C, nb));
            FLX_TRY(ensure(h, h->sx, nb));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)l));
            FLX_TRY(ensure(h, h->sr, sizeof(double) * (size_t)l));
            FLX_TRY(ensure(h, h->sc, sizeof(double) * (size_t)l));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(h2d(h->C.p, B, n, nrhs, *LDB));
            if (f == 'F') {
                FLX_TRY(h2d(h->sb.p, AF, n, n, *LDAF));
                if (n > 0)
                    FLX_TRY(hipMemcpy(h->ipiv.p, IPIV, sizeof(int) * n, hipMemcpyHostToDevice));
                if (rowequ) FLX_TRY(h2d(h->sr.p, R, n, 1, l));
                if (colequ) FLX_TRY(h2d(h->sc.p, C, n, 1, l));
            }
            double rcond = 0.0, rpvgrw = 0.0;
            int info = 0;
            int s = flx_dgesvx(h, f, t != 'N', n, nrhs, (double *)h->A.p, l, (double *)h->sb.p,
                               l, (int *)h->ipiv.p, &eq, (double *)h->sr.p, (double *)h->sc.p,
                               (double *)h->C.p, l, (double *)h->sx.p, l, &rcond, FERR, BERR,
                               &rpvgrw, &info);
            if (s) return s;
            FLX_RET(flx_sync(h));
            *RCOND = rcond;
            WORK[0] = rpvgrw;
            *INFO = info;
            *EQUED = eq;
            if (f == 'E' && eq != 'N') FLX_TRY(d2h(A, h->A.p, n, n, *LDA));
            if (f != 'F') {
                FLX_TRY(d2h(AF, h->sb.p, n, n, *LDAF));
                if (n > 0)
                    FLX_TRY(hipMemcpy(IPIV, h->ipiv.p, sizeof(int) * n, hipMemcpyDeviceToHost));
            }
            if (f == 'E' && n > 0) {
                FLX_TRY(d2h(R, h->sr.p, n, 1, l));
                FLX_TRY(d2h(C, h->sc.p, n, 1, l));
            }
            const bool rowq = eq == 'R' || eq == 'B', colq = eq == 'C' || eq == 'B';
            if ((t == 'N') ? rowq : colq) FLX_TRY(d2h(B, h->C.p, n, nrhs, *LDB));
            if (info == 0 || info == n + 1) FLX_TRY(d2h(X, h->sx.p, n, nrhs, *LDX));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgesvx_", st);
}

extern "C" void dgetri_(const int *N, double *A, const int *LDA, const int *IPIV, double *WORK,
                        const int *LWORK, int *INFO)
{
    const int n = *N;
    // workspace query: the reference reports n * nb with nb = 64 (ilaenv GETRI)
    const int lwopt = (n > 1 ? n : 1) * 64;
    *INFO = 0;
    if (n < 0) *INFO = -1;
    else if (*LDA < (n > 1 ? n : 1)) *INFO = -3;
    else if (*LWORK < (n > 1 ? n : 1) && *LWORK != -1) *INFO = -6;
    if (*INFO != 0) {
        int i = -*INFO;
        xerbla_("DGETRI", &i, 6);
        return;
    }
    if (WORK) WORK[0] = (double)lwopt;
    if (*LWORK == -1 || n == 0) return;
    flx_context *h = default_handle();
    int st = h ? FLX_SUCCESS : FLX_HIP_ERROR;
    if (h && n > LU_N_MAX) st = FLX_UNSUPPORTED;
    if (st == FLX_SUCCESS) {
        auto run = [&]() -> int {
            FLX_TRY(ensure(h, h->A, sizeof(double) * (size_t)n * n));
            FLX_TRY(ensure(h, h->ipiv, sizeof(int) * (size_t)n));
            FLX_TRY(h2d(h->A.p, A, n, n, *LDA));
            FLX_TRY(hipMemcpy(h->ipiv.p, IPIV, sizeof(int) * n, hipMemcpyHostToDevice));
            int info = 0;
            int s = flx_dgetri(h, n, (double *)h->A.p, n, (const int *)h->ipiv.p, &info);
            if (s) return s;
            *INFO = info;
            if (info == 0) FLX_TRY(d2h(A, h->A.p, n, n, *LDA));
            return FLX_SUCCESS;
        };
        st = run();
    }
    report("dgetri_", st);
}

// Fortran BLAS seam, the convention of FLENS's own export
// (flens/blas/interface/blas/gemm.cc:121-229): argument checks with info
// 1, 2, 3, 4, 5, 8, 10, 13 through xerbla_, then blas::mm's generic order
// (C scaled by beta first, so beta = 0 computes 0 * C).
extern "C" void dgemm_(const char *TRANSA, const char *TRANSB, const int *M, const int *N,
                       const int *K, const double *ALPHA, const double *A, const int *LDA,
                       const double *B, const int *LDB, const double *BETA, double *C,
                       const int *LDC)
{
    const char ta = up(TRANSA), tb = up(TRANSB);
    const int m = *M, n = *N, k = *K;
    const int nrowa = (ta == 'N') ? m : k, nrowb = (tb == 'N') ? k : n;
    int info = 0;
    if (ta != 'N' && ta != 'C' && ta != 'T') info = 1;
    else if (tb != 'N' && tb != 'C' && tb != 'T') info = 2;
    else if (m < 0) info = 3;
    else if (n < 0) info = 4;
    else if (k < 0) info = 5;
    else if (*LDA < (nrowa > 1 ? nrowa : 1)) info = 8;
    else if (*LDB < (nrowb > 1 ? nrowb : 1)) info = 10;
    else if (*LDC < (m > 1 ? m : 1)) info = 13;
    if (info) {
        xerbla_("DGEMM ", &info, 6);
        return;
    }
    if (m == 0 || n == 0) return;
    report("dgemm_", host_gemm(ta != 'N', tb != 'N', m, n, k, *ALPHA, A, *LDA, B, *LDB, *BETA, C,
                               *LDC, 0));
}
