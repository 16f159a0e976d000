// solve_mixed.hip -- mixed-precision solver: low-precision LU + FP64
// refinement, with FP64 LU as the fallback.
//
//   flx_dsgesv / dsgesv_  netlib dsgesv semantics (dsgesv.f:131-337, the
//                         routine FLENS reaches as cxxlapack::sgesv,
//                         cxxlapack/netlib/interface/lapack.in.h:10611-10623):
//                         FP32 factor, classic refinement, ITERMAX = 30,
//                         stop when ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n)
//                         for every column, ITER < 0 codes for the fallback.
//   flx_dgesv_mixed       the same with a choice of factor precision (FP32,
//                         BF16 or FP16 trailing updates) and of refinement
//                         (classic or GMRES-IR), stopping on the normwise
//                         backward error ||r|| / (||A|| ||x|| + ||b||) <= tol.
//
// FP64 appears only where the refinement needs it: residuals r = b - A x on
// the caller's A, solution updates and norms.  The factor is applied as the
// preconditioner M^{-1} (ir.hip).
#include <chrono>
#include <cstring>
#include <cmath>
#include <vector>
#include "context.h"

using namespace flx;

namespace {

constexpr double EPS64 = 1.1102230246251565e-16;   // 2^-53, lamch(Eps)
constexpr int SLOT_ANRM = 0, SLOT_VN = 1, SLOT_T = 2, SLOT_COL = 5;   // scalar slots

struct Opts {
    int prec = MIX_FP32, refine = FLX_REFINE_CLASSIC, itermax = 30, gmres_m = 40;
    int pivot = FLX_PIVOT_TOURNAMENT;
    double tol = 0.0, gmres_tol = 1e-10;
    int fallback = 1, nb = 256, graph = 1;
    bool dsgesv = false;
};

double bits_to_double(unsigned long long b)
{
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

struct Solver {
    flx_context *h;
    hipStream_t s;
    int n, nrhs;
    const double *A;
    int lda;
    const double *B;
    int ldb;
    double *X;
    int ldx;
    Opts o;
    PrecFactor f;
    MixWork mw;
    long ldw = 0;
    unsigned long long *scal = nullptr;
    std::vector<unsigned long long> hs;   // host copy of the scalar slots

    unsigned long long *slot(int k) { return scal + k; }
    unsigned long long *col_slot(int c, int q) { return scal + SLOT_COL + 3 * c + q; }

    hipError_t read_slots()
    {
        hipError_t e = hipMemcpyAsync(hs.data(), scal, sizeof(unsigned long long) * hs.size(),
                                      hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(s);
    }
    double hslot(int k) const { return bits_to_double(hs[k]); }
    double hcol(int c, int q) const { return bits_to_double(hs[SLOT_COL + 3 * c + q]); }

    int alloc()
    {
        MixBufs &m = h->mix;
        FLX_TRY(init_work(h));
        ldw = round_ld(n);
        const size_t nodes = tslu_nodes(n);
        FLX_TRY(ensure(h, m.W32, sizeof(float) * (size_t)ldw * n));
        FLX_TRY(ensure(h, m.perm, sizeof(int) * n));
        FLX_TRY(ensure(h, m.cand, sizeof(int) * nodes * 32));
        FLX_TRY(ensure(h, m.ccount, sizeof(int) * nodes));
        FLX_TRY(ensure_zeroed(h, m.counters, sizeof(unsigned) * nodes));
        FLX_TRY(ensure(h, m.u11, sizeof(float) * MIX_U11_FLOATS));
        FLX_TRY(ensure(h, m.win, sizeof(int) * 32));
        FLX_TRY(ensure(h, m.list, sizeof(int) * PANEL_LIST_SLICE));
        FLX_TRY(ensure(h, m.zero, sizeof(int) * 4));
        FLX_TRY(ensure(h, m.lpmax, sizeof(unsigned) * 4));
        FLX_TRY(ensure(h, m.inv, sizeof(int) * n));
        FLX_TRY(ensure(h, m.mv, sizeof(int) * (1 + 2 * LU_BLOCK_MOVES)));
        FLX_TRY(ensure(h, m.stage, sizeof(float) * MIX_STAGE_ROWS * MIX_STAGE_LD));
        FLX_TRY(ensure(h, m.stage_perm, sizeof(int) * MIX_STAGE_ROWS));
        if (!m.side) {
            FLX_TRY(hipStreamCreateWithFlags(&m.side, hipStreamNonBlocking));
            FLX_TRY(hipEventCreateWithFlags(&m.ev_a, hipEventDisableTiming));
            FLX_TRY(hipEventCreateWithFlags(&m.ev_r, hipEventDisableTiming));
        }
        if (o.prec != MIX_FP32) {
            FLX_TRY(ensure(h, m.la, 2 * lp_pack_elems(n, o.nb)));
            FLX_TRY(ensure(h, m.ub, 2 * lp_pack_elems(n, o.nb)));
        }
        FLX_TRY(ensure(h, m.dinv, sizeof(double) * trsv_inv_doubles(n)));
        FLX_TRY(ensure(h, m.y32, sizeof(float) * n));
        FLX_TRY(ensure(h, m.y64, sizeof(double) * n));
        if (m.flags.cap < sizeof(uint64_t) * trsv_words(n)) m.epoch = 0;
        FLX_TRY(ensure_zeroed(h, m.flags, sizeof(uint64_t) * trsv_words(n)));
        FLX_TRY(ensure(h, m.err, sizeof(int) * 4));
        FLX_TRY(ensure(h, m.dr, sizeof(double) * n));
        FLX_TRY(ensure(h, m.dc, sizeof(double) * n));
        const bool rbt = o.pivot == FLX_PIVOT_RBT && rbt_supported(n);
        if (rbt && m.rbt_n != n) {
            // the butterflies of order n, fixed seeds (the same for every call)
            FLX_TRY(ensure(h, m.rbt, sizeof(double) * 4 * (size_t)n));
            std::vector<double> r(4 * (size_t)n);
            for (int b = 0; b < 4; ++b) rbt_values(n, 0x5eed0000ull + (unsigned long long)b,
                                                   r.data() + (size_t)b * n);
            FLX_TRY(hipMemcpyAsync(m.rbt.p, r.data(), sizeof(double) * r.size(),
                                   hipMemcpyHostToDevice, s));
            FLX_TRY(hipStreamSynchronize(s));
            m.rbt_n = n;
        }
        FLX_TRY(ensure(h, m.part, sizeof(double) * ir_part_doubles(n)));
        const int nslots = SLOT_COL + 3 * nrhs;
        FLX_TRY(ensure(h, m.scal, sizeof(unsigned long long) * nslots));
        FLX_TRY(ensure(h, m.R, sizeof(double) * (size_t)n * nrhs));
        FLX_TRY(ensure(h, m.t, sizeof(double) * n));
        if (o.refine != FLX_REFINE_CLASSIC) {
            FLX_TRY(ensure(h, m.V, sizeof(double) * (size_t)n * (o.gmres_m + 1)));
            FLX_TRY(ensure(h, m.w, sizeof(double) * n));
            FLX_TRY(ensure(h, m.t, sizeof(double) * n));
            FLX_TRY(ensure(h, m.zeros, sizeof(double) * n));
            FLX_TRY(ensure(h, m.mpart, sizeof(double) * mdot_part_doubles(n, o.gmres_m + 1)));
            FLX_TRY(ensure(h, m.hbuf, sizeof(double) * 4 * (o.gmres_m + 2)));
        }
        scal = (unsigned long long *)m.scal.p;
        hs.assign(nslots, 0ull);
        mw.perm = (int *)m.perm.p;
        mw.cand = (int *)m.cand.p;
        mw.ccount = (int *)m.ccount.p;
        mw.counters = (unsigned *)m.counters.p;
        mw.u11 = (float *)m.u11.p;
        mw.win = (int *)m.win.p;
        mw.list = (int *)m.list.p;
        mw.zero = (int *)m.zero.p;
        mw.lpmax = (unsigned *)m.lpmax.p;
        mw.la = m.la.p;
        mw.ub = m.ub.p;
        mw.prof = h->prof;
        const bool look = h->opt.lookahead != 0;
        mw.inv = look ? (int *)m.inv.p : nullptr;
        mw.mv = (int *)m.mv.p;
        mw.side = m.side;
        mw.ev_a = m.ev_a;
        mw.ev_r = m.ev_r;
        mw.side_groups = h->mix_side_groups;
        mw.stage = look ? (float *)m.stage.p : nullptr;
        mw.stage_perm = look ? (int *)m.stage_perm.p : nullptr;
        f.n = n;
        f.W = (const float *)m.W32.p;
        f.ldw = ldw;
        f.perm = (const int *)m.perm.p;
        f.dr = (o.prec == MIX_FP16) ? (const double *)m.dr.p : nullptr;
        f.dc = (o.prec == MIX_FP16) ? (const double *)m.dc.p : nullptr;
        f.rbt = rbt;
        if (rbt) {
            const double *r = (const double *)m.rbt.p;
            f.rb.u1 = r;
            f.rb.u2 = r + n;
            f.rb.v1 = r + 2 * (size_t)n;
            f.rb.v2 = r + 3 * (size_t)n;
        }
        f.y32 = (float *)m.y32.p;
        f.y64 = (double *)m.y64.p;
        f.dinv = (const double *)m.dinv.p;
        f.xt = (uint64_t *)m.flags.p;
        f.epoch = &m.epoch;
        f.err = (int *)m.err.p;
        return FLX_SUCCESS;
    }

    double *R(int c) { return (double *)h->mix.R.p + (size_t)c * n; }
    const double *Bc(int c) { return B + (size_t)c * ldb; }
    double *Xc(int c) { return X + (size_t)c * ldx; }

    // the factorisation's launches: tournament-pivoted, or without pivoting
    // for the butterfly-transformed matrix
    hipError_t run_factor(hipStream_t st)
    {
        MixBufs &m = h->mix;
        if (f.rbt) return getrf_rm_genp(st, mw, (float *)m.W32.p, ldw, n, o.prec, o.nb);
        return getrf_rm_mixed(st, mw, (float *)m.W32.p, ldw, n, o.prec, o.nb);
    }

    // The factorisation's launches depend only on (n, prec, nb) and the
    // workspace, so they are captured once into a graph and replayed, on a
    // private stream ordered after / before the caller's by events (the
    // caller's stream may be the legacy default stream, which cannot be
    // captured).  Recaptured when the key changes.
    int factor_graph()
    {
        MixBufs &m = h->mix;
        // every value a captured launch depends on: shape, precision, blocking,
        // the side stream's workgroup cap and every workspace address
        const size_t key[MIX_GKEY] = {
            (size_t)n, (size_t)ldw, (size_t)o.prec, (size_t)o.nb, (size_t)mw.side_groups,
            (size_t)f.rbt,
            (size_t)m.W32.p, (size_t)mw.perm, (size_t)mw.cand, (size_t)mw.ccount,
            (size_t)mw.counters, (size_t)mw.u11, (size_t)mw.win, (size_t)mw.list,
            (size_t)mw.zero, (size_t)mw.lpmax, (size_t)mw.la, (size_t)mw.ub, (size_t)mw.inv,
            (size_t)mw.mv, (size_t)mw.stage, (size_t)mw.stage_perm, (size_t)mw.side};
        if (!m.gstream) {
            FLX_TRY(hipStreamCreateWithFlags(&m.gstream, hipStreamNonBlocking));
            FLX_TRY(hipEventCreateWithFlags(&m.ev0, hipEventDisableTiming));
            FLX_TRY(hipEventCreateWithFlags(&m.ev1, hipEventDisableTiming));
        }
        if (!m.graph || memcmp(key, m.gkey, sizeof key) != 0) {
            if (m.graph) {
                FLX_TRY(hipStreamSynchronize(m.gstream));
                (void)hipGraphExecDestroy(m.graph);
                m.graph = nullptr;
            }
            hipGraph_t g = nullptr;
            FLX_TRY(hipStreamBeginCapture(m.gstream, hipStreamCaptureModeThreadLocal));
            hipError_t e = run_factor(m.gstream);
            hipGraph_t gg = nullptr;
            hipError_t e2 = hipStreamEndCapture(m.gstream, &gg);
            if (e != hipSuccess) { if (gg) (void)hipGraphDestroy(gg); return hip_status(e); }
            if (e2 != hipSuccess) return hip_status(e2);
            g = gg;
            e = hipGraphInstantiate(&m.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { m.graph = nullptr; return hip_status(e); }
            memcpy(m.gkey, key, sizeof key);
        }
        FLX_TRY(hipEventRecord(m.ev0, s));
        FLX_TRY(hipStreamWaitEvent(m.gstream, m.ev0, 0));
        FLX_TRY(hipGraphLaunch(m.graph, m.gstream));
        FLX_TRY(hipEventRecord(m.ev1, m.gstream));
        FLX_TRY(hipStreamWaitEvent(s, m.ev1, 0));
        return FLX_SUCCESS;
    }

    // factor: 0 ok, -2 FP32 overflow in A or B, -3 exact zero pivot
    int factor(int *code)
    {
        MixBufs &m = h->mix;
        FLX_TRY(zero_slots(s, scal, SLOT_COL + 3 * nrhs));
        FLX_TRY(hipMemsetAsync(m.err.p, 0, sizeof(int) * 4, s));
        FLX_TRY(fill_int(s, mw.zero, 1, INT_MAX));
        FLX_TRY(norm_inf(s, A, lda, n, (double *)m.part.p, slot(SLOT_ANRM)));
        for (int c = 0; c < nrhs; ++c) FLX_TRY(vmax(s, Bc(c), n, col_slot(c, 2)));
        if (o.prec == MIX_FP16)
            FLX_TRY(equilibrate(s, A, lda, n, (double *)m.part.p, (double *)m.dr.p,
                                (double *)m.dc.p));
        if (f.rbt)
            FLX_TRY(rbt_conv(s, A, lda, n, f.dr, f.dc, f.rb, (float *)m.W32.p, ldw,
                             (int *)m.err.p + 1));
        else
            FLX_TRY(conv_d2s_rm(s, A, lda, n, f.dr, f.dc, (float *)m.W32.p, ldw,
                                (int *)m.err.p + 1));
        if (o.graph && !h->prof) {
            FLX_RET(factor_graph());
        } else {
            FLX_TRY(run_factor(s));
        }
        FLX_TRY(trsv_diag_inv(s, f.W, ldw, n, (double *)m.dinv.p));
        int flags[4] = {0, 0, 0, 0}, zero = INT_MAX;
        FLX_TRY(hipMemcpyAsync(flags, m.err.p, sizeof flags, hipMemcpyDeviceToHost, s));
        FLX_TRY(hipMemcpyAsync(&zero, mw.zero, sizeof(int), hipMemcpyDeviceToHost, s));
        FLX_TRY(read_slots());
        *code = 0;
        bool bovf = false;
        for (int c = 0; c < nrhs; ++c) bovf |= hcol(c, 2) > 3.4028234663852886e38;
        if (flags[1] || bovf) *code = -2;
        else if (zero != INT_MAX) *code = -3;
        return FLX_SUCCESS;
    }

    // residuals and norms of every column; returns 1 when all converged
    int check(bool *conv, double *eta_max)
    {
        FLX_TRY(zero_slots(s, scal + SLOT_COL, 3 * nrhs));
        for (int c = 0; c < nrhs; ++c) {
            FLX_TRY(gemv_resid(s, A, lda, n, Xc(c), Bc(c), R(c), (double *)h->mix.part.p,
                               col_slot(c, 0)));
            FLX_TRY(vmax(s, Xc(c), n, col_slot(c, 1)));
            FLX_TRY(vmax(s, Bc(c), n, col_slot(c, 2)));
        }
        FLX_TRY(read_slots());
        const double anrm = hslot(SLOT_ANRM);
        const double cte = anrm * EPS64 * std::sqrt((double)n);
        bool all = true;
        double em = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            const double rn = hcol(c, 0), xn = hcol(c, 1), bn = hcol(c, 2);
            const double denom = anrm * xn + bn;
            // a NaN anywhere (residual, iterate, right-hand side) is never
            // convergence: NaN / NaN would otherwise read as 0 below
            const bool nan = std::isnan(rn) || std::isnan(xn) || std::isnan(denom);
            const double eta = nan ? NAN
                             : denom > 0.0 ? rn / denom : (rn > 0.0 ? INFINITY : 0.0);
            em = std::isnan(eta) || std::isnan(em) ? NAN : std::max(em, eta);
            const bool ok = o.dsgesv ? !(rn > xn * cte) && !nan : eta <= o.tol;
            all = all && ok;
        }
        *conv = all;
        *eta_max = em;
        return FLX_SUCCESS;
    }

    int precond(const double *v, double *z, int add)
    {
        FLX_TRY(zero_slots(s, slot(SLOT_VN), 1));
        FLX_TRY(prec_apply(s, f, v, z, add, slot(SLOT_VN), nullptr));
        return FLX_SUCCESS;
    }

    // kappa_inf lower-bound estimate from the factor: ||A||_inf times
    // max(||A^{-1} e||_inf, ||A^{-1} sign(A^{-1} e)||_inf), e = ones.  The
    // factor is applied in FP32 (an order-of-magnitude estimate; the factor's
    // own error dominates the arithmetic's), about half the time of the FP64
    // application.
    int estimate(double *kappa)
    {
        MixBufs &m = h->mix;
        double *v = R(0);   // R holds nothing yet
        double *w = (double *)m.t.p;
        FLX_TRY(fill_const(s, v, n, 1.0));
        FLX_TRY(zero_slots(s, slot(SLOT_VN), 1));
        FLX_TRY(prec_apply(s, f, v, w, 0, slot(SLOT_VN), nullptr));
        FLX_TRY(sign_of(s, w, v, n));
        FLX_TRY(zero_slots(s, slot(SLOT_T), 2));
        FLX_TRY(vmax(s, w, n, slot(SLOT_T)));
        FLX_TRY(zero_slots(s, slot(SLOT_VN), 1));
        FLX_TRY(prec_apply(s, f, v, w, 0, slot(SLOT_VN), nullptr));
        FLX_TRY(vmax(s, w, n, slot(SLOT_T + 1)));
        FLX_TRY(read_slots());
        const double ainv = std::max(hslot(SLOT_T), hslot(SLOT_T + 1));
        *kappa = hslot(SLOT_ANRM) * ainv;
        return FLX_SUCCESS;
    }

    // classic refinement: returns iterations (>= 0) or -itermax-1
    int classic(int *iters, double *eta)
    {
        bool conv = false;
        for (int c = 0; c < nrhs; ++c) {
            int st = precond(Bc(c), Xc(c), 0);
            if (st) return st;
        }
        int st = check(&conv, eta);
        if (st) return st;
        if (conv) { *iters = 0; return FLX_SUCCESS; }
        for (int it = 1; it <= o.itermax; ++it) {
            for (int c = 0; c < nrhs; ++c) {
                st = precond(R(c), Xc(c), 1);
                if (st) return st;
            }
            st = check(&conv, eta);
            if (st) return st;
            if (conv) { *iters = it; return FLX_SUCCESS; }
        }
        *iters = -o.itermax - 1;
        return FLX_SUCCESS;
    }

    // left-preconditioned GMRES on M^{-1} A d = M^{-1} r for column c; x_c += d
    int gmres(int c, int *inner)
    {
        MixBufs &m = h->mix;
        const int mm = o.gmres_m;
        double *V = (double *)m.V.p, *w = (double *)m.w.p, *t = (double *)m.t.p;
        double *hb = (double *)m.hbuf.p;   // [0..mm+1) h1, [mm+1..) h2, then norms
        double *mp = (double *)m.mpart.p;
        std::vector<double> H((size_t)(mm + 1) * mm, 0.0), g(mm + 1, 0.0), cs(mm), sn(mm);
        std::vector<double> hh(2 * (mm + 1) + 2);
        FLX_TRY(prec_apply64(s, f, R(c), V, 0));
        FLX_TRY(mdot(s, V, n, n, 1, V, mp, hb + 2 * (mm + 1)));
        FLX_TRY(hipMemcpyAsync(hh.data(), hb + 2 * (mm + 1), sizeof(double), hipMemcpyDeviceToHost, s));
        FLX_TRY(hipStreamSynchronize(s));
        const double beta = std::sqrt(hh[0]);
        if (!(beta > 0.0) || !std::isfinite(beta)) { *inner = 0; return FLX_SUCCESS; }
        FLX_TRY(vscale_copy(s, V, 1.0 / beta, V, n));
        g[0] = beta;
        int k = 0;
        for (; k < mm; ++k) {
            double *vk = V + (size_t)k * n;
            FLX_TRY(gemv_resid(s, A, lda, n, vk, (const double *)m.zeros.p, t,
                               (double *)m.part.p, slot(SLOT_T)));       // t = -A v_k
            FLX_TRY(prec_apply64(s, f, t, w, 0));                         // w = -M^{-1} A v_k
            FLX_TRY(vscale_copy(s, w, -1.0, w, n));
            // classical Gram-Schmidt, twice
            FLX_TRY(mdot(s, V, n, n, k + 1, w, mp, hb));
            FLX_TRY(maxpy(s, V, n, n, k + 1, hb, w));
            FLX_TRY(mdot(s, V, n, n, k + 1, w, mp, hb + (mm + 1)));
            FLX_TRY(maxpy(s, V, n, n, k + 1, hb + (mm + 1), w));
            FLX_TRY(mdot(s, w, n, n, 1, w, mp, hb + 2 * (mm + 1)));
            FLX_TRY(hipMemcpyAsync(hh.data(), hb, sizeof(double) * (2 * (mm + 1) + 1),
                                   hipMemcpyDeviceToHost, s));
            FLX_TRY(hipStreamSynchronize(s));
            double *hk = &H[(size_t)k * (mm + 1)];
            for (int i = 0; i <= k; ++i) hk[i] = hh[i] + hh[mm + 1 + i];
            const double hn = std::sqrt(hh[2 * (mm + 1)]);
            hk[k + 1] = hn;
            if (hn > 0.0 && std::isfinite(hn))
                FLX_TRY(vscale_copy(s, w, 1.0 / hn, V + (size_t)(k + 1) * n, n));
            // Givens rotations
            for (int i = 0; i < k; ++i) {
                const double a = hk[i], b = hk[i + 1];
                hk[i] = cs[i] * a + sn[i] * b;
                hk[i + 1] = -sn[i] * a + cs[i] * b;
            }
            const double a = hk[k], b = hk[k + 1];
            const double r = std::hypot(a, b);
            cs[k] = r > 0.0 ? a / r : 1.0;
            sn[k] = r > 0.0 ? b / r : 0.0;
            hk[k] = r;
            hk[k + 1] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            if (!std::isfinite(g[k + 1])) return FLX_SUCCESS;   // leaves x unchanged
            if (std::fabs(g[k + 1]) <= o.gmres_tol * beta || !(hn > 0.0)) { ++k; break; }
        }
        // y = H^{-1} g, x_c += V y
        std::vector<double> y(k, 0.0), ny(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
            double v = g[i];
            for (int j = i + 1; j < k; ++j) v -= H[(size_t)j * (mm + 1) + i] * y[j];
            const double d = H[(size_t)i * (mm + 1) + i];
            y[i] = d != 0.0 ? v / d : 0.0;
        }
        for (int i = 0; i < k; ++i) ny[i] = -y[i];   // maxpy subtracts
        FLX_TRY(hipMemcpyAsync(hb, ny.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        FLX_TRY(hipStreamSynchronize(s));   // ny is a host temporary
        FLX_TRY(maxpy(s, V, n, n, k, hb, Xc(c)));
        *inner = k;
        return FLX_SUCCESS;
    }

    int gmres_ir(int *iters, int *inner_total, double *eta)
    {
        MixBufs &m = h->mix;
        FLX_TRY(hipMemsetAsync(m.zeros.p, 0, sizeof(double) * n, s));
        for (int c = 0; c < nrhs; ++c) FLX_TRY(prec_apply64(s, f, Bc(c), Xc(c), 0));
        bool conv = false;
        *inner_total = 0;
        for (int it = 0; it <= o.itermax; ++it) {
            int st = check(&conv, eta);
            if (st) return st;
            if (conv) { *iters = it; return FLX_SUCCESS; }
            if (it == o.itermax || std::isnan(*eta)) break;
            for (int c = 0; c < nrhs; ++c) {
                int inner = 0;
                st = gmres(c, &inner);
                if (st) return st;
                *inner_total += inner;
            }
        }
        *iters = -o.itermax - 1;
        return FLX_SUCCESS;
    }

    // FP64 LU + solve: the factorisation runs on a copy F of A, so that the
    // backward error of the fallback's X is measured against A itself; the
    // factors then replace A, as dsgesv's fallback leaves them.  *info =
    // LAPACK info of the FP64 factorisation, *eta its solution's backward
    // error (NaN when info > 0: no solution).
    int fallback64(int *info, int *ipiv, double *eta)
    {
        MixBufs &m = h->mix;
        FLX_TRY(copy_cm(s, B, ldb, X, ldx, n, nrhs));
        FLX_TRY(ensure(h, h->info, sizeof(int)));
        FLX_TRY(ensure(h, h->ipiv, sizeof(int) * n));
        FLX_TRY(ensure(h, m.F, sizeof(double) * (size_t)n * n));
        double *F = (double *)m.F.p;
        int *piv = ipiv ? ipiv : (int *)h->ipiv.p;
        FLX_RET(dgetrf_dev(h, n, n, const_cast<double *>(A), lda, piv, (int *)h->info.p, true,
                           F, n));
        FLX_TRY(hipMemcpyAsync(info, h->info.p, sizeof(int), hipMemcpyDeviceToHost, s));
        FLX_RET(flx_sync(h));   // a panel wait that gave up is this call's error
        *eta = NAN;
        if (*info == 0) {
            FLX_RET(dgetrs_dev(h, 0, n, nrhs, F, n, piv, (const int *)h->perm.p, X, ldx));
            FLX_RET(flx_sync(h));
            bool conv = false;
            const bool keep = o.dsgesv;
            o.dsgesv = false;
            const int r = check(&conv, eta);
            o.dsgesv = keep;
            if (r) return r;
        }
        FLX_TRY(copy_cm(s, F, n, const_cast<double *>(A), lda, n, n));
        return FLX_SUCCESS;
    }

    // LAPACK-style interchanges from the factor's row order (host, O(n))
    int perm_to_ipiv(int *ipiv_dev)
    {
        std::vector<int> perm(n), ip(n), cur(n), where(n);
        FLX_TRY(hipMemcpyAsync(perm.data(), mw.perm, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        FLX_TRY(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) { cur[i] = i; where[i] = i; }
        for (int i = 0; i < n; ++i) {
            const int p = where[perm[i]];
            ip[i] = p + 1;
            const int ri = cur[i], rp = cur[p];
            cur[i] = rp; cur[p] = ri;
            where[rp] = i; where[ri] = p;
        }
        FLX_TRY(hipMemcpyAsync(ipiv_dev, ip.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        FLX_TRY(hipStreamSynchronize(s));
        return FLX_SUCCESS;
    }
};

bool valid_dims(int n, int nrhs, int lda, int ldb, int ldx)
{
    const int m1 = n > 1 ? n : 1;
    return n >= 0 && nrhs >= 0 && lda >= m1 && ldb >= m1 && ldx >= m1;
}

} // namespace

static int run_mixed(flx_context *h, int n, int nrhs, double *A, int lda, const double *B,
                     int ldb, double *X, int ldx, const Opts &o, flx_mixed_stats *st,
                     int *ipiv, int *iter_out, int *info_out)
{
    using clk = std::chrono::steady_clock;
    Solver S{h, h->stream, n, nrhs, A, lda, B, ldb, X, ldx, o, {}, {}};
    if (o.tol <= 0.0) S.o.tol = std::sqrt((double)n) * EPS64;
    int r = S.alloc();
    if (r) return r;
    const auto t0 = clk::now();
    int code = 0;
    r = S.factor(&code);
    if (r) return r;
    const auto t1 = clk::now();
    int iters = 0, inner = 0;
    double eta = NAN, kappa = 0.0;
    int refine = o.refine;
    if (code == 0 && refine == FLX_REFINE_AUTO) {
        r = S.estimate(&kappa);
        if (r) return r;
        const double u = (o.prec == MIX_FP32) ? 5.960464477539063e-08
                         : (o.prec == MIX_FP16) ? 4.8828125e-04 : 3.90625e-03;
        // kappa_inf overestimates kappa_2 (by up to n); the thresholds are
        // set where the refinement, with the factor already paid for, is
        // still expected to cost less than an FP64 LU
        const double ku = kappa * u;
        if (!(ku < 100.0)) code = -4;                     // FP64 at once
        else if (ku < 2.0) { refine = FLX_REFINE_CLASSIC; S.o.itermax = std::max(S.o.itermax, 60); }
        else refine = FLX_REFINE_GMRES;
    }
    if (code == 0) {
        r = (refine == FLX_REFINE_GMRES) ? S.gmres_ir(&iters, &inner, &eta)
                                         : S.classic(&iters, &eta);
        if (r) return r;
        int terr[4] = {0, 0, 0, 0};
        FLX_TRY(hipMemcpy(terr, h->mix.err.p, sizeof terr, hipMemcpyDeviceToHost));
        if (terr[0]) return FLX_SYNC_TIMEOUT;
        code = iters;
    }
    const auto t2 = clk::now();
    int info = 0, fell = 0;
    if (code < 0 && (o.dsgesv || o.fallback || code == -4)) {   // -4: the ladder chose FP64
        r = S.fallback64(&info, ipiv, &eta);
        if (r) return r;
        fell = 1;
    } else if (code >= 0 && ipiv) {
        r = S.perm_to_ipiv(ipiv);
        if (r) return r;
    }
    const auto t3 = clk::now();
    if (iter_out) *iter_out = code;
    if (info_out) *info_out = info;
    if (st) {
        st->iter = code;
        st->inner_iters = inner;
        st->fallback = fell;
        st->info = info;
        st->backward_error = eta;
        st->anorm = S.hslot(SLOT_ANRM);
        st->kappa_est = kappa;
        st->factor_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->refine_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        st->fallback_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    }
    return FLX_SUCCESS;
}

extern "C" int flx_dgesv_mixed(flx_handle h, int n, int nrhs, double *A, int lda,
                               const double *B, int ldb, double *X, int ldx,
                               const flx_mixed_opts *opt, flx_mixed_stats *st)
{
    ProfBind pb_(h);
    if (!h || !valid_dims(n, nrhs, lda, ldb, ldx)) return FLX_INVALID_ARG;
    Opts o;
    if (opt) {
        if (opt->prec < 0 || opt->prec > 2 || opt->refine < 0 || opt->refine > 2 ||
            opt->itermax < 0 || opt->gmres_m < 0 || opt->gmres_m > 400)
            return FLX_INVALID_ARG;
        o.prec = opt->prec;
        o.refine = opt->refine;
        o.itermax = opt->itermax;
        o.gmres_m = opt->gmres_m > 0 ? opt->gmres_m : 40;
        o.tol = opt->tol;
        o.gmres_tol = opt->gmres_tol > 0.0 ? opt->gmres_tol : 1e-10;
        o.fallback = opt->fallback;
        o.nb = opt->nb > 0 ? opt->nb : 256;
        o.graph = opt->graph;
        o.pivot = opt->pivot;
    }
    if (st) memset(st, 0, sizeof *st);
    if (n == 0 || nrhs == 0) return FLX_SUCCESS;
    if (n > 65536) return FLX_UNSUPPORTED;
    if (o.fallback && n > LU_N_MAX) return FLX_UNSUPPORTED;
    return run_mixed(h, n, nrhs, A, lda, B, ldb, X, ldx, o, st, nullptr, nullptr, nullptr);
}

extern "C" int flx_dsgesv(flx_handle h, int n, int nrhs, double *A, int lda, int *ipiv,
                          const double *B, int ldb, double *X, int ldx, int *iter, int *info)
{
    ProfBind pb_(h);
    if (!h || !valid_dims(n, nrhs, lda, ldb, ldx)) return FLX_INVALID_ARG;
    if (iter) *iter = 0;
    if (info) *info = 0;
    if (n == 0) return FLX_SUCCESS;
    if (n > LU_N_MAX) return FLX_UNSUPPORTED;
    Opts o;
    o.dsgesv = true;
    o.itermax = 30;
    o.nb = h->opt.nb;
    return run_mixed(h, n, nrhs, A, lda, B, ldb, X, ldx, o, nullptr, ipiv, iter, info);
}
