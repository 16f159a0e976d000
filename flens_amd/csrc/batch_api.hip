// batch_api.hip -- plans for the batched solves and GEMMs (include/flx.h).
//
// A plan fixes one batch: its problem table (ordered by decreasing cost for
// the persistent workgroups' queue), the launch shape chosen by the
// autotuner, and a HIP graph holding the whole batch as one persistent
// launch, replayed by flx_batch_execute on a private stream ordered after
// and before the handle's stream by events.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>
#include "context.h"

using namespace flx;

namespace {

struct BatchProbH {   // layout of flx::BatchProb (batched.hip)
    long a, b;
    int n, nrhs, piv, id;
};
struct GemmProbH {    // layout of flx::GemmProb (batched.hip)
    long a, b, c;
    int m, n, k, tile0;
    double alpha, beta;
};

} // namespace

struct flx_batch {
    flx_context *h = nullptr;
    int kind = 0;               // 0 solves, 1 GEMMs
    int mode = 0, itmax = 5;
    int count = 0;
    void *probs = nullptr;      // device problem table
    int *counter = nullptr;
    // solves
    double *A = nullptr, *B = nullptr, *R = nullptr;
    float *W32 = nullptr;
    int *ipiv = nullptr, *info = nullptr, *iters = nullptr;
    long a_total = 0, b_total = 0;
    // GEMMs
    const double *GA = nullptr, *GB = nullptr;
    double *GC = nullptr;
    long c_total = 0;
    int ntiles[3] = {0, 0, 0};
    void *gprobs[3] = {nullptr, nullptr, nullptr};   // tables per tile shape
    // launch shape
    int tile_cfg = 1, grid = 0;
    double tuned_ms = 0.0;
    BatchOut out;
    // graph
    hipStream_t gs = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int cu_count(flx_context *h)
{
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus;
}

static hipError_t launch(flx_batch *p, hipStream_t s, int tile_cfg, int grid, double *A,
                         double *B, double *C)
{
    if (p->kind == 0)
        return batched_gesv(s, p->mode, p->probs, p->count, p->counter, A, B, p->W32, p->R,
                            p->ipiv, p->info, p->iters, p->itmax, grid, p->out);
    return batched_gemm(s, tile_cfg, p->gprobs[tile_cfg], p->count, p->ntiles[tile_cfg],
                        p->counter, p->GA, p->GB, C, grid);
}

static int capture(flx_batch *p)
{
    if (!p->gs) {
        FLX_TRY(hipStreamCreateWithFlags(&p->gs, hipStreamNonBlocking));
        FLX_TRY(hipEventCreateWithFlags(&p->ev0, hipEventDisableTiming));
        FLX_TRY(hipEventCreateWithFlags(&p->ev1, hipEventDisableTiming));
    }
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
    }
    hipGraph_t g = nullptr;
    FLX_TRY(hipStreamBeginCapture(p->gs, hipStreamCaptureModeThreadLocal));
    hipError_t e = launch(p, p->gs, p->tile_cfg, p->grid, p->A, p->B, p->GC);
    hipError_t e2 = hipStreamEndCapture(p->gs, &g);
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_status(e); }
    if (e2 != hipSuccess) return hip_status(e2);
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { p->exec = nullptr; return hip_status(e); }
    return FLX_SUCCESS;
}

// Times each candidate launch shape on copies of the plan's inputs (the
// solves work in place) and keeps the fastest; cfgs = tile shapes (GEMMs).
static int autotune(flx_batch *p, bool tune)
{
    const int cus = cu_count(p->h);
    std::vector<int> grids = {cus, 2 * cus, 4 * cus};
    std::vector<int> cfgs = {0};
    if (p->kind == 1) cfgs = {0, 1, 2};
    if (!tune) {
        p->grid = 2 * cus;
        p->tile_cfg = (p->kind == 1) ? 1 : 0;
        return FLX_SUCCESS;
    }
    hipStream_t s = p->h->stream;
    double *tA = nullptr, *tB = nullptr, *tC = nullptr;
    int *saved_piv = nullptr;
    if (p->kind == 0) {
        FLX_TRY(hipMalloc(&tA, sizeof(double) * std::max(1L, p->a_total)));
        FLX_TRY(hipMalloc(&tB, sizeof(double) * std::max(1L, p->b_total)));
    } else {
        FLX_TRY(hipMalloc(&tC, sizeof(double) * std::max(1L, p->c_total)));
    }
    (void)saved_piv;
    hipEvent_t e0, e1;
    FLX_TRY(hipEventCreate(&e0));
    FLX_TRY(hipEventCreate(&e1));
    double best = 1e30;
    int bcfg = cfgs[0], bgrid = grids[0];
    for (int cfg : cfgs)
        for (int grid : grids) {
            float ms_best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                if (p->kind == 0) {
                    FLX_TRY(hipMemcpyAsync(tA, p->A, sizeof(double) * p->a_total,
                                           hipMemcpyDeviceToDevice, s));
                    FLX_TRY(hipMemcpyAsync(tB, p->B, sizeof(double) * p->b_total,
                                           hipMemcpyDeviceToDevice, s));
                } else {
                    FLX_TRY(hipMemcpyAsync(tC, p->GC, sizeof(double) * p->c_total,
                                           hipMemcpyDeviceToDevice, s));
                }
                FLX_TRY(hipEventRecord(e0, s));
                FLX_TRY(launch(p, s, cfg, grid, tA, tB, tC));
                FLX_TRY(hipEventRecord(e1, s));
                FLX_TRY(hipEventSynchronize(e1));
                float ms = 0.f;
                FLX_TRY(hipEventElapsedTime(&ms, e0, e1));
                ms_best = std::min(ms_best, ms);
            }
            if (ms_best < best) { best = ms_best; bcfg = cfg; bgrid = grid; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (tA) (void)hipFree(tA);
    if (tB) (void)hipFree(tB);
    if (tC) (void)hipFree(tC);
    p->tile_cfg = bcfg;
    p->grid = bgrid;
    p->tuned_ms = best;
    return FLX_SUCCESS;
}

extern "C" int flx_batch_gesv_create(flx_handle h, flx_batch_t *out, int count, const int *n,
                                     const int *nrhs, double *A, double *B, int *ipiv, int *info,
                                     int *iters, int mode, int itmax, int tune,
                                     const flx_batch_extra *extra)
{
    if (!h || !out || count < 0 || (count > 0 && (!n || !nrhs || !A || !B || !ipiv || !info)) ||
        mode < 0 || mode > 2 || itmax < 0 || itmax > 100)
        return FLX_INVALID_ARG;
    if (mode == 2 && (!extra || !extra->selector)) return FLX_INVALID_ARG;
    if (batched_prob_bytes() != sizeof(BatchProbH)) return FLX_HIP_ERROR;
    std::vector<BatchProbH> tab(count);
    long ao = 0, bo = 0;
    int po = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1 || n[i] > 512 || nrhs[i] < 0 || nrhs[i] > 8) return FLX_UNSUPPORTED;
        tab[i] = BatchProbH{ao, bo, n[i], nrhs[i], po, i};
        ao += (long)n[i] * n[i];
        bo += (long)n[i] * nrhs[i];
        po += n[i];
    }
    // largest first: the queue then balances the persistent workgroups
    std::stable_sort(tab.begin(), tab.end(), [](const BatchProbH &x, const BatchProbH &y) {
        return (double)x.n * x.n * x.n + 2.0 * x.n * x.n * x.nrhs >
               (double)y.n * y.n * y.n + 2.0 * y.n * y.n * y.nrhs;
    });
    flx_batch *p = new (std::nothrow) flx_batch();
    if (!p) return FLX_ALLOC_FAILED;
    p->h = h;
    p->kind = 0;
    p->mode = mode;
    p->itmax = itmax;
    p->count = count;
    p->A = A; p->B = B; p->ipiv = ipiv; p->info = info; p->iters = iters;
    p->a_total = ao;
    p->b_total = bo;
    if (extra) {
        p->out.sel = extra->selector;
        p->out.feat = extra->features;
        p->out.tsys = extra->ticks;
        p->out.choice = extra->choice;
    }
    auto fail = [&](int st) { flx_batch_destroy(p); return st; };
    if (hipMalloc(&p->probs, sizeof(BatchProbH) * std::max(1, count)) != hipSuccess ||
        hipMalloc(&p->counter, sizeof(int)) != hipSuccess)
        return fail(FLX_ALLOC_FAILED);
    if (mode >= 1 && (hipMalloc(&p->W32, sizeof(float) * std::max(1L, ao)) != hipSuccess ||
                      hipMalloc(&p->R, sizeof(double) * std::max(1L, bo)) != hipSuccess))
        return fail(FLX_ALLOC_FAILED);
    if (count > 0 &&
        hipMemcpy(p->probs, tab.data(), sizeof(BatchProbH) * count, hipMemcpyHostToDevice) !=
            hipSuccess)
        return fail(FLX_HIP_ERROR);
    int st = autotune(p, tune != 0);
    if (st) return fail(st);
    st = capture(p);
    if (st) return fail(st);
    *out = p;
    return FLX_SUCCESS;
}

extern "C" int flx_batch_gemm_create(flx_handle h, flx_batch_t *out, int count, const int *m,
                                     const int *n, const int *k, const double *alpha,
                                     const double *beta, const double *A, const double *B,
                                     double *C, int tune)
{
    if (!h || !out || count < 0 ||
        (count > 0 && (!m || !n || !k || !alpha || !beta || !A || !B || !C)))
        return FLX_INVALID_ARG;
    for (int i = 0; i < count; ++i)
        if (m[i] < 0 || n[i] < 0 || k[i] < 0) return FLX_INVALID_ARG;
    if (batched_gemm_prob_bytes() != sizeof(GemmProbH)) return FLX_HIP_ERROR;
    flx_batch *p = new (std::nothrow) flx_batch();
    if (!p) return FLX_ALLOC_FAILED;
    p->h = h;
    p->kind = 1;
    p->count = count;
    p->GA = A; p->GB = B; p->GC = C;
    auto fail = [&](int st) { flx_batch_destroy(p); return st; };
    // offsets in the caller's order, problems ordered by decreasing cost
    std::vector<long> ao(count), bo(count), co(count);
    long a = 0, b = 0, c = 0;
    for (int i = 0; i < count; ++i) {
        ao[i] = a; bo[i] = b; co[i] = c;
        a += (long)m[i] * k[i];
        b += (long)k[i] * n[i];
        c += (long)m[i] * n[i];
    }
    p->c_total = c;
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (double)m[x] * n[x] * k[x] > (double)m[y] * n[y] * k[y];
    });
    if (hipMalloc(&p->counter, sizeof(int)) != hipSuccess) return fail(FLX_ALLOC_FAILED);
    for (int cfg = 0; cfg < 3; ++cfg) {
        int tm = 0, tn = 0;
        batched_gemm_tile(cfg, &tm, &tn);
        std::vector<GemmProbH> tab;
        tab.reserve(count);
        int tiles = 0;
        for (int i : order) {
            if (m[i] == 0 || n[i] == 0) continue;
            GemmProbH g{ao[i], bo[i], co[i], m[i], n[i], k[i], tiles, alpha[i], beta[i]};
            tab.push_back(g);
            tiles += ((m[i] + tm - 1) / tm) * ((n[i] + tn - 1) / tn);
        }
        p->ntiles[cfg] = tiles;
        if (hipMalloc(&p->gprobs[cfg], sizeof(GemmProbH) * std::max<size_t>(1, tab.size())) !=
            hipSuccess)
            return fail(FLX_ALLOC_FAILED);
        if (!tab.empty() && hipMemcpy(p->gprobs[cfg], tab.data(), sizeof(GemmProbH) * tab.size(),
                                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(FLX_HIP_ERROR);
        if (cfg == 0) p->count = (int)tab.size();
    }
    int st = autotune(p, tune != 0);
    if (st) return fail(st);
    st = capture(p);
    if (st) return fail(st);
    *out = p;
    return FLX_SUCCESS;
}

extern "C" int flx_batch_execute(flx_batch_t p)
{
    if (!p || !p->exec) return FLX_INVALID_ARG;
    hipStream_t s = p->h->stream;
    FLX_TRY(hipEventRecord(p->ev0, s));
    FLX_TRY(hipStreamWaitEvent(p->gs, p->ev0, 0));
    FLX_TRY(hipGraphLaunch(p->exec, p->gs));
    FLX_TRY(hipEventRecord(p->ev1, p->gs));
    FLX_TRY(hipStreamWaitEvent(s, p->ev1, 0));
    return FLX_SUCCESS;
}

extern "C" int flx_batch_config(flx_batch_t p, int *cfg, double *tuned_ms)
{
    if (!p || !cfg) return FLX_INVALID_ARG;
    cfg[0] = p->tile_cfg;
    cfg[1] = p->grid;
    if (tuned_ms) *tuned_ms = p->tuned_ms;
    return FLX_SUCCESS;
}

extern "C" int flx_batch_destroy(flx_batch_t p)
{
    if (!p) return FLX_INVALID_ARG;
    if (p->gs) (void)hipStreamSynchronize(p->gs);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->gs) (void)hipStreamDestroy(p->gs);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    if (p->probs) (void)hipFree(p->probs);
    for (void *g : p->gprobs)
        if (g) (void)hipFree(g);
    if (p->counter) (void)hipFree(p->counter);
    if (p->W32) (void)hipFree(p->W32);
    if (p->R) (void)hipFree(p->R);
    delete p;
    return FLX_SUCCESS;
}
