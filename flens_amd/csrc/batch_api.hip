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
    int n, nrhs, piv, id, cfg, pad;
};

// size bins of the solve autotuner: n in [8 * 2^b, 16 * 2^b), b = 0..5
// (n < 8 in bin 0, n = 512 in bin 5)
constexpr int SOLVE_BINS = 6;
inline int solve_bin(int n)
{
    int b = 0;
    while (b < SOLVE_BINS - 1 && n >= (16 << b)) ++b;
    return b;
}
struct GemmProbH {    // layout of flx::GemmProb (batched.hip)
    long a, b, c;
    int m, n, k, tile0, cfg, id;
    double alpha, beta;
};

// size bins of the GEMM autotuner: max(m, n) in [16 * 2^b, 32 * 2^b),
// b = 0..5 (below 16 in bin 0, 512 and above in bin 5)
inline int gemm_bin(int m, int n)
{
    const int d = m > n ? m : n;
    int b = 0;
    while (b < SOLVE_BINS - 1 && d >= (32 << b)) ++b;
    return b;
}

} // namespace

struct flx_batch {
    flx_context *h = nullptr;
    int kind = 0;               // 0 solves, 1 GEMMs
    int mode = 0, itmax = 5;
    int policy = 0;             // batched_gesv policy of the mode
    std::vector<BatchProbH> tab;   // solves: the problem table (host copy)
    int bin_tile[SOLVE_BINS] = {0, 0, 0, 0, 0, 0};
    double *G = nullptr;        // scratch of the low-precision ladders
    int Ggrid = 0;
    double *stats = nullptr;    // row statistics per system (pre-pass)
    float *selT = nullptr;      // the selector transposed for the kernel
    unsigned long long *gticks = nullptr;   // GEMM tuning: per-problem time
    int *cls = nullptr;         // the selector's class per system
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
    int ntiles = 0;
    std::vector<GemmProbH> gtab;   // problem table (host copy), tile0 per the cfgs
    void *gprobs = nullptr;
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
        return batched_gesv(s, p->policy, p->probs, p->count, p->counter, A, B, p->W32, p->R,
                            grid <= p->Ggrid ? p->G : nullptr, p->ipiv, p->info, p->iters,
                            p->itmax, grid, p->out);
    (void)tile_cfg;
    return batched_gemm(s, p->gprobs, p->count, p->ntiles, p->counter, p->GA, p->GB, C, grid,
                        p->gticks);
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

// device buffers and events of one tuning session, released on every exit
struct TuneScratch {
    std::vector<void *> bufs;
    std::vector<hipEvent_t> evs;
    ~TuneScratch()
    {
        for (void *b : bufs) (void)hipFree(b);
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    }
    template <typename T> hipError_t alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc(reinterpret_cast<void **>(p), sizeof(T) * std::max<size_t>(1, n));
        if (e == hipSuccess) bufs.push_back(*p);
        return e;
    }
    hipError_t event(hipEvent_t *e)
    {
        hipError_t r = hipEventCreate(e);
        if (r == hipSuccess) evs.push_back(*e);
        return r;
    }
};

// one timed run of a candidate on copies of the inputs (the solves work in
// place); returns milliseconds in *ms
static int timed_run(flx_batch *p, TuneScratch &ts, int cfg, int grid, double *tA, double *tB,
                     double *tC, float *ms)
{
    hipStream_t s = p->h->stream;
    hipEvent_t e0, e1;
    FLX_TRY(ts.event(&e0));
    FLX_TRY(ts.event(&e1));
    if (p->kind == 0) {
        FLX_TRY(hipMemcpyAsync(tA, p->A, sizeof(double) * p->a_total, hipMemcpyDeviceToDevice, s));
        FLX_TRY(hipMemcpyAsync(tB, p->B, sizeof(double) * p->b_total, hipMemcpyDeviceToDevice, s));
    } else {
        FLX_TRY(hipMemcpyAsync(tC, p->GC, sizeof(double) * p->c_total, hipMemcpyDeviceToDevice, s));
    }
    FLX_TRY(hipEventRecord(e0, s));
    FLX_TRY(launch(p, s, cfg, grid, tA, tB, tC));
    FLX_TRY(hipEventRecord(e1, s));
    FLX_TRY(hipEventSynchronize(e1));
    FLX_TRY(hipEventElapsedTime(ms, e0, e1));
    return FLX_SUCCESS;
}

// tile0 of every GEMM problem from its cfg; uploads the table
static int upload_gemm_table(flx_batch *p)
{
    int tiles = 0;
    for (auto &g : p->gtab) {
        int tm = 0, tn = 0;
        batched_gemm_tile(g.cfg, &tm, &tn);
        g.tile0 = tiles;
        tiles += ((g.m + tm - 1) / tm) * ((g.n + tn - 1) / tn);
    }
    p->ntiles = tiles;
    if (!p->gtab.empty())
        FLX_TRY(hipMemcpy(p->gprobs, p->gtab.data(), sizeof(GemmProbH) * p->gtab.size(),
                          hipMemcpyHostToDevice));
    return FLX_SUCCESS;
}

static int upload_table(flx_batch *p)
{
    if (p->count > 0)
        FLX_TRY(hipMemcpy(p->probs, p->tab.data(), sizeof(BatchProbH) * p->count,
                          hipMemcpyHostToDevice));
    return FLX_SUCCESS;
}

// The autotuner.  Solves: every system runs once with each tile
// configuration; the in-kernel times, summed per size bin (solve_bin),
// give each bin its tile (the table's tile bits), then the persistent grid
// is timed.  GEMMs: tile shape x grid, timed as whole batches.
static int autotune(flx_batch *p, bool tune)
{
    const int cus = cu_count(p->h);
    if (!tune) {
        p->grid = 2 * cus;
        p->tile_cfg = (p->kind == 1) ? 1 : 0;
        return FLX_SUCCESS;
    }
    TuneScratch ts;
    double *tA = nullptr, *tB = nullptr, *tC = nullptr;
    if (p->kind == 0) {
        FLX_TRY(ts.alloc(&tA, p->a_total));
        FLX_TRY(ts.alloc(&tB, p->b_total));
    } else {
        FLX_TRY(ts.alloc(&tC, p->c_total));
    }
    double best = 1e30;
    int bcfg = 0, bgrid = 2 * cus;
    if (p->kind == 0) {
        long long *ticks = nullptr;
        FLX_TRY(ts.alloc(&ticks, p->count));
        const BatchOut saved = p->out;
        BatchOut tout;
        tout.tsys = ticks;
        tout.sel = saved.sel;
        tout.stats = saved.stats;
        tout.selT = saved.selT;
        tout.cls = saved.cls;
        p->out = tout;
        double sum[2][SOLVE_BINS] = {};
        std::vector<long long> h(p->count);
        int st = FLX_SUCCESS;
        for (int tile = 0; tile < 2 && st == FLX_SUCCESS; ++tile) {
            for (auto &e : p->tab) e.cfg = (e.cfg & 3) | (tile << 2);
            st = upload_table(p);
            float ms = 0.f;
            if (st == FLX_SUCCESS) st = timed_run(p, ts, 0, 2 * cus, tA, tB, tC, &ms);
            if (st == FLX_SUCCESS &&
                hipMemcpy(h.data(), ticks, sizeof(long long) * p->count, hipMemcpyDeviceToHost) !=
                    hipSuccess)
                st = FLX_HIP_ERROR;
            if (st == FLX_SUCCESS)
                for (const auto &e : p->tab) sum[tile][solve_bin(e.n)] += (double)h[e.id];
        }
        p->out = saved;
        FLX_RET(st);
        for (int b = 0; b < SOLVE_BINS; ++b) p->bin_tile[b] = sum[1][b] < sum[0][b] ? 1 : 0;
        for (auto &e : p->tab) e.cfg = (e.cfg & 3) | (p->bin_tile[solve_bin(e.n)] << 2);
        FLX_RET(upload_table(p));
        for (int grid : {cus, 2 * cus}) {
            float ms_best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                float ms = 0.f;
                FLX_RET(timed_run(p, ts, 0, grid, tA, tB, tC, &ms));
                ms_best = std::min(ms_best, ms);
            }
            if (ms_best < best) { best = ms_best; bgrid = grid; }
        }
    } else {
        // every problem under each tile shape, its tiles' in-kernel time
        // summed per size bin (gemm_bin), then the grid
        unsigned long long *gt = nullptr;
        FLX_TRY(ts.alloc(&gt, p->gtab.size()));
        double sum[3][SOLVE_BINS] = {};
        std::vector<unsigned long long> hv(p->gtab.size());
        for (int cfg = 0; cfg < 3; ++cfg) {
            for (auto &g : p->gtab) g.cfg = cfg;
            FLX_RET(upload_gemm_table(p));
            FLX_TRY(hipMemset(gt, 0, sizeof(unsigned long long) * std::max<size_t>(1, hv.size())));
            p->gticks = gt;
            float ms = 0.f;
            const int st = timed_run(p, ts, 0, 2 * cus, tA, tB, tC, &ms);
            p->gticks = nullptr;
            FLX_RET(st);
            if (!hv.empty())
                FLX_TRY(hipMemcpy(hv.data(), gt, sizeof(unsigned long long) * hv.size(),
                                  hipMemcpyDeviceToHost));
            for (size_t i = 0; i < p->gtab.size(); ++i)
                sum[cfg][gemm_bin(p->gtab[i].m, p->gtab[i].n)] += (double)hv[i];
        }
        for (int b = 0; b < SOLVE_BINS; ++b) {
            int bc = 0;
            for (int cfg = 1; cfg < 3; ++cfg)
                if (sum[cfg][b] < sum[bc][b]) bc = cfg;
            p->bin_tile[b] = bc;
        }
        for (auto &g : p->gtab) g.cfg = p->bin_tile[gemm_bin(g.m, g.n)];
        FLX_RET(upload_gemm_table(p));
        for (int grid : {cus, 2 * cus, 4 * cus}) {
            float ms_best = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                float ms = 0.f;
                FLX_RET(timed_run(p, ts, 0, grid, tA, tB, tC, &ms));
                ms_best = std::min(ms_best, ms);
            }
            if (ms_best < best) { best = ms_best; bgrid = grid; }
        }
        bcfg = -1;   // per size bin
    }
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
        mode < 0 || mode > FLX_BATCH_TABLE || itmax < 0 || itmax > 100)
        return FLX_INVALID_ARG;
    if (mode == FLX_BATCH_LEARNED && (!extra || !extra->selector)) return FLX_INVALID_ARG;
    if (mode == FLX_BATCH_TABLE && (!extra || !extra->cfg)) return FLX_INVALID_ARG;
    if (batched_prob_bytes() != sizeof(BatchProbH)) return FLX_HIP_ERROR;
    // the ladder of every system: fixed by the mode, or the caller's table
    static const int LADDER_OF_MODE[] = {0, 1, 0, 2, 3, 1, 0};
    std::vector<BatchProbH> tab(count);
    long ao = 0, bo = 0;
    int po = 0;
    for (int i = 0; i < count; ++i) {
        if (n[i] < 1 || n[i] > 512 || nrhs[i] < 0 || nrhs[i] > 8) return FLX_UNSUPPORTED;
        int cfg = LADDER_OF_MODE[mode];
        if (mode == FLX_BATCH_TABLE) {
            cfg = extra->cfg[i];
            if (cfg < 0 || cfg > 7) return FLX_INVALID_ARG;
        }
        tab[i] = BatchProbH{ao, bo, n[i], nrhs[i], po, i, cfg, 0};
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
    p->policy = mode == FLX_BATCH_LEARNED ? 1 : mode == FLX_BATCH_RULE ? 2 : 0;
    p->itmax = itmax;
    p->count = count;
    p->tab = std::move(tab);
    p->A = A; p->B = B; p->ipiv = ipiv; p->info = info; p->iters = iters;
    p->a_total = ao;
    p->b_total = bo;
    if (extra) {
        p->out.sel = extra->selector;
        p->out.feat = extra->features;
        p->out.tsys = extra->ticks;
        p->out.choice = extra->choice;
        p->out.inner = extra->inner;
        p->out.used = extra->used;
    }
    auto fail = [&](int st) { flx_batch_destroy(p); return st; };
    if (hipMalloc(&p->probs, sizeof(BatchProbH) * std::max(1, count)) != hipSuccess ||
        hipMalloc(&p->counter, sizeof(int)) != hipSuccess)
        return fail(FLX_ALLOC_FAILED);
    if (mode != FLX_BATCH_FP64) {
        p->Ggrid = 2 * cu_count(h);
        if (hipMalloc(&p->W32, sizeof(float) * std::max(1L, ao)) != hipSuccess ||
            hipMalloc(&p->R, sizeof(double) * std::max(1L, bo)) != hipSuccess ||
            hipMalloc(&p->G, sizeof(double) * batched_scratch_doubles(p->Ggrid)) != hipSuccess)
            return fail(FLX_ALLOC_FAILED);
    }
    if (mode != FLX_BATCH_FP64 || p->out.feat) {
        if (hipMalloc(&p->stats, sizeof(double) * 3 * std::max(1, count)) != hipSuccess)
            return fail(FLX_ALLOC_FAILED);
        p->out.stats = p->stats;
    }
    if (mode == FLX_BATCH_LEARNED) {
        if (hipMalloc(&p->selT, sizeof(float) * FLX_MLP_PARAMS) != hipSuccess ||
            hipMalloc(&p->cls, sizeof(int) * std::max(1, count)) != hipSuccess)
            return fail(FLX_ALLOC_FAILED);
        p->out.selT = p->selT;
        p->out.cls = p->cls;
    }
    int st = upload_table(p);
    if (st) return fail(st);
    st = autotune(p, tune != 0);
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
    p->gtab.reserve(count);
    for (int i : order) {
        if (m[i] == 0 || n[i] == 0) continue;
        // id: the position in the table (the tuner's per-problem time slot)
        const int pos = (int)p->gtab.size();
        p->gtab.push_back(GemmProbH{ao[i], bo[i], co[i], m[i], n[i], k[i], 0, 1, pos, alpha[i],
                                    beta[i]});
    }
    p->count = (int)p->gtab.size();
    for (int b = 0; b < SOLVE_BINS; ++b) p->bin_tile[b] = 1;
    if (hipMalloc(&p->gprobs, sizeof(GemmProbH) * std::max<size_t>(1, p->gtab.size())) != hipSuccess)
        return fail(FLX_ALLOC_FAILED);
    {
        const int st0 = upload_gemm_table(p);
        if (st0) return fail(st0);
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

extern "C" int flx_batch_bin_tiles(flx_batch_t p, int *tiles, int nbins)
{
    if (!p || !tiles || nbins < SOLVE_BINS) return FLX_INVALID_ARG;
    for (int b = 0; b < SOLVE_BINS; ++b) tiles[b] = p->bin_tile[b];
    return SOLVE_BINS;
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
    if (p->gprobs) (void)hipFree(p->gprobs);
    if (p->counter) (void)hipFree(p->counter);
    if (p->W32) (void)hipFree(p->W32);
    if (p->R) (void)hipFree(p->R);
    if (p->G) (void)hipFree(p->G);
    if (p->stats) (void)hipFree(p->stats);
    if (p->selT) (void)hipFree(p->selT);
    if (p->cls) (void)hipFree(p->cls);
    delete p;
    return FLX_SUCCESS;
}
