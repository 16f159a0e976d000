// noalloc.hip -- the FP64 LU entry point that allocates nothing.
//
// flx_dgetrf_noalloc runs the same kernels as flx_dgetrf (transpose to the
// row-major working copy, cooperative panels, interchanges, trsm, trailing
// GEMM, transpose back; flens/lapack/ge/trf.tcc:59-152 semantics) with every
// scratch buffer carved out of one file-scope device array, so a caller that
// cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc) can factor matrices up to the array's size.
//
// The array is per device: at a device's first call its address is looked
// up once (hipGetSymbolAddress, with that device current) and kept; kernels
// only ever receive that address.  Every call on a device shares the array,
// so calls that overlap on one device race; the panel's exchange words stay
// valid across calls because the panel generation counter is kept per device
// too.
#include <climits>
#include <mutex>
#include "context.h"

using namespace flx;

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)FLX_NOALLOC_SCRATCH_MIB << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

struct DeviceScratch {
    unsigned char *ptr = nullptr;   // device address of scratch on this device
    uint32_t gen = 0;               // panel generation counter of the exchange words
};
std::mutex g_mu;
DeviceScratch g_dev[NOALLOC_MAX_DEVICES];

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// byte offsets of the working buffers inside the scratch array
struct Layout {
    size_t exch, flags, list, lsc, perm, piv0, W, total;
};

Layout layout(int m, int n)
{
    Layout L;
    const int mn = m < n ? m : n;
    size_t o = 0;
    L.exch = o;  o = align256(o + sizeof(uint64_t) * lu_exch_words());
    L.flags = o; o = align256(o + sizeof(int) * 4);
    L.list = o;  o = align256(o + sizeof(int) * PANEL_LIST_SLICE * PANEL_GROUPS_MAX);
    L.lsc = o;   o = align256(o + sizeof(double) * lu_lsc_doubles(m));
    L.perm = o;  o = align256(o + sizeof(int) * (size_t)m);
    L.piv0 = o;  o = align256(o + sizeof(int) * (size_t)mn);
    L.W = o;     o = align256(o + sizeof(double) * (size_t)round_ld(n) * (size_t)m);
    L.total = o;
    return L;
}

__global__ void zero_words_kernel(uint64_t *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// 1-based pivots; INFO = 0 when no zero pivot was found, INT_MIN when a
// bounded wait of the panels gave up (the factors are then not valid)
__global__ void noalloc_finish_kernel(const int *__restrict__ piv0, int *__restrict__ ipiv,
                                      int mn, int *__restrict__ info,
                                      const int *__restrict__ flags)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < mn) ipiv[i] = piv0[i] + 1;
    if (i == 0) {
        if (flags[0]) *info = INT_MIN;
        else if (*info == INT_MAX) *info = 0;
    }
}

} // namespace

extern "C" size_t flx_noalloc_bytes(int m, int n)
{
    if (m <= 0 || n <= 0) return 0;
    return layout(m, n).total;
}

extern "C" int flx_dgetrf_noalloc(flx_handle h, int m, int n, double *A, int lda, int *ipiv,
                                  int *info)
{
    if (!h || m < 0 || n < 0 || lda < (m > 1 ? m : 1) || !info || (m > 0 && n > 0 && !ipiv))
        return FLX_INVALID_ARG;
    hipStream_t s = h->stream;
    if (m == 0 || n == 0) {
        FLX_TRY(fill_int(s, info, 1, 0));
        return FLX_SUCCESS;
    }
    const Layout L = layout(m, n);
    if (L.total > NOALLOC_SCRATCH_BYTES) return FLX_SCRATCH_LIMIT;
    int dev = -1;
    FLX_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES || dev != h->device) return FLX_INVALID_ARG;

    std::lock_guard<std::mutex> lock(g_mu);
    DeviceScratch &d = g_dev[dev];
    if (!d.ptr) {
        // set-up, once per device, with that device current
        void *p = nullptr;
        FLX_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return FLX_HIP_ERROR;
        // exchange words start with no tag (a fill kernel: capture-safe)
        const size_t words = lu_exch_words();
        hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256),
                           0, s, (uint64_t *)((unsigned char *)p + L.exch), words);
        FLX_TRY(hipGetLastError());
        d.ptr = (unsigned char *)p;
    }
    unsigned char *base = d.ptr;
    LuWork wk;
    wk.exch = (uint64_t *)(base + L.exch);
    wk.exch_words = lu_exch_words();
    wk.flags = (int *)(base + L.flags);
    wk.list = (int *)(base + L.list);
    wk.lsc = (double *)(base + L.lsc);
    wk.perm = (int *)(base + L.perm);
    wk.perm_len = m;
    wk.gen = d.gen;
    // no lookahead (it needs a second stream and events, created on demand);
    // the results are the same
    LuOptions opt = h->opt;
    opt.lookahead = 0;
    double *W = (double *)(base + L.W);
    int *piv0 = (int *)(base + L.piv0);
    const long ldw = round_ld(n);
    const int mn = m < n ? m : n;
    FLX_TRY(fill_int(s, wk.flags, 4, 0));         // the last call's flag is not this one's
    FLX_TRY(fill_int(s, info, 1, INT_MAX));
    FLX_TRY(transpose_f64(s, A, lda, W, ldw, m, n));
    hipError_t e = getrf_rm_f64(s, wk, W, ldw, m, n, piv0, info, opt);
    d.gen = wk.gen;
    FLX_TRY(e);
    FLX_TRY(transpose_f64(s, W, ldw, A, lda, n, m));
    hipLaunchKernelGGL(noalloc_finish_kernel, dim3((mn + 255) / 256), dim3(256), 0, s,
                       (const int *)piv0, ipiv, mn, info, (const int *)wk.flags);
    FLX_TRY(hipGetLastError());
    return FLX_SUCCESS;
}
