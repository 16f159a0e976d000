// getrf.hip -- blocked right-looking LU with partial pivoting (FP64).
//
// Replaces flens::lapack::trf (flens/lapack/ge/trf.tcc:59-152) and its
// unblocked panel tf2 (flens/lapack/ge/tf2.tcc:56-136), laswp
// (flens/lapack/la/laswp.tcc:54-93) and the two BLAS-3 calls of every block
// step: sm (trsm, trf.tcc:136-138) and mm (trailing update, :144-146).
//
// Layout: the factorisation runs on a ROW-major working copy W of A (the
// column-major caller layout is transposed on entry and exit).  Row
// interchanges -- the traffic partial pivoting generates every column --
// then move contiguous row segments instead of one 8-byte element per
// 128-byte line, and a panel row is a contiguous run.  The trailing update
// W22 -= W21*W12 is the column-major GEMM C(n'xm') -= A(n'xk)*B(kxm') on the
// same memory, so it runs on the MFMA GEMM unchanged.
//
// Blocking: outer blocks of nb columns (trailing-update K = nb); inside an
// outer block, sub-panels of ib <= 64 columns are factored by one
// cooperative kernel each.
#include <climits>
#include <cfloat>
#include <type_traits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

// ---------------------------------------------------------------------------
// Panel: cooperative unblocked LU of an mp x w row-major panel (w <= W).
//
// One thread owns one panel row and keeps it in registers (fully unrolled
// column loop, so every register index is static).  Per column j:
//   1. block-local pivot search (wave shuffles + LDS), first max on ties;
//   2. every workgroup posts its best candidate (|v|, row, the row's data)
//      and the owner of row j posts row j, as sequence-tagged 64-bit words;
//   3. every workgroup reads all posts (one agent-scope round trip, no
//      fence, no counter), picks the same winner, swaps, scales and applies
//      the rank-1 update to its own rows.
// The posts are double-buffered by column parity: a workgroup can only
// overwrite a slot after every workgroup has posted the next column, which
// each does only after it finished reading the previous one.
// ---------------------------------------------------------------------------
constexpr int PANEL_T = 256;
constexpr int PANEL_GMAX = 64;          // panel height <= 64*256 = 16384
constexpr int PANEL_WMAX = 64;
constexpr uint32_t SPIN_LIMIT = 1u << 20;   // ~1 s of polling per word

// words per parity: headers [GMAX][4], rows [GMAX][2*WMAX], diag [2*WMAX]
constexpr size_t EX_HDR = 0;
constexpr size_t EX_ROWS = PANEL_GMAX * 4;
constexpr size_t EX_DIAG = EX_ROWS + (size_t)PANEL_GMAX * 2 * PANEL_WMAX;
constexpr size_t EX_PAR = EX_DIAG + 2 * PANEL_WMAX;

// Compile-time loop: the body sees its index as a constant, so arrays
// indexed by it stay in registers.
template <int I, int N, typename F>
FLX_DEV void static_for(F &&f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

size_t lu_exch_words() { return 2 * EX_PAR; }

template <int W>
__global__ __launch_bounds__(PANEL_T) void panel_kernel(
    double *__restrict__ P, long ldw, int mp, int w, uint64_t *__restrict__ ex,
    uint32_t gen, int *__restrict__ piv, int piv_base, int *__restrict__ list,
    int *__restrict__ info, int info_base, int *__restrict__ err)
{
    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int r = g * PANEL_T + tid;
    const bool valid = r < mp;
    const int mn = mp < w ? mp : w;

    __shared__ double s_cand[PANEL_GMAX][W];
    __shared__ double s_diag[W];
    __shared__ double s_habs[PANEL_GMAX];
    __shared__ int s_hidx[PANEL_GMAX];
    __shared__ double s_wabs[4];
    __shared__ int s_widx[4];
    __shared__ int s_top[W];
    __shared__ int s_orow[W];
    __shared__ int s_osrc[W];
    __shared__ int s_ocnt;
    __shared__ int s_fail;

    double row[W];
#pragma unroll
    for (int c = 0; c < W; ++c)
        row[c] = (valid && c < w) ? P[(long)r * ldw + c] : 0.0;

    if (g == 0)
        for (int i = tid; i < W; i += PANEL_T) s_top[i] = i;
    if (tid == 0) { s_ocnt = 0; s_fail = 0; }
    int first_zero = 0;
    __syncthreads();

    static_for<0, W>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if (j >= mn) return;
        // ---- 1. local pivot candidate --------------------------------------
        const bool act = valid && r >= j;
        double a = act ? fabs(row[j]) : -1.0;
        int ia = act ? r : INT_MAX;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            double a2 = __shfl_xor(a, o, 64);
            int i2 = __shfl_xor(ia, o, 64);
            if (cand_better(a2, i2, a, ia)) { a = a2; ia = i2; }
        }
        if (lane == 0) { s_wabs[wave] = a; s_widx[wave] = ia; }
        __syncthreads();
        double ba = s_wabs[0];
        int bi = s_widx[0];
#pragma unroll
        for (int q = 1; q < 4; ++q)
            if (cand_better(s_wabs[q], s_widx[q], ba, bi)) { ba = s_wabs[q]; bi = s_widx[q]; }
        if (r == bi) {
#pragma unroll
            for (int c = 0; c < W; ++c) s_cand[g][c] = row[c];
        }
        if (r == j) {
#pragma unroll
            for (int c = 0; c < W; ++c) s_diag[c] = row[c];
        }
        __syncthreads();

        // ---- 2./3. all-gather across workgroups ---------------------------
        int win = bi, wg = g;
        if (G > 1) {
            const uint32_t seq = (gen << 8) | (uint32_t)(j + 1);
            uint64_t *base = ex + (size_t)(j & 1) * EX_PAR;
            const int diag_owner = j / PANEL_T;
            if (tid < 2 * W) {
                const int c = tid >> 1;
                uint64_t b = (uint64_t)__double_as_longlong(s_cand[g][c]);
                uint64_t half = (tid & 1) ? (b & 0xffffffffull) : (b >> 32);
                st_agent(base + EX_ROWS + (size_t)g * 2 * PANEL_WMAX + tid,
                         ((uint64_t)seq << 32) | half);
                if (g == diag_owner) {
                    uint64_t d = (uint64_t)__double_as_longlong(s_diag[c]);
                    uint64_t dh = (tid & 1) ? (d & 0xffffffffull) : (d >> 32);
                    st_agent(base + EX_DIAG + tid, ((uint64_t)seq << 32) | dh);
                }
            }
            if (tid == 0) {
                uint64_t b = (uint64_t)__double_as_longlong(ba);
                st_agent(base + EX_HDR + g * 4 + 0, ((uint64_t)seq << 32) | (b >> 32));
                st_agent(base + EX_HDR + g * 4 + 1, ((uint64_t)seq << 32) | (b & 0xffffffffull));
                st_agent(base + EX_HDR + g * 4 + 2, ((uint64_t)seq << 32) | (uint32_t)bi);
            }
            // read every post: units = G headers + G*W row values + W diag
            const int units = G + G * W + W;
            for (int u = tid; u < units; u += PANEL_T) {
                const uint64_t *p0;
                int nw;
                if (u < G) {
                    p0 = base + EX_HDR + u * 4; nw = 3;
                } else if (u < G + G * W) {
                    int gg = (u - G) / W, c = (u - G) % W;
                    if (gg == g) continue;
                    p0 = base + EX_ROWS + (size_t)gg * 2 * PANEL_WMAX + 2 * c; nw = 2;
                } else {
                    if (g == diag_owner) continue;
                    p0 = base + EX_DIAG + 2 * (u - G - G * W); nw = 2;
                }
                uint64_t v[3];
                for (int q = 0; q < nw; ++q) {
                    uint32_t spins = 0;
                    uint64_t x = ld_agent(p0 + q);
                    while (seq_of(x) != seq) {
                        // give up for good once any wait of this workgroup failed
                        if (*(volatile int *)&s_fail || ++spins > SPIN_LIMIT) { s_fail = 1; break; }
                        __builtin_amdgcn_s_sleep(1);
                        x = ld_agent(p0 + q);
                    }
                    v[q] = x;
                }
                if (u < G) {
                    s_habs[u] = join_double(v[0], v[1]);
                    s_hidx[u] = (int)(uint32_t)(v[2] & 0xffffffffull);
                } else if (u < G + G * W) {
                    int gg = (u - G) / W, c = (u - G) % W;
                    s_cand[gg][c] = join_double(v[0], v[1]);
                } else {
                    s_diag[u - G - G * W] = join_double(v[0], v[1]);
                }
            }
            __syncthreads();
            double wa = s_habs[0];
            win = s_hidx[0];
            for (int q = 1; q < G; ++q)
                if (cand_better(s_habs[q], s_hidx[q], wa, win)) { wa = s_habs[q]; win = s_hidx[q]; }
            // a winner outside the panel can only come from a failed wait:
            // keep every later index in bounds and report the failure
            if (win < j || win >= mp) { win = j; s_fail = 1; }
            wg = win / PANEL_T;
        }

        // ---- swap, scale, rank-1 update ------------------------------------
        const double *u = s_cand[wg];
        const double pv = u[j];
        if (win != j) {
            if (r == j) {
#pragma unroll
                for (int c = 0; c < W; ++c) row[c] = u[c];
            } else if (r == win) {
#pragma unroll
                for (int c = 0; c < W; ++c) row[c] = s_diag[c];
            }
        }
        if (valid && r > j) {
            double l = row[j];
            if (pv != 0.0) {
                l = (fabs(pv) >= DBL_MIN) ? l * (1.0 / pv) : l / pv;
                row[j] = l;
            }
#pragma unroll
            for (int c = j + 1; c < W; ++c) row[c] = fma(-l, u[c], row[c]);
        }
        if (g == 0) {
            if (tid == 0) {
                piv[j] = piv_base + win;
                if (pv == 0.0 && first_zero == 0) first_zero = j + 1;
            }
            if (wave == 1 && win != j) {
                // track the composed permutation of this panel's swaps
                if (win < w) {
                    if (lane == 0) {
                        int t = s_top[j]; s_top[j] = s_top[win]; s_top[win] = t;
                    }
                } else {
                    const int cnt = s_ocnt;
                    int found = -1;
                    for (int q0 = 0; q0 < cnt; q0 += 64) {
                        int q = q0 + lane;
                        unsigned long long m = __ballot(q < cnt && s_orow[q] == win);
                        if (m) { found = q0 + __ffsll(m) - 1; break; }
                    }
                    if (lane == 0) {
                        if (found < 0) {
                            found = cnt;
                            s_orow[cnt] = win;
                            s_osrc[cnt] = win;
                            s_ocnt = cnt + 1;
                        }
                        int t = s_top[j]; s_top[j] = s_osrc[found]; s_osrc[found] = t;
                    }
                }
            }
        }
        __syncthreads();
    });

    if (valid) {
#pragma unroll
        for (int c = 0; c < W; ++c)
            if (c < w) P[(long)r * ldw + c] = row[c];
    }
    if (g == 0 && tid == 0) {
        // list = [count, dst_0, src_0, dst_1, src_1, ...] in absolute rows
        int cnt = 0;
        for (int k = 0; k < mn; ++k)
            if (s_top[k] != k) {
                list[1 + 2 * cnt] = piv_base + k;
                list[2 + 2 * cnt] = piv_base + s_top[k];
                ++cnt;
            }
        for (int q = 0; q < s_ocnt; ++q)
            if (s_osrc[q] != s_orow[q]) {
                list[1 + 2 * cnt] = piv_base + s_orow[q];
                list[2 + 2 * cnt] = piv_base + s_osrc[q];
                ++cnt;
            }
        list[0] = cnt;
        if (first_zero) atomicMin(info, info_base + first_zero);
    }
    if (tid == 0 && s_fail) atomicExch(err, 1);
}

// ---------------------------------------------------------------------------
// Row permutation of a row-major matrix outside one panel:
//   W[dst_e][c] = W_old[src_e][c]   for every list entry e,
// for columns c in [0, c1) U [c2, n).  Each workgroup owns a 64-column
// chunk: it reads every source segment into LDS, then writes (no hazard
// between chunks).  This is laswp (flens/lapack/la/laswp.tcc:54-93) for
// both column ranges of trf.tcc:126 and :132 at once.
// ---------------------------------------------------------------------------
constexpr int RP_CW = 64;
constexpr int RP_MAXE = 2 * PANEL_WMAX;

__global__ __launch_bounds__(256) void rowperm_kernel(double *__restrict__ Wm,
                                                      long ldw, int m, int c1, int c2,
                                                      int n,
                                                      const int *__restrict__ list,
                                                      int *__restrict__ perm,
                                                      unsigned long long *__restrict__ moved)
{
    __shared__ double tile[RP_MAXE][RP_CW];
    __shared__ int s_dst[RP_MAXE], s_src[RP_MAXE];
    const int cnt = min(list[0], RP_MAXE);
    if (cnt <= 0) return;
    const int nleft = (c1 + RP_CW - 1) / RP_CW;
    int chunk = blockIdx.x;
    int cbeg, cend;
    if (chunk < nleft) { cbeg = chunk * RP_CW; cend = min(c1, cbeg + RP_CW); }
    else { cbeg = c2 + (chunk - nleft) * RP_CW; cend = min(n, cbeg + RP_CW); }
    for (int e = threadIdx.x; e < cnt; e += 256) {
        int d = list[1 + 2 * e], sr = list[2 + 2 * e];
        bool ok = d >= 0 && d < m && sr >= 0 && sr < m;
        s_dst[e] = ok ? d : 0;
        s_src[e] = ok ? sr : 0;   // an invalid entry degenerates to row 0 -> row 0
    }
    __syncthreads();
    const int cw = cend - cbeg;
    if (moved && threadIdx.x == 0) atomicAdd(moved, (unsigned long long)cnt * cw);
    for (int q = threadIdx.x; q < cnt * RP_CW; q += 256) {
        int e = q / RP_CW, c = q % RP_CW;
        if (c < cw) tile[e][c] = Wm[(long)s_src[e] * ldw + cbeg + c];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < cnt * RP_CW; q += 256) {
        int e = q / RP_CW, c = q % RP_CW;
        if (c < cw) Wm[(long)s_dst[e] * ldw + cbeg + c] = tile[e][c];
    }
    // block 0 also composes the global row permutation
    if (blockIdx.x == 0 && perm) {
        __syncthreads();
        int v[2] = {0, 0};
        for (int e = threadIdx.x, k = 0; e < cnt; e += 256, ++k) v[k] = perm[s_src[e]];
        __syncthreads();
        for (int e = threadIdx.x, k = 0; e < cnt; e += 256, ++k) perm[s_dst[e]] = v[k];
    }
}

// permutation-only variant, used when no column lies outside the panel
__global__ void perm_compose_kernel(const int *__restrict__ list, int *__restrict__ perm, int m)
{
    const int cnt = min(list[0], RP_MAXE);
    int v[2] = {0, 0};
    for (int e = threadIdx.x, k = 0; e < cnt; e += 256, ++k) {
        int sr = list[2 + 2 * e];
        v[k] = (sr >= 0 && sr < m) ? perm[sr] : 0;
    }
    __syncthreads();
    for (int e = threadIdx.x, k = 0; e < cnt; e += 256, ++k) {
        int d = list[1 + 2 * e];
        if (d >= 0 && d < m) perm[d] = v[k];
    }
}

// ---------------------------------------------------------------------------
// Row-major unit-lower triangular solve X <- L^{-1} X (L k x k, X k x ncols),
// blas::sm(Left, NoTrans, 1, A11.lowerUnit(), A12) of trf.tcc:136-138.
// One workgroup per 32-column chunk; X stays in LDS; 16-row diagonal blocks
// are solved in place, then folded into the rows below.
// ---------------------------------------------------------------------------
constexpr int TR_CW = 32;
constexpr int TR_KMAX = 256;

__global__ __launch_bounds__(256) void trsm_rm_llu_kernel(const double *__restrict__ L,
                                                          long ldl, double *__restrict__ X,
                                                          long ldx, int k, int ncols)
{
    __shared__ double xs[TR_KMAX][TR_CW + 1];
    const int c0 = blockIdx.x * TR_CW;
    const int cw = min(TR_CW, ncols - c0);
    const int tid = threadIdx.x;
    const int tc = tid % TR_CW, tg = tid / TR_CW;   // 8 row groups
    for (int q = tid; q < k * TR_CW; q += 256) {
        int rr = q / TR_CW, c = q % TR_CW;
        xs[rr][c] = (c < cw) ? X[(long)rr * ldx + c0 + c] : 0.0;
    }
    __syncthreads();
    for (int kb = 0; kb < k; kb += 16) {
        const int bs = min(16, k - kb);
        // diagonal block: one thread per column, sequential rows
        if (tid < TR_CW) {
            for (int i = 1; i < bs; ++i) {
                double s = xs[kb + i][tc];
                const double *lr = L + (long)(kb + i) * ldl + kb;
                for (int q = 0; q < i; ++q) s = fma(-lr[q], xs[kb + q][tc], s);
                xs[kb + i][tc] = s;
            }
        }
        __syncthreads();
        // rows below
        double xb[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) xb[q] = (q < bs) ? xs[kb + q][tc] : 0.0;
        for (int rr = kb + bs + tg; rr < k; rr += 8) {
            const double *lr = L + (long)rr * ldl + kb;
            double s = xs[rr][tc];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q < bs) s = fma(-lr[q], xb[q], s);
            xs[rr][tc] = s;
        }
        __syncthreads();
    }
    for (int q = tid; q < k * TR_CW; q += 256) {
        int rr = q / TR_CW, c = q % TR_CW;
        if (c < cw) X[(long)rr * ldx + c0 + c] = xs[rr][c];
    }
}

// ---------------------------------------------------------------------------
// out(cols_in x rows_in) = in(rows_in x cols_in)^T; both column-major views
// (equivalently: column-major <-> row-major conversion).  32x32 LDS tiles.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void transpose_kernel(const double *__restrict__ in,
                                                        long ldi,
                                                        double *__restrict__ out,
                                                        long ldo, int rows, int cols)
{
    __shared__ double t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;   // bx: rows, by: cols
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int q = ty; q < 32; q += 8) {
        int i = bx + tx, j = by + q;
        if (i < rows && j < cols) t[q][tx] = in[(long)j * ldi + i];
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        int j = by + tx, i = bx + q;
        if (i < rows && j < cols) out[(long)i * ldo + j] = t[tx][q];
    }
}

hipError_t transpose_f64(hipStream_t s, const double *in, long ldi, double *out,
                         long ldo, int rows_in, int cols_in)
{
    if (rows_in <= 0 || cols_in <= 0) return hipSuccess;
    ProfScope ps(tl_prof, s, PROF_TRANSPOSE, 0.0, 16.0 * rows_in * cols_in);
    dim3 grid((rows_in + 31) / 32, (cols_in + 31) / 32);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, in, ldi, out, ldo,
                       rows_in, cols_in);
    return hipGetLastError();
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void iota_kernel(int *p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

hipError_t fill_int(hipStream_t s, int *p, int n, int v)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

__global__ void piv_out_kernel(const int *piv0, int *ipiv, int n, int *info)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ipiv[i] = piv0[i] + 1;
    if (i == 0 && info && *info == INT_MAX) *info = 0;
}

hipError_t piv0_to_ipiv(hipStream_t s, const int *piv0, int *ipiv, int n, int *info)
{
    hipLaunchKernelGGL(piv_out_kernel, dim3((n + 255) / 256 > 0 ? (n + 255) / 256 : 1),
                       dim3(256), 0, s, piv0, ipiv, n, info);
    return hipGetLastError();
}

__global__ void zero_u64_kernel(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Host driver.
// ---------------------------------------------------------------------------
#define FLX_CHECK(x)                                                          \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) return e_;                                      \
    } while (0)

static hipError_t launch_panel(hipStream_t s, LuWork &wk, double *P, long ldw,
                               int mp, int w, int *piv, int piv_base, int *info,
                               int ib)
{
    int G = (mp + PANEL_T - 1) / PANEL_T;
    if (G > PANEL_GMAX) return hipErrorInvalidValue;
    wk.gen = (wk.gen + 1) & 0xffffffu;
    if (wk.gen == 0) {
        // sequence space wrapped: clear the slots so no stale tag can match
        hipLaunchKernelGGL(zero_u64_kernel, dim3((wk.exch_words + 255) / 256), dim3(256), 0, s,
                           wk.exch, wk.exch_words);
        wk.gen = 1;
    }
    // algorithmic panel traffic: the mp x w panel read and written once
    const double pf = (double)w * w * mp - (double)w * w * w / 3.0;
    ProfScope ps(wk.prof, s, PROF_PANEL, pf, 16.0 * mp * w);
    if (ib <= 32)
        hipLaunchKernelGGL(panel_kernel<32>, dim3(G), dim3(PANEL_T), 0, s, P, ldw, mp, w,
                           wk.exch, wk.gen, piv, piv_base, wk.list, info, piv_base,
                           wk.flags);
    else
        hipLaunchKernelGGL(panel_kernel<64>, dim3(G), dim3(PANEL_T), 0, s, P, ldw, mp, w,
                           wk.exch, wk.gen, piv, piv_base, wk.list, info, piv_base,
                           wk.flags);
    return hipGetLastError();
}

static hipError_t launch_trsm_rm(hipStream_t s, Profiler *prof, const double *L, long ldl,
                                 double *X, long ldx, int k, int ncols)
{
    if (ncols <= 0 || k <= 0) return hipSuccess;
    if (k > TR_KMAX) return hipErrorInvalidValue;
    ProfScope ps(prof, s, PROF_TRSM_RM, (double)k * k * ncols,
                 8.0 * (2.0 * k * ncols + 0.5 * k * k));
    hipLaunchKernelGGL(trsm_rm_llu_kernel, dim3((ncols + TR_CW - 1) / TR_CW), dim3(256), 0,
                       s, L, ldl, X, ldx, k, ncols);
    return hipGetLastError();
}

// W[r0.., c0..c0+nc) -= W[r0.., k0..k0+kk) * W[k0..k0+kk, c0..c0+nc)
static hipError_t rm_update(hipStream_t s, Profiler *prof, double *W, long ldw, int r0,
                            int nr, int c0, int nc, int k0, int kk)
{
    if (nr <= 0 || nc <= 0 || kk <= 0) return hipSuccess;
    ProfScope ps(prof, s, PROF_GEMM, 2.0 * nr * nc * kk,
                 8.0 * ((double)nc * kk + (double)kk * nr + 2.0 * nc * nr));
    return gemm_f64(s, 0, 0, nc, nr, kk, -1.0, W + (long)k0 * ldw + c0, ldw,
                    W + (long)r0 * ldw + k0, ldw, 1.0, W + (long)r0 * ldw + c0, ldw, 0);
}

hipError_t getrf_rm_f64(hipStream_t s, LuWork &wk, double *W, long ldw, int m, int n,
                        int *piv0, int *info, const LuOptions &opt)
{
    const int mn = m < n ? m : n;
    if (m <= 0 || n <= 0) return hipSuccess;
    int nb = opt.nb, ib = opt.ib;
    if (ib > PANEL_WMAX) ib = PANEL_WMAX;
    if (ib < 1) ib = 32;
    if (nb < ib) nb = ib;
    if (nb > TR_KMAX) nb = TR_KMAX;
    nb = (nb / ib) * ib;
    if (m > PANEL_GMAX * PANEL_T) return hipErrorInvalidValue;
    // the composed permutation, when requested, starts as the identity
    if (wk.perm) {
        hipLaunchKernelGGL(iota_kernel, dim3((m + 255) / 256), dim3(256), 0, s, wk.perm, m);
    }
    for (int J = 0; J < mn; J += nb) {
        const int jb = (mn - J < nb) ? mn - J : nb;
        for (int j = J; j < J + jb; j += ib) {
            const int w = (J + jb - j < ib) ? J + jb - j : ib;
            FLX_CHECK(launch_panel(s, wk, W + (long)j * ldw + j, ldw, m - j, w, piv0 + j, j,
                                   info, ib));
            // apply this sub-panel's interchanges to every other column
            const int nleft = (j + RP_CW - 1) / RP_CW;
            const int nright = (n - (j + w) + RP_CW - 1) / RP_CW;
            if (nleft + nright > 0) {
                ProfScope ps(wk.prof, s, PROF_ROWPERM, 0.0, 0.0);
                hipLaunchKernelGGL(rowperm_kernel, dim3(nleft + nright), dim3(256), 0, s, W,
                                   ldw, m, j, j + w, n, wk.list, wk.perm, wk.moved);
            } else if (wk.perm) {
                hipLaunchKernelGGL(perm_compose_kernel, dim3(1), dim3(256), 0, s, wk.list,
                                   wk.perm, m);
            }
            FLX_CHECK(hipGetLastError());
            // inner update of the rest of the outer block
            const int cr = J + jb - (j + w);
            if (cr > 0) {
                FLX_CHECK(launch_trsm_rm(s, wk.prof, W + (long)j * ldw + j, ldw,
                                         W + (long)j * ldw + j + w, ldw, w, cr));
                FLX_CHECK(rm_update(s, wk.prof, W, ldw, j + w, m - j - w, j + w, cr, j, w));
            }
        }
        const int nr = n - (J + jb);
        if (nr > 0) {
            FLX_CHECK(launch_trsm_rm(s, wk.prof, W + (long)J * ldw + J, ldw,
                                     W + (long)J * ldw + J + jb, ldw, jb, nr));
            FLX_CHECK(rm_update(s, wk.prof, W, ldw, J + jb, m - J - jb, J + jb, nr, J, jb));
        }
    }
    return hipSuccess;
}

} // namespace flx
