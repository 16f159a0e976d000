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
constexpr int PANEL_W = 32;             // panel width of the cooperative kernel
constexpr int PANEL_ROWS_MAX = LU_N_MAX;   // panel height limit
constexpr int PANEL_WMAX = 64;

// ---------------------------------------------------------------------------
// Row permutation of a row-major matrix outside one panel:
//   W[dst_e][c] = W_old[src_e][c]   for every list entry e,
// for columns c in [0, c1) U [c2, n).  Each workgroup owns a 64-column
// chunk: it reads every source segment into registers, then writes (no
// hazard between chunks).  This is laswp (flens/lapack/la/laswp.tcc:54-93) for
// both column ranges of trf.tcc:126 and :132 at once.
// ---------------------------------------------------------------------------
constexpr int RP_CW = 64;
constexpr int RP_MAXE = 2 * PANEL_WMAX;

// The interchange list's slices gathered into LDS with every load issued at
// once: 256 threads x SL slots cover PANEL_GROUPS_MAX slices of up to 64
// entries, slice offsets by a wave prefix sum.  Returns the entry count.
FLX_DEV int load_list_par(const int *__restrict__ list, int G, int m, int *s_dst, int *s_src,
                          int *s_cnt)
{
    const int t = threadIdx.x;
    constexpr int SL = PANEL_GROUPS_MAX * 64 / 256;   // slots per thread
    static_assert(PANEL_GROUPS_MAX * 64 == 256 * SL, "slot map");
    int d[SL], sr[SL];
#pragma unroll
    for (int q = 0; q < SL; ++q) {
        const int slot = t + 256 * q, g = slot >> 6, e = slot & 63;
        const int *sl = list + (size_t)(g < G ? g : 0) * PANEL_LIST_SLICE;
        d[q] = sl[1 + 2 * e];
        sr[q] = sl[2 + 2 * e];
    }
    static_assert(PANEL_GROUPS_MAX == 64, "one wave scans the slice counts");
    if (t < 64) {
        // counts, their exclusive prefix sums (slice offsets), the total
        const int c = (t < G) ? min(list[(size_t)t * PANEL_LIST_SLICE], 64) : 0;
        int inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(inc, o, 64);
            if (t >= o) inc += y;
        }
        s_cnt[t] = c;
        s_cnt[PANEL_GROUPS_MAX + 1 + t] = inc - c;
        if (t == 63) s_cnt[PANEL_GROUPS_MAX] = min(inc, RP_MAXE);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SL; ++q) {
        const int slot = t + 256 * q, g = slot >> 6, e = slot & 63;
        if (e < s_cnt[g]) {
            const int o = s_cnt[PANEL_GROUPS_MAX + 1 + g] + e;
            const bool ok = d[q] >= 0 && d[q] < m && sr[q] >= 0 && sr[q] < m;
            if (o < RP_MAXE) {
                s_dst[o] = ok ? d[q] : 0;   // an invalid entry degenerates to row 0 -> row 0
                s_src[o] = ok ? sr[q] : 0;
            }
        }
    }
    __syncthreads();
    return s_cnt[PANEL_GROUPS_MAX];
}

template <typename T>
__global__ __launch_bounds__(256) void rowperm_kernel(T *__restrict__ Wm, long ldw, int m,
                                                      int c1, int c2, int n,
                                                      const int *__restrict__ list, int G,
                                                      int *__restrict__ perm,
                                                      unsigned long long *__restrict__ moved)
{
    critical_path_priority();
    __shared__ int s_dst[RP_MAXE], s_src[RP_MAXE], s_cnt[2 * PANEL_GROUPS_MAX + 1];
    const int cnt = load_list_par(list, G, m, s_dst, s_src, s_cnt);
    if (cnt <= 0) return;
    const int nleft = (c1 + RP_CW - 1) / RP_CW;
    int chunk = blockIdx.x;
    int cbeg, cend;
    if (chunk < nleft) { cbeg = chunk * RP_CW; cend = min(c1, cbeg + RP_CW); }
    else { cbeg = c2 + (chunk - nleft) * RP_CW; cend = min(n, cbeg + RP_CW); }
    const int cw = cend - cbeg;
    if (moved && threadIdx.x == 0) atomicAdd(moved, (unsigned long long)cnt * cw);
    // every source segment is read before any destination is written:
    // all loads issued at once into registers, completed, a barrier, then
    // the stores
    constexpr int PER = RP_MAXE * RP_CW / 256;
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int q = threadIdx.x + 256 * i, e = q / RP_CW, c = q % RP_CW;
        v[i] = (e < cnt && c < cw) ? Wm[(long)s_src[e] * ldw + cbeg + c] : T(0);
    }
    __builtin_amdgcn_s_waitcnt(0xf70);   // vmcnt(0)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int q = threadIdx.x + 256 * i, e = q / RP_CW, c = q % RP_CW;
        if (e < cnt && c < cw) Wm[(long)s_dst[e] * ldw + cbeg + c] = v[i];
    }
    // block 0 also composes the global row permutation
    if (blockIdx.x == 0 && perm) {
        __syncthreads();
        int v = 0;
        if (threadIdx.x < cnt) v = perm[s_src[threadIdx.x]];
        __syncthreads();
        if (threadIdx.x < cnt) perm[s_dst[threadIdx.x]] = v;
    }
}

// permutation-only variant, used when no column lies outside the panel
__global__ __launch_bounds__(256) void perm_compose_kernel(const int *__restrict__ list, int G,
                                                           int *__restrict__ perm, int m)
{
    __shared__ int s_dst[RP_MAXE], s_src[RP_MAXE], s_cnt[2 * PANEL_GROUPS_MAX + 1];
    const int cnt = load_list_par(list, G, m, s_dst, s_src, s_cnt);
    int v = 0;
    if (threadIdx.x < cnt) v = perm[s_src[threadIdx.x]];
    __syncthreads();
    if (threadIdx.x < cnt) perm[s_dst[threadIdx.x]] = v;
}

// ---------------------------------------------------------------------------
// Fused panel step of the blocked LU (lookahead path), after the
// cooperative panel kernel has factored and interchanged the panel's own
// columns j .. j+w of the outer block [J, J+jb):
//   lu_stage_kernel (one workgroup per stage row, one more): the
//     interchange list's source rows (block columns) into stage rows 0..,
//     the panel's top w rows as they will be after the interchange into
//     stage rows LU_STAGE_TOP.., the moved rows' perm entries and
//     destinations; the extra workgroup inverts the unit-lower L11 (when its
//     entries are finite) into the stage's LU_STAGE_INV part;
//   lu_apply_kernel (32 rows per workgroup): the interchange of the block's
//     other columns, U12 = L11^-1 (top rows' right part) -- laswp, the
//     blas::sm of trf.tcc:136-138 -- and A22 -= L21 U12 (trf.tcc:140-145)
//     for the block's columns right of the panel, in one launch.  U12 is the
//     product of the inverse's lower triangle with the top rows (every
//     workgroup, no dependent chain), or, when L11 holds a non-finite
//     entry, the substitution itself (the reference's NaN / Inf pattern).
// A workgroup writes only its own rows and reads other rows only from the
// stage (or the panel's columns, which it does not write), so workgroups
// need no ordering among themselves.
// ---------------------------------------------------------------------------
constexpr int LA_ROWS = 32;                 // rows per apply workgroup
constexpr int LA_TPR = 256 / LA_ROWS;       // threads per row
constexpr int LA_CR = LU_STAGE_LD - 32;     // widest right part (block 256, panel >= 32)
static_assert(LU_STAGE_TOP == RP_MAXE, "stage layout");

// One workgroup per stage row: blockIdx.x < LU_STAGE_TOP = list entry,
// above = top row blockIdx.x - LU_STAGE_TOP.  Workgroup 0 also writes the
// destinations (count at [LU_STAGE_TOP]) and perm entries.
__global__ __launch_bounds__(256) void lu_stage_kernel(const double *__restrict__ W, long ldw,
                                                       int m, int J, int jb, int j, int w,
                                                       const int *__restrict__ list, int G,
                                                       const int *__restrict__ perm,
                                                       double *__restrict__ stage,
                                                       int *__restrict__ sdst,
                                                       int *__restrict__ sperm,
                                                       unsigned long long *__restrict__ moved)
{
    critical_path_priority();
    __shared__ int s_dst[RP_MAXE], s_src[RP_MAXE], s_cnt[2 * PANEL_GROUPS_MAX + 1];
    __shared__ int s_row;
    const int t = threadIdx.x, b = blockIdx.x;
    if (b == LU_STAGE_ROWS) {
        // inv(L11) of the w x w unit-lower block at W[j.., j..] (rows and
        // columns beyond w: identity), row-major at stage + LU_STAGE_INV,
        // then a flag: 1 when every strict-lower entry is finite
        __shared__ double s_l[32][33];
        __shared__ int s_bad;
        if (t == 0) s_bad = 0;
        __syncthreads();
        double lv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q, r = e >> 5, c = e & 31;
            lv[q] = (r < w && c < r) ? W[(long)(j + r) * ldw + j + c] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            if (!__builtin_isfinite(lv[q])) s_bad = 1;
            s_l[e >> 5][e & 31] = lv[q];
        }
        __syncthreads();
        double *inv = stage + LU_STAGE_INV;
        if (t < 32) {
            // column t of the inverse: x = e_t, x_i -= L(i, k) x_k (k < i)
            double x[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) x[i] = i == t ? 1.0 : 0.0;
#pragma unroll
            for (int k = 0; k < 31; ++k)
#pragma unroll
                for (int i = k + 1; i < 32; ++i) x[i] = fma(-s_l[i][k], x[k], x[i]);
#pragma unroll
            for (int i = 0; i < 32; ++i) inv[i * 32 + t] = x[i];
        }
        if (t == 0) inv[32 * 32] = s_bad ? 0.0 : 1.0;
        return;
    }
    const int cnt = load_list_par(list, G, m, s_dst, s_src, s_cnt);
    if (b == 0) {
        if (t < cnt) {
            sdst[t] = s_dst[t];
            sperm[t] = perm ? perm[s_src[t]] : 0;
            sperm[LU_STAGE_TOP + t] = s_src[t];   // the moved row's original position
        }
        if (t == 0) sdst[LU_STAGE_TOP] = cnt;
        if (t == 0 && moved) atomicAdd(moved, (unsigned long long)cnt * (jb - w));
    }
    int src = -1;
    if (b < LU_STAGE_TOP) {
        if (b < cnt) src = s_src[b];
    } else {
        const int k = b - LU_STAGE_TOP;
        if (k < w) {
            // top row j + k: the source of the entry moving into it, else itself
            if (t == 0) s_row = j + k;
            __syncthreads();
            if (t < cnt && s_dst[t] == j + k) s_row = s_src[t];
            __syncthreads();
            src = s_row;
        }
    }
    if (src < 0 || t >= jb) return;
    stage[(long)b * LU_STAGE_LD + t] = W[(long)src * ldw + J + t];
}

// CR: the widest right part the instance handles (the LDS arrays scale with
// it, so at nb <= 128 a workgroup fits beside a trailing-update GEMM's on
// one compute unit)
template <int CR>
__global__ __launch_bounds__(256) void lu_apply_kernel(double *__restrict__ W, long ldw, int m,
                                                       int J, int jb, int j, int w,
                                                       const double *__restrict__ stage,
                                                       const int *__restrict__ sdst,
                                                       const int *__restrict__ sperm,
                                                       int *__restrict__ perm,
                                                       const double *__restrict__ lsc,
                                                       int defer_l)
{
    static_assert(LA_ROWS * CR % 256 == 0 && CR % LA_TPR == 0 && CR <= CR, "apply tile");
    constexpr int NQ = LA_ROWS * CR / 256;   // right-part elements per thread
    critical_path_priority();
    __shared__ double s_l[32][33];               // inv(L11) or L11 shifted (s_l[k][i] = L(k+1+i, k))
    __shared__ double s12[32][CR + 1];        // U12
    __shared__ double s_a[LA_ROWS][CR + 1];   // the tile's right parts
    __shared__ double sx[LA_ROWS][33];           // the tile's L21 rows
    __shared__ int s_slot[LA_ROWS];
    __shared__ int s_rsrc[LA_ROWS];              // the row's original position (moved rows)
    const int t = threadIdx.x;
    const int jo = j - J;
    const int cr = jb - jo - w;                  // columns right of the panel
    const int i0 = j + blockIdx.x * LA_ROWS;     // first row of the tile
    // Every load that does not depend on the interchange list is issued at
    // once: the list's count and entries, the inverse's flag, L11 (shifted:
    // s_l[k][i] = L(k+1+i, k)) and its inverse (row-major), the tile's L21
    // rows (panel columns, final) and this thread's column of the top rows.
    const int cnt = sdst[LU_STAGE_TOP];
    const int sd = sdst[t < LU_STAGE_TOP ? t : 0];
    const int ss = sperm[LU_STAGE_TOP + (t < LU_STAGE_TOP ? t : 0)];
    const bool useinv = stage[LU_STAGE_INV + 32 * 32] != 0.0;
    double lv[4], iv[4], lx[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = t + 256 * q, k = e >> 5, i = e & 31, r = k + 1 + i;
        lv[q] = (r < w) ? W[(long)(j + (r < 32 ? r : 0)) * ldw + j + k] : 0.0;
        iv[q] = stage[LU_STAGE_INV + e];
        const int rr = e >> 5, kk = e & 31, ii = i0 + rr;
        lx[q] = (!defer_l && ii < m && kk < w && ii - j >= w) ? W[(long)ii * ldw + j + kk] : 0.0;
    }
    double x[32];
#pragma unroll
    for (int k = 0; k < 32; ++k)
        x[k] = (t < cr && k < w) ? stage[(long)(LU_STAGE_TOP + k) * LU_STAGE_LD + jo + w + t]
                                 : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int e = t + 256 * q;
        s_l[e >> 5][e & 31] = useinv ? iv[q] : lv[q];
        sx[e >> 5][e & 31] = lx[q];
    }
    if (t < LA_ROWS) s_slot[t] = -1;
    __syncthreads();
    if (t < cnt) {
        const int d = sd - i0;
        if (d >= 0 && d < LA_ROWS) {
            s_slot[d] = t;
            s_rsrc[d] = ss;
        }
    }
    __syncthreads();
    if (defer_l) {
        // the panel left the tile's multiplier rows in lsc (column-major by
        // original row, mp = m - j rows): row i's are those of the row that
        // moved into it, or its own
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q, rr = e >> 5, kk = e & 31, ii = i0 + rr;
            const int src = s_slot[rr] >= 0 ? s_rsrc[rr] : ii;
            lx[q] = (ii < m && kk < w && ii - j >= w) ? lsc[(long)kk * (m - j) + (src - j)] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            sx[e >> 5][e & 31] = lx[q];
        }
    }
    // the tile's right parts (moved rows from the stage), every load issued
    // before any is used
    {
        double av[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + 256 * q, r = e / CR, c = e % CR, i = i0 + r;
            const int sl = s_slot[r];
            const double *src = (sl >= 0) ? stage + (long)sl * LU_STAGE_LD + jo + w
                                          : W + (long)(i < m ? i : j) * ldw + j + w;
            av[q] = (c < cr && i < m) ? src[c] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int e = t + 256 * q;
            s_a[e / CR][e % CR] = av[q];
        }
    }
    if (useinv && cr > 0 && t < cr) {
        // U12 = inv(L11) A12 over the inverse's lower triangle only (row r
        // takes the top rows k <= r, as the substitution does), every entry
        // independent, one column per thread in registers
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            double v = x[r];
#pragma unroll
            for (int k = 0; k < r; ++k) v = fma(s_l[r][k], x[k], v);
            s12[r][t] = v;
        }
    }
    // else U12 by forward substitution with the unit-lower L11, one column
    // per thread; x rotates (register 0 is the finished entry of the step),
    // so the rolled loop indexes registers by constants only
    if (!useinv && cr > 0 && t < cr) {
#pragma unroll 1
        for (int k = 0; k < 32; ++k) {
            const double x0 = x[0];
            // only the real rows below k take the update: the finished
            // (rotated) entries and the rows beyond the panel just move, so
            // an Inf / NaN in x0 reaches exactly the rows the reference's
            // substitution reaches (cxxblas trsv: x_i -= sum_{q<i} L(i,q) x_q)
#pragma unroll
            for (int i = 0; i < 31; ++i) {
                const double t = fma(-s_l[k][i], x0, x[i + 1]);
                x[i] = (k + 1 + i < w) ? t : x[i + 1];
            }
            x[31] = x0;
        }
#pragma unroll
        for (int k = 0; k < 32; ++k) s12[k][t] = x[k];
    }
    __syncthreads();
    const int r = t / LA_TPR, part = t % LA_TPR;
    const int i = i0 + r;
    if (i >= m) return;
    const int slot = s_slot[r];
    double *row = W + (long)i * ldw + J;
    if (slot >= 0) {
        // the left part (a multiple of 32 columns) of a row that moves in
        const double *src = stage + (long)slot * LU_STAGE_LD;
        constexpr int NL = CR / LA_TPR;
        double lv[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int c = part + LA_TPR * q;
            lv[q] = src[c < jo ? c : 0];
        }
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int c = part + LA_TPR * q;
            if (c < jo) row[c] = lv[q];
        }
        if (part == 0 && perm) perm[i] = sperm[slot];
    }
    const int p = i - j;
    if (defer_l && p >= w) {
        // the row's multipliers, which the panel left in lsc
#pragma unroll
        for (int q = 0; q < 32 / LA_TPR; ++q) {
            const int c = part + LA_TPR * q;
            if (c < w) row[jo + c] = sx[r][c];
        }
    }
    if (cr <= 0) return;
    double *rr = row + jo + w;
    if (p < w) {
        for (int c = part; c < cr; c += LA_TPR) rr[c] = s12[p][c];
        return;
    }
    double xr[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) xr[k] = sx[r][k];
#pragma unroll 1
    for (int c = part; c < cr; c += LA_TPR) {
        double v = s_a[r][c];
#pragma unroll
        for (int k = 0; k < 32; ++k) v = fma(-xr[k], s12[k][c], v);
        rr[c] = v;
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
constexpr int TR_BS = 16;
constexpr int TR_FOLD = 4;   // rows folded together per thread

template <typename T>
__global__ __launch_bounds__(256) void trsm_rm_llu_kernel(const T *__restrict__ L,
                                                          long ldl, T *__restrict__ X,
                                                          long ldx, int k, int ncols)
{
    __shared__ T xs[TR_KMAX][TR_CW + 1];
    __shared__ T ls[TR_KMAX][TR_BS + 1];   // L(kb.., kb..kb+15) strip
    const int tid = threadIdx.x;
    const int tc = tid % TR_CW, tg = tid / TR_CW;   // 32 columns x 8 row groups
    // workgroup b solves the column chunks b, b + grid, ...
    for (int c0 = blockIdx.x * TR_CW; c0 < ncols; c0 += gridDim.x * TR_CW) {
        const int cw = min(TR_CW, ncols - c0);
        if (c0 != (int)blockIdx.x * TR_CW) __syncthreads();
        for (int q = tid; q < k * TR_CW; q += 256) {
            int rr = q / TR_CW, c = q % TR_CW;
            xs[rr][c] = (c < cw) ? X[(long)rr * ldx + c0 + c] : T(0);
        }
        for (int kb = 0; kb < k; kb += TR_BS) {
            const int bs = min(TR_BS, k - kb);
            // stage the strip of L below (and including) the diagonal block
            for (int q = tid; q < (k - kb) * TR_BS; q += 256) {
                int rr = q / TR_BS, cc = q % TR_BS;
                ls[rr][cc] = (cc < bs) ? L[(long)(kb + rr) * ldl + kb + cc] : T(0);
            }
            __syncthreads();
            // diagonal block: column tc held in registers by threads tg == 0
            if (tg == 0) {
                T xb[TR_BS];
#pragma unroll
                for (int i = 0; i < TR_BS; ++i) xb[i] = (i < bs) ? xs[kb + i][tc] : T(0);
#pragma unroll
                for (int i = 1; i < TR_BS; ++i) {
                    T v = xb[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) v = fma(-ls[i][q], xb[q], v);
                    xb[i] = v;
                }
#pragma unroll
                for (int i = 0; i < TR_BS; ++i)
                    if (i < bs) xs[kb + i][tc] = xb[i];
            }
            __syncthreads();
            // fold the solved block into the rows below, four rows (four
            // independent chains, each in the order above) per pass
            T xb[TR_BS];
#pragma unroll
            for (int q = 0; q < TR_BS; ++q) xb[q] = xs[kb + q < k ? kb + q : kb][tc];
            for (int r0 = kb + bs + tg; r0 < k; r0 += 8 * TR_FOLD) {
                T v[TR_FOLD];
                int rc[TR_FOLD];
#pragma unroll
                for (int u = 0; u < TR_FOLD; ++u) {
                    rc[u] = min(r0 + 8 * u, k - 1);
                    v[u] = xs[rc[u]][tc];
                }
#pragma unroll
                for (int q = 0; q < TR_BS; ++q)
#pragma unroll
                    for (int u = 0; u < TR_FOLD; ++u) v[u] = fma(-ls[rc[u] - kb][q], xb[q], v[u]);
#pragma unroll
                for (int u = 0; u < TR_FOLD; ++u)
                    if (r0 + 8 * u < k) xs[r0 + 8 * u][tc] = v[u];
            }
            __syncthreads();
        }
        for (int q = tid; q < k * TR_CW; q += 256) {
            int rr = q / TR_CW, c = q % TR_CW;
            if (c < cw) X[(long)rr * ldx + c0 + c] = xs[rr][c];
        }
    }
}

// For k <= TR_FULLK the lower part of L is staged in LDS once, beside the
// chunk's X, so the dependent steps of the substitution wait on LDS only.
// Layout: 16-row blocks R, block R holding its rows' first 16 (R + 1)
// columns (through its diagonal block) from offset 256 R (R + 1) / 2, with
// zeros on and above the diagonal and outside k x k -- every read below is
// in range and branch-free, and a zero stands where trsm_rm_llu_kernel's
// strip holds one, so the arithmetic is that kernel's, operation for
// operation.
constexpr int TR_FULLK = 128;     // FP64 (KMAX = 256 fits LDS for FP32 at CW = 8)

// CW columns per workgroup, 256 / CW row groups folding the rows below
template <typename T, int CW, int KMAX>
__global__ __launch_bounds__(256) void trsm_rm_llu_full_kernel(const T *__restrict__ L,
                                                               long ldl, T *__restrict__ X,
                                                               long ldx, int k, int ncols)
{
    constexpr int NG = 256 / CW;
    constexpr int FB = KMAX / TR_BS;                              // 16-row blocks
    constexpr int LPAD = TR_BS * TR_BS * FB * (FB + 1) / 2;
    static_assert(256 % CW == 0 && KMAX % NG == 0 && KMAX % TR_BS == 0, "column chunk");
    __shared__ T xs[KMAX][CW + 1];
    __shared__ T lp[LPAD];
    const int tid = threadIdx.x;
    const int tc = tid % CW, tg = tid / CW;   // CW columns x NG row groups
    // block R: 256 (R + 1) entries, R + 1 per thread; every block's loads
    // are issued before any is stored (one memory latency in all)
    {
        T lv[FB * (FB + 1) / 2];
        StaticFor<0, FB>::run([&](auto RC) {
            constexpr int R = decltype(RC)::value;
            constexpr int W = TR_BS * (R + 1), V0 = R * (R + 1) / 2;
#pragma unroll
            for (int u = 0; u <= R; ++u) {
                const int e = tid + 256 * u, r = R * TR_BS + e / W, c = e % W;
                lv[V0 + u] = (r < k && c < r) ? L[(long)r * ldl + c] : T(0);
            }
        });
        StaticFor<0, FB>::run([&](auto RC) {
            constexpr int R = decltype(RC)::value;
            constexpr int BASE = TR_BS * TR_BS * R * (R + 1) / 2, V0 = R * (R + 1) / 2;
#pragma unroll
            for (int u = 0; u <= R; ++u) lp[BASE + tid + 256 * u] = lv[V0 + u];
        });
    }
    // workgroup b solves the column chunks b, b + grid, ...
    for (int c0 = blockIdx.x * CW; c0 < ncols; c0 += gridDim.x * CW) {
        const int cw = min(CW, ncols - c0);
        if (c0 != (int)blockIdx.x * CW) __syncthreads();
        // rows tg, tg + NG, ... of the chunk, all loads issued first
        {
            T v[KMAX / NG];
#pragma unroll
            for (int u = 0; u < KMAX / NG; ++u) {
                const int rr = tg + NG * u;
                v[u] = (rr < k && tc < cw) ? X[(long)rr * ldx + c0 + tc] : T(0);
            }
#pragma unroll
            for (int u = 0; u < KMAX / NG; ++u) xs[tg + NG * u][tc] = v[u];
        }
        __syncthreads();
        for (int kb = 0; kb < k; kb += TR_BS) {
            const int bs = min(TR_BS, k - kb);
            const int R = kb / TR_BS;
            const T *lb = lp + TR_BS * TR_BS * R * (R + 1) / 2 + kb;   // block R, column kb
            const int lw = TR_BS * (R + 1);
            // diagonal block: column tc held in registers by threads tg == 0
            if (tg == 0) {
                T xb[TR_BS];
#pragma unroll
                for (int i = 0; i < TR_BS; ++i) {
                    const T t = xs[kb + i][tc];
                    xb[i] = (i < bs) ? t : T(0);
                }
#pragma unroll
                for (int i = 1; i < TR_BS; ++i) {
                    // row i's coefficients, read as whole 16-entry rows
                    T co[TR_BS];
#pragma unroll
                    for (int q = 0; q < TR_BS; ++q) co[q] = lb[i * lw + q];
                    T v = xb[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) v = fma(-co[q], xb[q], v);
                    xb[i] = v;
                }
#pragma unroll
                for (int i = 0; i < TR_BS; ++i) xs[kb + i][tc] = xb[i];
            }
            __syncthreads();
            // fold the solved block into the rows below, four rows (four
            // independent chains, each in the order above) per pass
            T xb[TR_BS];
#pragma unroll
            for (int q = 0; q < TR_BS; ++q) xb[q] = xs[kb + q < k ? kb + q : kb][tc];
            for (int r0 = kb + bs + tg; r0 < k; r0 += NG * TR_FOLD) {
                T v[TR_FOLD];
                const T *lr[TR_FOLD];
#pragma unroll
                for (int u = 0; u < TR_FOLD; ++u) {
                    const int rr = min(r0 + NG * u, k - 1), Rr = rr / TR_BS;
                    lr[u] = lp + TR_BS * TR_BS * Rr * (Rr + 1) / 2 + (rr % TR_BS) * TR_BS * (Rr + 1) +
                            kb;
                    v[u] = xs[rr][tc];
                }
#pragma unroll
                for (int q = 0; q < TR_BS; ++q)
#pragma unroll
                    for (int u = 0; u < TR_FOLD; ++u) v[u] = fma(-lr[u][q], xb[q], v[u]);
#pragma unroll
                for (int u = 0; u < TR_FOLD; ++u)
                    if (r0 + NG * u < k) xs[r0 + NG * u][tc] = v[u];
            }
            __syncthreads();
        }
        for (int q = tid; q < k * CW; q += 256) {
            int rr = q / CW, c = q % CW;
            if (c < cw) X[(long)rr * ldx + c0 + c] = xs[rr][c];
        }
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


// ---------------------------------------------------------------------------
// Host driver.
// ---------------------------------------------------------------------------

template <typename T>
hipError_t trsm_rm_llu(hipStream_t s, Profiler *prof, const T *L, long ldl, T *X, long ldx,
                       int k, int ncols, int max_groups)
{
    if (ncols <= 0 || k <= 0) return hipSuccess;
    if (k > TR_KMAX) return hipErrorInvalidValue;
    ProfScope ps(prof, s, PROF_TRSM_RM, (double)k * k * ncols,
                 sizeof(T) * (2.0 * k * ncols + 0.5 * k * k));
    int groups = (ncols + TR_CW - 1) / TR_CW;
    if (max_groups > 0 && groups > max_groups) groups = max_groups;
    const int g8 = (ncols + 7) / 8;
    if (k <= TR_FULLK && ncols <= 4 * TR_CW) {
        // a narrow solve (the next block's columns): 8-column chunks, so
        // more workgroups share the rows below each diagonal block
        hipLaunchKernelGGL((trsm_rm_llu_full_kernel<T, 8, TR_FULLK>), dim3(g8), dim3(256), 0, s,
                           L, ldl, X, ldx, k, ncols);
        return hipGetLastError();
    }
    if (k <= TR_FULLK) {
        hipLaunchKernelGGL((trsm_rm_llu_full_kernel<T, TR_CW, TR_FULLK>), dim3(groups), dim3(256),
                           0, s, L, ldl, X, ldx, k, ncols);
        return hipGetLastError();
    }
    if constexpr (sizeof(T) == 4) {
        if (ncols <= 8 * TR_CW) {
            // FP32 at nb = 256 (the mixed factor's next block): L's 256 rows
            // fit LDS beside 8-column chunks
            hipLaunchKernelGGL((trsm_rm_llu_full_kernel<T, 8, 2 * TR_FULLK>), dim3(g8),
                               dim3(256), 0, s, L, ldl, X, ldx, k, ncols);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL(trsm_rm_llu_kernel<T>, dim3(groups), dim3(256), 0, s, L, ldl, X, ldx, k,
                       ncols);
    return hipGetLastError();
}
template hipError_t trsm_rm_llu<double>(hipStream_t, Profiler *, const double *, long, double *,
                                        long, int, int, int);
template hipError_t trsm_rm_llu<float>(hipStream_t, Profiler *, const float *, long, float *,
                                       long, int, int, int);

template <typename T>
hipError_t rowperm_rm(hipStream_t s, Profiler *prof, T *W, long ldw, int m, int c1, int c2,
                      int n, const int *list, int G, int *perm, unsigned long long *moved)
{
    const int nleft = (c1 + RP_CW - 1) / RP_CW;
    const int nright = (n - c2 + RP_CW - 1) / RP_CW;
    if (nleft + nright > 0) {
        ProfScope ps(prof, s, PROF_ROWPERM, 0.0, 0.0);
        hipLaunchKernelGGL(rowperm_kernel<T>, dim3(nleft + nright), dim3(256), 0, s, W, ldw, m,
                           c1, c2, n, list, G, perm, moved);
    } else if (perm) {
        hipLaunchKernelGGL(perm_compose_kernel, dim3(1), dim3(256), 0, s, list, G, perm, m);
    }
    return hipGetLastError();
}
template hipError_t rowperm_rm<double>(hipStream_t, Profiler *, double *, long, int, int, int,
                                       int, const int *, int, int *, unsigned long long *);
template hipError_t rowperm_rm<float>(hipStream_t, Profiler *, float *, long, int, int, int, int,
                                      const int *, int, int *, unsigned long long *);

hipError_t iota(hipStream_t s, int *p, int n)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

// W[r0.., c0..c0+nc) -= W[r0.., k0..k0+kk) * W[k0..k0+kk, c0..c0+nc)
static hipError_t rm_update(hipStream_t s, Profiler *prof, double *W, long ldw, int r0,
                            int nr, int c0, int nc, int k0, int kk, int max_groups = 0)
{
    if (nr <= 0 || nc <= 0 || kk <= 0) return hipSuccess;
    ProfScope ps(prof, s, PROF_GEMM, 2.0 * nr * nc * kk,
                 8.0 * ((double)nc * kk + (double)kk * nr + 2.0 * nc * nr));
    return gemm_f64(s, 0, 0, nc, nr, kk, -1.0, W + (long)k0 * ldw + c0, ldw,
                    W + (long)r0 * ldw + k0, ldw, 1.0, W + (long)r0 * ldw + c0, ldw, 0,
                    max_groups);
}

// ---------------------------------------------------------------------------
// Deferred row interchanges of one outer block: the block's sub-panels move
// rows only inside the block's columns; afterwards the composed permutation
// moves the same rows in every other column at once.  inv is the inverse of
// perm before the block, so a row now at position p came from inv[perm[p]].
// ---------------------------------------------------------------------------
__global__ void inv_perm_kernel(const int *__restrict__ perm, int *__restrict__ inv, int m,
                                int *__restrict__ reset)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && reset) *reset = 0;   // the block's move count, before its moves are listed
    if (p < m) {
        const int o = perm[p];
        if (o >= 0 && o < m) inv[o] = p;
    }
}

__global__ void block_moves_kernel(const int *__restrict__ perm, const int *__restrict__ inv,
                                   int J, int m, int *__restrict__ mv)
{
    const int p = J + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const int o = perm[p];
    if (o < 0 || o >= m) return;
    const int src = inv[o];
    if (src != p && src >= 0 && src < m) {
        const int e = atomicAdd(&mv[0], 1);
        if (e < LU_BLOCK_MOVES) {
            mv[1 + 2 * e] = p;
            mv[2 + 2 * e] = src;
        }
    }
}

static_assert(2 * TR_KMAX <= LU_BLOCK_MOVES, "an outer block moves at most 2*nb rows");
constexpr int AM_CW = 16;   // columns per workgroup

// W[dst_e][c] = W_old[src_e][c] for the columns [0, cs) and [ce, n)
template <typename T>
__global__ __launch_bounds__(256) void apply_moves_kernel(T *__restrict__ Wm, long ldw, int m,
                                                          int cs, int ce, int n,
                                                          const int *__restrict__ mv,
                                                          unsigned long long *__restrict__ moved)
{
    critical_path_priority();
    __shared__ int s_dst[LU_BLOCK_MOVES], s_src[LU_BLOCK_MOVES];
    const int cnt = min(mv[0], LU_BLOCK_MOVES);
    if (cnt <= 0) return;
    // chunks cover [0, cs) then [ce, n)
    const int nleft = (cs + AM_CW - 1) / AM_CW;
    const int chunk = blockIdx.x;
    int c0, c1;
    if (chunk < nleft) { c0 = chunk * AM_CW; c1 = min(cs, c0 + AM_CW); }
    else { c0 = ce + (chunk - nleft) * AM_CW; c1 = min(n, c0 + AM_CW); }
    const int cw = c1 - c0;
    if (cw <= 0) return;
    {
        int d[LU_BLOCK_MOVES / 256], sr[LU_BLOCK_MOVES / 256];
#pragma unroll
        for (int i = 0; i < LU_BLOCK_MOVES / 256; ++i) {
            const int e = threadIdx.x + 256 * i;
            d[i] = e < cnt ? mv[1 + 2 * e] : 0;
            sr[i] = e < cnt ? mv[2 + 2 * e] : 0;
        }
#pragma unroll
        for (int i = 0; i < LU_BLOCK_MOVES / 256; ++i) {
            const int e = threadIdx.x + 256 * i;
            const bool ok = d[i] >= 0 && d[i] < m && sr[i] >= 0 && sr[i] < m;
            if (e < cnt) {
                s_dst[e] = ok ? d[i] : 0;
                s_src[e] = ok ? sr[i] : 0;
            }
        }
    }
    __syncthreads();
    if (moved && threadIdx.x == 0) atomicAdd(moved, (unsigned long long)cnt * cw);
    // every source element of the chunk is read before any destination is
    // written (a destination row can be another move's source): all loads
    // issued at once into registers, completed, a barrier, then the stores
    constexpr int PER = LU_BLOCK_MOVES * AM_CW / 256;
    T v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int q = threadIdx.x + 256 * i, e = q / AM_CW, c = q % AM_CW;
        v[i] = (e < cnt && c < cw) ? Wm[(long)s_src[e] * ldw + c0 + c] : T(0);
    }
    __builtin_amdgcn_s_waitcnt(0xf70);   // vmcnt(0)
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int q = threadIdx.x + 256 * i, e = q / AM_CW, c = q % AM_CW;
        if (e < cnt && c < cw) Wm[(long)s_dst[e] * ldw + c0 + c] = v[i];
    }
}

hipError_t inv_perm(hipStream_t s, const int *perm, int *inv, int m, int *reset)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(inv_perm_kernel, dim3((m + 255) / 256), dim3(256), 0, s, perm, inv, m,
                       reset);
    return hipGetLastError();
}

// the block's move list (rows p >= J whose content moved), mv[0] cleared by
// the block's inv_perm(..., mv)
static hipError_t list_moves(hipStream_t s, const int *perm, const int *inv, int J, int m, int *mv)
{
    const int rows = m - J;
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(block_moves_kernel, dim3((rows + 255) / 256), dim3(256), 0, s, perm, inv,
                       J, m, mv);
    return hipGetLastError();
}

// the listed moves applied to the columns [0, c1) and [c2, c3)
template <typename T>
static hipError_t apply_moves(hipStream_t s, Profiler *prof, T *W, long ldw, int m, int c1, int c2,
                              int c3, const int *mv, unsigned long long *moved)
{
    const int nleft = (c1 + AM_CW - 1) / AM_CW;
    const int nright = c3 > c2 ? (c3 - c2 + AM_CW - 1) / AM_CW : 0;
    if (nleft + nright <= 0) return hipSuccess;
    ProfScope ps(prof, s, PROF_ROWPERM, 0.0, 0.0);
    hipLaunchKernelGGL(apply_moves_kernel<T>, dim3(nleft + nright), dim3(256), 0, s, W, ldw, m, c1,
                       c2, c3, mv, moved);
    return hipGetLastError();
}

template <typename T>
hipError_t defer_moves(hipStream_t s, Profiler *prof, T *W, long ldw, int m, int n, int J,
                       int jb, const int *perm, const int *inv, int *mv,
                       unsigned long long *moved)
{
    if (m - J <= 0) return hipSuccess;
    FLX_CHECK(list_moves(s, perm, inv, J, m, mv));
    return apply_moves<T>(s, prof, W, ldw, m, J, J + jb, n, mv, moved);
}
template hipError_t defer_moves<double>(hipStream_t, Profiler *, double *, long, int, int, int,
                                        int, const int *, const int *, int *,
                                        unsigned long long *);
template hipError_t defer_moves<float>(hipStream_t, Profiler *, float *, long, int, int, int, int,
                                       const int *, const int *, int *, unsigned long long *);

int lu_block_width(const LuOptions &opt)
{
    int nb = opt.nb, ib = opt.ib;
    if (ib > PANEL_W) ib = PANEL_W;
    if (ib < 1) ib = PANEL_W;
    if (nb < ib) nb = ib;
    if (nb > TR_KMAX) nb = TR_KMAX;
    return (nb / ib) * ib;
}

hipError_t getrf_rm_f64(hipStream_t s, LuWork &wk, double *W, long ldw, int m, int n,
                        int *piv0, int *info, const LuOptions &opt)
{
    const int mn = m < n ? m : n;
    const bool look = opt.lookahead && wk.perm && wk.inv && wk.mv && wk.side && wk.ev_a &&
                      wk.ev_r;
    bool rest_pending = false, next_pending = false;
    if (wk.side_pending) {
        // side-stream work the caller queued (ev_r): the first block end
        // waits for it, or at once without the lookahead (and on every
        // early return)
        wk.side_pending = false;
        const bool run = m > 0 && n > 0 && m <= PANEL_ROWS_MAX;
        if (look && run) rest_pending = true;
        else FLX_CHECK(hipStreamWaitEvent(s, wk.ev_r, 0));
    }
    if (m <= 0 || n <= 0) return hipSuccess;
    int ib = opt.ib;
    if (ib > PANEL_W) ib = PANEL_W;
    if (ib < 1) ib = PANEL_W;
    const int nb = lu_block_width(opt);
    if (m > PANEL_ROWS_MAX) return hipErrorInvalidValue;
    // the composed permutation, when requested, starts as the identity
    if (wk.perm) FLX_CHECK(iota(s, wk.perm, m));
    for (int J = 0; J < mn; J += nb) {
        const int jb = (mn - J < nb) ? mn - J : nb;
        // the block's move list, double-buffered by block parity: the side
        // stream may still read the previous block's
        int *mvb = wk.mv + (size_t)((J / nb) & 1) * (1 + 2 * LU_BLOCK_MOVES);
        if (look) FLX_CHECK(inv_perm(s, wk.perm, wk.inv, m, mvb));
        for (int j = J; j < J + jb; j += ib) {
            const int w = (J + jb - j < ib) ? J + jb - j : ib;
            // with the fused step, the panel writes only its pivot rows and
            // the step places the other rows' multipliers
            const bool fused = look && wk.stage && jb <= LU_STAGE_LD && w == PANEL_W;
            FLX_CHECK(launch_panel(s, wk, W + (long)j * ldw + j, ldw, m - j, w, piv0 + j, j,
                                   info, ib, fused ? 1 : 0));
            // the block's other columns may still be taking the previous
            // block's update on the side stream
            if (next_pending) {
                FLX_CHECK(hipStreamWaitEvent(s, wk.ev_n, 0));
                next_pending = false;
            }
            if (fused) {
                // the rest of the panel step in two launches: the stage, then
                // interchange + U12 + in-block update together
                {
                    ProfScope ps(wk.prof, s, PROF_ROWPERM, 0.0, 16.0 * 64 * jb);
                    hipLaunchKernelGGL(lu_stage_kernel, dim3(LU_STAGE_ROWS + 1), dim3(256), 0, s,
                                       (const double *)W, ldw, m, J, jb, j, w,
                                       (const int *)wk.list, panel_groups(m - j),
                                       (const int *)wk.perm, wk.stage, wk.sdst, wk.sperm,
                                       wk.moved);
                }
                const int cr = J + jb - (j + w);
                ProfScope ps(wk.prof, s, PROF_TRSM_RM, 2.0 * (m - j - w) * (double)w * cr,
                             16.0 * (m - j) * (double)(cr > 0 ? cr : 0));
                const dim3 ag((m - j + LA_ROWS - 1) / LA_ROWS);
                if (jb - PANEL_W <= 32)
                    hipLaunchKernelGGL(lu_apply_kernel<32>, ag, dim3(256), 0, s, W, ldw, m, J, jb,
                                       j, w, (const double *)wk.stage, (const int *)wk.sdst,
                                       (const int *)wk.sperm, wk.perm, (const double *)wk.lsc, 1);
                else if (jb - PANEL_W <= 96)
                    hipLaunchKernelGGL(lu_apply_kernel<96>, ag, dim3(256), 0, s, W, ldw, m, J, jb,
                                       j, w, (const double *)wk.stage, (const int *)wk.sdst,
                                       (const int *)wk.sperm, wk.perm, (const double *)wk.lsc, 1);
                else
                    hipLaunchKernelGGL(lu_apply_kernel<LA_CR>, ag, dim3(256), 0, s, W, ldw, m, J,
                                       jb, j, w, (const double *)wk.stage, (const int *)wk.sdst,
                                       (const int *)wk.sperm, wk.perm, (const double *)wk.lsc, 1);
                FLX_CHECK(hipGetLastError());
                continue;
            }
            if (look) {
                // interchanges inside the block's columns now, the rest deferred
                FLX_CHECK(rowperm_rm(s, wk.prof, W + J, ldw, m, j - J, j + w - J, jb, wk.list,
                                     panel_groups(m - j), wk.perm, wk.moved));
            } else {
                FLX_CHECK(rowperm_rm(s, wk.prof, W, ldw, m, j, j + w, n, wk.list,
                                     panel_groups(m - j), wk.perm, wk.moved));
            }
            // inner update of the rest of the outer block
            const int cr = J + jb - (j + w);
            if (cr > 0) {
                FLX_CHECK(trsm_rm_llu(s, wk.prof, (const double *)W + (long)j * ldw + j, ldw,
                                      W + (long)j * ldw + j + w, ldw, w, cr));
                FLX_CHECK(rm_update(s, wk.prof, W, ldw, j + w, m - j - w, j + w, cr, j, w));
            }
        }
        const int nr = n - (J + jb);
        if (!look) {
            if (nr > 0) {
                FLX_CHECK(trsm_rm_llu(s, wk.prof, (const double *)W + (long)J * ldw + J, ldw,
                                      W + (long)J * ldw + J + jb, ldw, jb, nr));
                FLX_CHECK(rm_update(s, wk.prof, W, ldw, J + jb, m - J - jb, J + jb, nr, J, jb));
            }
            continue;
        }
        // the previous block's side-stream work must be done before its
        // columns move again
        if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, wk.ev_r, 0));
        rest_pending = false;
        FLX_CHECK(list_moves(s, wk.perm, wk.inv, J, m, mvb));
        // The block's interchanges reach the next block's columns on the main
        // stream (its panels need them), every other column on the side
        // stream ahead of that stream's updates.  Then the next block's first
        // panel columns are updated on the critical path; its other columns,
        // then the rest of the trailing matrix, on the side stream beside that
        // panel (the next block's fused steps wait for the former, its
        // deferred moves for all of it).
        const int nl = nr > 0 ? (nr < nb ? nr : nb) : 0;
        const int c0 = J + jb + nl;
        // While the trailing matrix is large its update on the side stream
        // takes longer than the next block's panels, and the block end waits
        // for it: those blocks' side launches get an eighth of the compute
        // units more.  Measured: at n = 8192, nb = 128 up to 0.8 n rows
        // (38.65 -> 38.2 ms per solve); at n = 16384, nb = 256 up to 0.6 n
        // (142.0 -> 133.2 ms); more workgroups, or for more blocks, is slower.
        const bool early = (long)(m - J - jb) * 10 > (long)m * (nb >= 256 ? 6 : 8);
        const int sg = early && wk.side_groups_early > 0 ? wk.side_groups_early : wk.side_groups;
        FLX_CHECK(apply_moves<double>(s, wk.prof, W, ldw, m, 0, J + jb, c0, mvb, wk.moved));
        const bool side_moves = J > 0 || n > c0;
        const bool split = wk.ev_n && nl > PANEL_W;
        if (side_moves || split) {
            FLX_CHECK(hipEventRecord(wk.ev_a, s));
            FLX_CHECK(hipStreamWaitEvent(wk.side, wk.ev_a, 0));
        }
        if (side_moves)
            FLX_CHECK(apply_moves<double>(wk.side, wk.prof, W, ldw, m, J, c0, n, mvb, wk.moved));
        if (nl > 0) {
            const int nf = split ? PANEL_W : nl;
            const double *L11 = (const double *)W + (long)J * ldw + J;
            FLX_CHECK(trsm_rm_llu(s, wk.prof, L11, ldw, W + (long)J * ldw + J + jb, ldw, jb, nf));
            FLX_CHECK(rm_update(s, wk.prof, W, ldw, J + jb, m - J - jb, J + jb, nf, J, jb));
            if (split) {
                FLX_CHECK(trsm_rm_llu(wk.side, wk.prof, L11, ldw, W + (long)J * ldw + J + jb + nf,
                                      ldw, jb, nl - nf, sg));
                FLX_CHECK(rm_update(wk.side, wk.prof, W, ldw, J + jb, m - J - jb, J + jb + nf,
                                    nl - nf, J, jb, sg));
                FLX_CHECK(hipEventRecord(wk.ev_n, wk.side));
                next_pending = true;
            }
            if (n > c0) {
                FLX_CHECK(trsm_rm_llu(wk.side, wk.prof, L11, ldw, W + (long)J * ldw + c0, ldw, jb,
                                      n - c0, sg));
                FLX_CHECK(rm_update(wk.side, wk.prof, W, ldw, J + jb, m - J - jb, c0, n - c0, J,
                                    jb, sg));
            }
        }
        if (side_moves || split) {
            FLX_CHECK(hipEventRecord(wk.ev_r, wk.side));
            rest_pending = true;
        }
    }
    // (columns right of the last block, when n > m, may still be on the side stream)
    if (next_pending) FLX_CHECK(hipStreamWaitEvent(s, wk.ev_n, 0));
    if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, wk.ev_r, 0));
    return hipSuccess;
}

} // namespace flx
