// mixed.hip -- low-precision LU factorisation for iterative refinement.
//
// The factor of A (FP64, column-major, caller's) is computed in a row-major
// FP32 working copy W: FP32 panels and triangular solves, and a trailing
// update in FP32 (v_mfma_f32_16x16x4_f32) or in BF16 / FP16 with FP32
// accumulation (v_mfma_f32_16x16x32_bf16 / _f16).  It is the factorisation
// step of a mixed-precision solver in the style of LAPACK's dsgesv (netlib
// dsgesv.f:131-337, reached from FLENS as cxxlapack::sgesv -> dsgesv_,
// cxxlapack/netlib/interface/lapack.in.h:10611-10623).  Refinement (FP64
// residuals, classic and GMRES) lives in ir.hip.
//
// Pivoting: tournament pivoting ("CALU", Grigori, Demmel & Xiang 2011).
// The pivot rows of a 32-column panel are chosen by partial-pivoting
// eliminations on blocks of 256 rows (leaves), whose 32 winners each go up
// a reduction tree of fan-in 8 where the same elimination on the stacked
// candidates' original rows picks 32 again.  One kernel launch per panel:
// the last workgroup to finish a tree node's children (an atomic arrival
// count, no spinning) climbs to that node.  The root's elimination yields
// L11\U11 of the panel directly; L21 = A21 U11^-1 is one independent
// substitution per row.  This replaces the per-column grid-wide pivot
// exchange of partial pivoting (one device round trip per column) with
// about one per panel.  The factor is only a preconditioner for FP64
// refinement, so it does not have to reproduce the reference's FP32 pivot
// sequence; its quality is checked by its backward error.
#include <cfloat>
#include <climits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// Tournament panel
// ---------------------------------------------------------------------------
constexpr int TS_B = 32;     // panel width (max)
constexpr int TS_R = 256;    // rows per leaf = threads per workgroup
constexpr int TS_F = 8;      // tree fan-in: 8 children x 32 candidates = 256 rows

constexpr int TS_SLD = 256;        // staged row length (floats): the widest block
constexpr int TS_SLOTS = 2 * TS_B; // staged rows: winners, then displaced top rows
static_assert(TS_SLD == MIX_STAGE_LD && TS_SLOTS == MIX_STAGE_ROWS, "stage layout");
// u11 buffer: L11\U11 (32 x 32); U11 right of the diagonal, row c shifted
// to start at column c + 1: uus[c][i] = U(c, c + 1 + i); L11 below the
// diagonal by columns, likewise: lts[k][i] = L(k + 1 + i, k) (both zero
// outside the triangle); 1 / diag(U11) (32)
constexpr int TS_UU = TS_B * TS_B, TS_LT = TS_UU + TS_B * TS_B, TS_DINV = TS_LT + TS_B * TS_B;
constexpr int TS_U11 = TS_DINV + TS_B;
static_assert(TS_U11 == MIX_U11_FLOATS, "u11 layout");
static_assert(TS_B % 32 == 0 && MIX_STAGE_LD % 32 == 0, "float4 chunks of the stage");
static_assert(TS_R == 256, "the selection key holds the thread index in 8 bits");

// 256 threads, one row each.  Selection key (32 bits): |v| with its low 8
// mantissa bits replaced by 255 - thread, so the wave maximum is one DPP
// reduction and ties go to the lower thread.  Dropping those bits lets a
// pivot within a factor 1 - 2^-15 of the column maximum win, which is
// immaterial for the stability of the factor.
struct TsShared {
    float4 row[2][4][TS_B / 4];   // per-wave candidate rows, double-buffered
    uint32_t key[2][4];
    int win[TS_B];                // selected positions in order
    float u[TS_B][TS_B + 1];      // root: the selected rows' L\U entries
};

FLX_DEV uint32_t ts_key32(float v, int t)
{
    uint32_t a = __float_as_uint(fabsf(v)) & ~0xffu;
    a = a < 0x100u ? 0x100u : a;           // nonzero for every live row
    return a | (uint32_t)(255 - t);
}

// Partial-pivoting elimination over the workgroup's rows (one per thread,
// the column loop unrolled at compile time so every register index is a
// constant, one barrier per column).  Selects up to w rows into sh.win
// (positions, in order) and returns how many.  When u11 is given (the root)
// every selected row's eliminated values -- L multipliers left of the
// diagonal, U entries from it -- are written to u11 in selection order.
FLX_DEV int ts_select(float (&r)[TS_B], bool act, int pos, int w, TsShared &sh,
                      float *__restrict__ u11)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int nsel = 0, sel = -1;
    StaticFor<0, TS_B>::run([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j < w) {
            const int par = j & 1;
            const uint32_t key = act ? ts_key32(r[j], t) : 0u;
            const uint32_t wk = wave_umax(key);
            if (key != 0u && key == wk) {
#pragma unroll
                for (int c = j / 4; c < TS_B / 4; ++c)
                    sh.row[par][wave][c] = make_float4(r[4 * c], r[4 * c + 1], r[4 * c + 2],
                                                       r[4 * c + 3]);
            }
            if (lane == 0) sh.key[par][wave] = wk;
            __syncthreads();
            uint32_t best = sh.key[par][0];
            int bw = 0;
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                const uint32_t kq = sh.key[par][q];
                if (kq > best) { best = kq; bw = q; }
            }
            if (best != 0u) {
                float pr[TS_B];
#pragma unroll
                for (int c = j / 4; c < TS_B / 4; ++c) {
                    const float4 v = sh.row[par][bw][c];
                    pr[4 * c] = v.x; pr[4 * c + 1] = v.y; pr[4 * c + 2] = v.z; pr[4 * c + 3] = v.w;
                }
                if (key == best) {
                    act = false;
                    sel = j;
                    sh.win[j] = pos;
                }
                nsel = j + 1;
                if (act) {
                    const float p = pr[j];
                    const float l = (p != 0.0f) ? r[j] / p : 0.0f;
                    r[j] = l;
#pragma unroll
                    for (int c = j + 1; c < TS_B; ++c) r[c] = fmaf(-l, pr[c], r[c]);
                }
            }
        }
    });
    if (u11 && sel >= 0) {
        float4 *dst = reinterpret_cast<float4 *>(u11 + sel * TS_B);
#pragma unroll
        for (int c = 0; c < TS_B / 4; ++c)
            dst[c] = make_float4(r[4 * c], r[4 * c + 1], r[4 * c + 2], r[4 * c + 3]);
#pragma unroll
        for (int c = 0; c < TS_B; ++c) sh.u[sel][c] = r[c];
    }
    __syncthreads();
    return nsel;
}

FLX_DEV int ld_agent_i(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
FLX_DEV void st_agent_i(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct TsTree {
    int nlev;        // levels, 0 = leaves
    int count[6];    // nodes per level
    int off[6];      // first node id of each level
};

// one panel row (w <= 32 entries, 128-byte aligned when w == 32) into r
FLX_DEV void ts_load_row(float (&r)[TS_B], const float *__restrict__ src, bool ok, int w)
{
    if (ok && w == TS_B) {
        const float4 *s4 = reinterpret_cast<const float4 *>(src);
#pragma unroll
        for (int c = 0; c < TS_B / 4; ++c) {
            const float4 v = s4[c];
            r[4 * c] = v.x; r[4 * c + 1] = v.y; r[4 * c + 2] = v.z; r[4 * c + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < TS_B; ++c) r[c] = (ok && c < w) ? src[c] : 0.0f;
    }
}

// P: panel top-left in W (row-major, ld ldw), mp rows x w (<= 32) columns.
// Outputs: u11 (TS_U11 floats: the 32 x 32 L11\U11 in selection order,
// then U11's rows and L11's columns each padded with zeros to 64 entries,
// then the 32 reciprocals of U11's diagonal, 0 for a zero pivot), win
// (selected positions, relative to the panel top), list (one interchange
// slice, absolute rows r0 + position), zero (first zero-pivot column,
// INT_MAX if none).
// With stage given, the root also copies the rows that move -- winner k to
// slot k, the k-th displaced top row to slot 32 + k -- over the jb columns of
// the outer block (Wb: the block's first column in the panel's top row,
// row stride ldw) into stage (TS_SLD floats per slot), and their entries of
// perm (absolute positions) into stage_perm, so the interchange can be
// applied in place by a later kernel whose workgroups run in any order.
// One launch per panel: the last workgroup to finish a tree node's
// children (an atomic arrival count, no spinning) climbs to that node.
__global__ __launch_bounds__(TS_R) void tslu_kernel(const float *__restrict__ P, long ldw, int mp,
                                                    int w, int r0, TsTree tree,
                                                    int *__restrict__ cand,
                                                    int *__restrict__ ccount,
                                                    unsigned *__restrict__ counters,
                                                    float *__restrict__ u11,
                                                    int *__restrict__ win,
                                                    int *__restrict__ list,
                                                    int *__restrict__ zero,
                                                    const float *__restrict__ Wb, int jb,
                                                    float *__restrict__ stage,
                                                    const int *__restrict__ perm,
                                                    int *__restrict__ stage_perm)
{
    critical_path_priority();
    __shared__ TsShared sh;
    __shared__ int s_last;
    __shared__ int s_D[TS_B], s_V[TS_B], s_n[2];
    const int t = threadIdx.x;
    int node = blockIdx.x;     // index within the level
    int lev = 0;
    float r[TS_B];
    int pos = node * TS_R + t;
    bool act = pos < mp;
    ts_load_row(r, P + (long)(act ? pos : 0) * ldw, act, w);
    int nsel;
    for (;;) {
        const bool root = (lev == tree.nlev - 1);
        nsel = ts_select(r, act, pos, w, sh, root ? u11 : nullptr);
        if (root) break;
        const int id = tree.off[lev] + node;
        if (t < nsel) st_agent_i(&cand[(long)id * TS_B + t], sh.win[t]);
        if (t == 0) st_agent_i(&ccount[id], nsel);
        __threadfence();
        __syncthreads();
        const int parent = node / TS_F;
        const int c0 = parent * TS_F;
        const int nchild = min(TS_F, tree.count[lev] - c0);
        const int pid = tree.off[lev + 1] + parent;
        if (t == 0) {
            const unsigned old = atomicAdd(&counters[pid], 1u);
            s_last = (old == (unsigned)(nchild - 1));
            if (s_last) atomicExch(&counters[pid], 0u);   // ready for the next launch
        }
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        // gather the children's candidates: thread t -> child t/32, slot t%32
        const int ch = t / TS_B, k = t % TS_B;
        int p = -1;
        if (ch < nchild) {
            const int cid = tree.off[lev] + c0 + ch;
            if (k < ld_agent_i(&ccount[cid])) p = ld_agent_i(&cand[(long)cid * TS_B + k]);
        }
        act = (p >= 0 && p < mp);
        pos = act ? p : INT_MAX;
        ts_load_row(r, P + (long)(act ? p : 0) * ldw, act, w);
        node = parent;
        ++lev;
    }
    // root.  u11's rows were written by their threads in ts_select, and
    // copied to sh.u (its closing barrier orders them before the reads
    // below); the rows >= nsel and the shifted triangles follow.
#pragma unroll
    for (int q = 0; q < TS_B * TS_B / TS_R; ++q) {
        const int e = t + q * TS_R;
        const int k = e / TS_B, i = e % TS_B;
        if (k >= nsel) u11[e] = 0.0f;
        const int c = k + 1 + i;
        u11[TS_UU + e] = (k < nsel && c < nsel) ? sh.u[k][c] : 0.0f;
        u11[TS_LT + e] = (c < nsel) ? sh.u[c][k] : 0.0f;
    }
    if (t < TS_B) {
        // pivot reciprocals; the first exactly-zero pivot column
        const float d = (t < nsel) ? sh.u[t][t] : 1.0f;
        u11[TS_DINV + t] = (t < nsel && d != 0.0f) ? 1.0f / d : 0.0f;
        if (d == 0.0f) atomicMin(zero, r0 + t);
    }
    // the interchange list (wave 0).  Rows k < nsel receive winner k; the
    // displaced top rows (not winners) fill the winners' vacated slots
    // (>= nsel), both in ascending order.
    if (t < 64) {
        const int k = t;
        const int wkp = (k < nsel) ? sh.win[k] : -1;
        if (k < nsel) win[k] = wkp;
        bool top_is_win = false;
        if (k < nsel)
            for (int q = 0; q < nsel; ++q) top_is_win |= (sh.win[q] == k);
        const bool inD = (k < nsel) && !top_is_win;
        const bool inV = (k < nsel) && (wkp >= nsel);
        const uint64_t bD = __ballot(inD);
        const int rankD = __popcll(bD & ((1ull << k) - 1ull));
        int rankV = 0;
        if (inV)
            for (int q = 0; q < nsel; ++q) rankV += (sh.win[q] >= nsel && sh.win[q] < wkp);
        const bool mv1 = (k < nsel) && (wkp != k);
        const uint64_t b1 = __ballot(mv1);
        const int slot1 = __popcll(b1 & ((1ull << k) - 1ull));
        if (mv1) {
            list[1 + 2 * slot1] = r0 + k;
            list[2 + 2 * slot1] = r0 + wkp;
        }
        if (inD) s_D[rankD] = k;
        if (inV) s_V[rankV] = wkp;
        if (k == 0) { s_n[0] = __popcll(b1); s_n[1] = __popcll(bD); }
    }
    __syncthreads();
    const int n1 = s_n[0], nD = s_n[1];
    if (t < nD) {
        list[1 + 2 * (n1 + t)] = r0 + s_V[t];
        list[2 + 2 * (n1 + t)] = r0 + s_D[t];
    }
    if (t == 0) list[0] = n1 + nD;
    if (!stage) return;
    // staging: 4 threads per slot, each a quarter of the row
    const int slot = t >> 2, part = t & 3;
    int src = -1;
    if (slot < nsel) src = sh.win[slot];
    else if (slot >= TS_B && slot - TS_B < nD) src = s_D[slot - TS_B];
    if (src < 0) return;
    const float *from = Wb + (long)src * ldw;
    float *to = stage + (long)slot * TS_SLD;
    // all loads first (unconditional, clamped in bounds), then the stores:
    // one memory latency, not one per chunk
    if ((jb & 3) == 0) {
        float4 v[TS_SLD / 16];
#pragma unroll
        for (int q = 0; q < TS_SLD / 16; ++q) {
            const int c = 4 * part + 16 * q;
            v[q] = *reinterpret_cast<const float4 *>(from + (c < jb ? c : 0));
        }
#pragma unroll
        for (int q = 0; q < TS_SLD / 16; ++q) {
            const int c = 4 * part + 16 * q;
            if (c < jb) *reinterpret_cast<float4 *>(to + c) = v[q];
        }
    } else {
        float v[TS_SLD / 4];
#pragma unroll
        for (int q = 0; q < TS_SLD / 4; ++q) {
            const int c = part + 4 * q;
            v[q] = from[c < jb ? c : 0];
        }
#pragma unroll
        for (int q = 0; q < TS_SLD / 4; ++q)
            if (part + 4 * q < jb) to[part + 4 * q] = v[q];
    }
    if (part == 0) stage_perm[slot] = perm[r0 + src];
}

// The rest of a panel step in one launch, once tslu_kernel has chosen the
// pivots and staged the moving rows (block columns [J, J + jb), panel at
// columns jo .. jo + w of the block, rows r0 .. r0 + mp):
//   U12 = L11^-1 (the winners' columns right of the panel)   (each workgroup)
//   row p < nsel:  L11\U11 row p | U12 row p, left part from the stage
//   other rows:    x = a U11^-1 (the L21 row), a22 -= x U12; a row that moves
//                  in (a winner's vacated slot) takes its data from the stage
// and perm follows the rows.  Workgroups write only their own rows and read
// other rows only from the stage, so they need no ordering among themselves.
// AP_ROWS rows per workgroup, 256 / AP_ROWS threads per row.
constexpr int AP_ROWS = 32;
constexpr int AP_TPR = 256 / AP_ROWS;
constexpr int AP_CR = TS_SLD - TS_B;   // widest right part

__global__ __launch_bounds__(256) void panel_apply_kernel(float *__restrict__ Wb, long ldw, int mp,
                                                          int jo, int w, int jb,
                                                          const float *__restrict__ u11,
                                                          const int *__restrict__ list, int r0,
                                                          const float *__restrict__ stage,
                                                          const int *__restrict__ stage_perm,
                                                          int *__restrict__ perm)
{
    critical_path_priority();
    __shared__ float4 s_tri[2][TS_B][TS_B / 4];  // uus, lts (shifted triangles)
    __shared__ float s_di[TS_B];                 // 1 / diag(U11)
    __shared__ float4 s12[TS_B][AP_CR / 4];      // U12
    __shared__ float sx[AP_ROWS][TS_B + 4];      // L21 rows of the tile
    __shared__ int s_slot[AP_ROWS];
    const int t = threadIdx.x;
    const int nsel = min(w, mp);
    const int cr = jb - jo - w;                  // columns right of the panel
    const int p0 = blockIdx.x * AP_ROWS;
    // the triangles' coefficients into LDS (every thread one or two loads)
    {
        const float4 *g = reinterpret_cast<const float4 *>(u11 + TS_UU);
        float4 *l = &s_tri[0][0][0];
#pragma unroll
        for (int q = 0; q < 2 * TS_B * TS_B / 4 / 256; ++q) l[t + 256 * q] = g[t + 256 * q];
        if (t < TS_B) s_di[t] = u11[TS_DINV + t];
    }
    __syncthreads();
    if (t < 64) {
        // wave 0: which rows of the tile move in, then the panel columns
        if (t < AP_ROWS) s_slot[t] = (p0 + t < nsel) ? p0 + t : -1;
        __builtin_amdgcn_wave_barrier();
        // rows moving in below the winners: the list's entries past the
        // winners' own, in order, are the displaced top rows (stage 32 + d)
        const int cnt = list[0];
        const int dl = (t < cnt) ? list[1 + 2 * t] - r0 : -1;
        const bool in = dl >= nsel;
        const uint64_t m = __ballot(in);
        const int d = __popcll(m & ((1ull << t) - 1ull));
        if (in && dl >= p0 && dl < p0 + AP_ROWS) s_slot[dl - p0] = TS_B + d;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // one thread per row
        const int p = p0 + t;
        if (t < AP_ROWS && p < mp) {
            float *dst = Wb + (long)p * ldw + jo;
            const int slot = s_slot[t];
            if (p < nsel) {
                float v[TS_B];
#pragma unroll
                for (int c = 0; c < TS_B; ++c) v[c] = u11[p * TS_B + c];
#pragma unroll
                for (int c = 0; c < TS_B; ++c)
                    if (c < w) dst[c] = v[c];
            } else {
                const float *src = (slot >= 0) ? stage + (long)slot * TS_SLD + jo : dst;
                float a[TS_B];
#pragma unroll
                for (int c = 0; c < TS_B; ++c) a[c] = (c < w) ? src[c] : 0.0f;
                // x U11 = a, right-looking and rotating: at step c register 0
                // holds a_c, finished as x_c, and x_c U(c, c + 1 + i) leaves
                // the entries still open, which shift down by one, so all
                // register indices are constants in a rolled loop (short
                // code: this kernel runs once per panel on the critical path)
#pragma unroll 2
                for (int c = 0; c < TS_B; ++c) {
                    float co[TS_B];
#pragma unroll
                    for (int q = 0; q < TS_B / 4; ++q) {
                        const float4 v = s_tri[0][c][q];
                        co[4 * q] = v.x; co[4 * q + 1] = v.y; co[4 * q + 2] = v.z; co[4 * q + 3] = v.w;
                    }
                    const float xc = a[0] * s_di[c];
#pragma unroll
                    for (int i = 0; i < TS_B - 1; ++i) a[i] = fmaf(-xc, co[i], a[i + 1]);
                    a[TS_B - 1] = xc;
                }
#pragma unroll
                for (int c = 0; c < TS_B; ++c) {
                    sx[t][c] = (c < nsel) ? a[c] : 0.0f;
                    if (c < w) dst[c] = a[c];
                }
            }
            if (slot >= 0) perm[r0 + p] = stage_perm[slot];
        }
    } else if (cr > 0) {
        // waves 1-3, beside wave 0: U12 by forward substitution with the
        // unit-lower L11, one column per thread (two for the first 32
        // threads when more than 192 columns remain), rotating like x above
        const int c1 = t - 64, c2 = c1 + 192;
        const bool two = cr > 192 && t < 128;      // uniform in each wave
        float x[TS_B], y[TS_B];
#pragma unroll
        for (int k = 0; k < TS_B; ++k)
            x[k] = (k < nsel && c1 < cr) ? stage[k * TS_SLD + jo + w + c1] : 0.0f;
        if (two) {
#pragma unroll
            for (int k = 0; k < TS_B; ++k)
                y[k] = (k < nsel && c2 < cr) ? stage[k * TS_SLD + jo + w + c2] : 0.0f;
#pragma unroll 2
            for (int k = 0; k < TS_B; ++k) {
                float co[TS_B];
#pragma unroll
                for (int c = 0; c < TS_B / 4; ++c) {
                    const float4 v = s_tri[1][k][c];
                    co[4 * c] = v.x; co[4 * c + 1] = v.y; co[4 * c + 2] = v.z; co[4 * c + 3] = v.w;
                }
                const float x0 = x[0], y0 = y[0];
#pragma unroll
                for (int i = 0; i < TS_B - 1; ++i) {
                    x[i] = fmaf(-co[i], x0, x[i + 1]);
                    y[i] = fmaf(-co[i], y0, y[i + 1]);
                }
                x[TS_B - 1] = x0;
                y[TS_B - 1] = y0;
            }
        } else {
#pragma unroll 2
            for (int k = 0; k < TS_B; ++k) {
                float co[TS_B];
#pragma unroll
                for (int c = 0; c < TS_B / 4; ++c) {
                    const float4 v = s_tri[1][k][c];
                    co[4 * c] = v.x; co[4 * c + 1] = v.y; co[4 * c + 2] = v.z; co[4 * c + 3] = v.w;
                }
                const float x0 = x[0];
#pragma unroll
                for (int i = 0; i < TS_B - 1; ++i) x[i] = fmaf(-co[i], x0, x[i + 1]);
                x[TS_B - 1] = x0;
            }
        }
        float *col = reinterpret_cast<float *>(&s12[0][0]);
        if (c1 < cr)
#pragma unroll
            for (int k = 0; k < TS_B; ++k) col[k * AP_CR + c1] = x[k];
        if (two && c2 < cr)
#pragma unroll
            for (int k = 0; k < TS_B; ++k) col[k * AP_CR + c2] = y[k];
    }
    __syncthreads();
    // left part (rows that move) and the right part, AP_TPR threads per row;
    // loads are issued together ahead of their use
    const int i = t / AP_TPR, part = t % AP_TPR;
    const int p = p0 + i;
    if (p >= mp) return;
    const int slot = s_slot[i];
    float *row = Wb + (long)p * ldw;
    if (slot >= 0 && jo > 0) {
        // jo is a multiple of 32: float4 chunks
        const float4 *from = reinterpret_cast<const float4 *>(stage + (long)slot * TS_SLD);
        float4 *to = reinterpret_cast<float4 *>(row);
        constexpr int NL = (TS_SLD - TS_B) / 4 / AP_TPR;
        float4 v[NL];
#pragma unroll
        for (int q = 0; q < NL; ++q) {
            const int c4 = part + AP_TPR * q;
            v[q] = from[c4 < jo / 4 ? c4 : 0];
        }
#pragma unroll
        for (int q = 0; q < NL; ++q)
            if (part + AP_TPR * q < jo / 4) to[part + AP_TPR * q] = v[q];
    }
    if (cr <= 0) return;
    float *rr = row + jo + w;
    if (p < nsel) {
        for (int c = part; c < cr; c += AP_TPR)
            rr[c] = reinterpret_cast<const float *>(&s12[0][0])[p * AP_CR + c];
        return;
    }
    const float *src = (slot >= 0) ? stage + (long)slot * TS_SLD + jo + w : rr;
    float xr[TS_B];
#pragma unroll
    for (int k = 0; k < TS_B; ++k) xr[k] = sx[i][k];
    if ((cr & 3) == 0 && (w & 3) == 0 && (ldw & 3) == 0) {
        // rolled, with the next chunk's load issued before this chunk's FMAs
        const int n4 = cr / 4;
        int c4 = part;
        float4 v = *reinterpret_cast<const float4 *>(src + 4 * (c4 < n4 ? c4 : 0));
#pragma unroll 1
        for (; c4 < n4; c4 += AP_TPR) {
            const int nx = c4 + AP_TPR;
            const float4 vn = *reinterpret_cast<const float4 *>(src + 4 * (nx < n4 ? nx : 0));
#pragma unroll
            for (int k = 0; k < TS_B; ++k) {
                const float4 u = s12[k][c4];
                v.x = fmaf(-xr[k], u.x, v.x);
                v.y = fmaf(-xr[k], u.y, v.y);
                v.z = fmaf(-xr[k], u.z, v.z);
                v.w = fmaf(-xr[k], u.w, v.w);
            }
            *reinterpret_cast<float4 *>(rr + 4 * c4) = v;
            v = vn;
        }
    } else {
        const float *u12 = reinterpret_cast<const float *>(&s12[0][0]);
#pragma unroll 1
        for (int c = part; c < cr; c += AP_TPR) {
            float v = src[c];
#pragma unroll
            for (int k = 0; k < TS_B; ++k) v = fmaf(-xr[k], u12[k * AP_CR + c], v);
            rr[c] = v;
        }
    }
}

// Writes the panel's factor in place, before the rows move: selected rows
// get their L11\U11 row, every other row its L21 row  x U11 = a.
__global__ __launch_bounds__(256) void tslu_finish_kernel(float *__restrict__ P, long ldw, int mp,
                                                          int w, const float *__restrict__ u11,
                                                          const int *__restrict__ win, int nsel)
{
    critical_path_priority();
    __shared__ float su[TS_B][TS_B + 1];
    __shared__ int sw[TS_B];
    const int t = threadIdx.x;
    for (int q = t; q < TS_B * TS_B; q += 256) su[q / TS_B][q % TS_B] = u11[q];
    if (t < TS_B) sw[t] = (t < nsel) ? win[t] : -1;
    __syncthreads();
    const int pos = blockIdx.x * 256 + t;
    if (pos >= mp) return;
    int k = -1;
    for (int q = 0; q < nsel; ++q) k = (sw[q] == pos) ? q : k;
    float *row = P + (long)pos * ldw;
    float x[TS_B];
    if (k >= 0) {
#pragma unroll
        for (int c = 0; c < TS_B; ++c) x[c] = su[k][c];
    } else {
        ts_load_row(x, row, true, w);
#pragma unroll
        for (int j = 0; j < TS_B; ++j) {
            if (j < nsel) {
                float v = x[j];
#pragma unroll
                for (int q = 0; q < j; ++q) v = fmaf(-x[q], su[q][j], v);
                const float d = su[j][j];
                x[j] = (d != 0.0f) ? v / d : v;
            }
        }
    }
    if (w == TS_B) {
        float4 *r4 = reinterpret_cast<float4 *>(row);
#pragma unroll
        for (int c = 0; c < TS_B / 4; ++c)
            r4[c] = make_float4(x[4 * c], x[4 * c + 1], x[4 * c + 2], x[4 * c + 3]);
    } else {
#pragma unroll
        for (int c = 0; c < TS_B; ++c)
            if (c < w) row[c] = x[c];
    }
}

size_t tslu_nodes(int mp)
{
    size_t total = 0;
    int c = (mp + TS_R - 1) / TS_R;
    for (;;) {
        total += c;
        if (c <= 1) break;
        c = (c + TS_F - 1) / TS_F;
    }
    return total;
}

static TsTree tslu_tree(int mp)
{
    TsTree t{};
    int c = (mp + TS_R - 1) / TS_R, off = 0;
    t.nlev = 0;
    for (;;) {
        t.count[t.nlev] = c;
        t.off[t.nlev] = off;
        off += c;
        ++t.nlev;
        if (c <= 1) break;
        c = (c + TS_F - 1) / TS_F;
    }
    return t;
}

// ---------------------------------------------------------------------------
// Low-precision trailing update  C (nr x nc, row-major FP32) += La * Ub^T
// with La (nr_pad x Kp) and Ub (nc_pad x Kp) packed 16-bit operands, k
// contiguous, so each MFMA operand fragment is one 16-byte load:
//   A operand: lane l holds La[i0 + l%16][k0 + 8(l/16) .. +7]
//   B operand: lane l holds Ub[j0 + l%16][k0 + 8(l/16) .. +7]
//   result item q of lane l: C[i0 + 4(l/16) + q][j0 + l%16]
// 256 threads = 4 waves in 2 x 2, a 64 x 64 sub-tile each (4 x 4 MFMAs).
// The update is bound by streaming C (8 bytes per element per step), not by
// the matrix cores, so the operands come straight from global memory/L2.
// ---------------------------------------------------------------------------
template <typename V>
FLX_DEV float4_t mfma_lp(V a, V b, float4_t c);
template <>
__device__ __forceinline__ float4_t mfma_lp<bf16x8_t>(bf16x8_t a, bf16x8_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
template <>
__device__ __forceinline__ float4_t mfma_lp<f16x8_t>(f16x8_t a, f16x8_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// operand scale: 2^(14 - e) with max = f * 2^e, f in [0.5, 1): the largest
// magnitude lands in [2^13, 2^14), well inside FP16's range (65504).
FLX_DEV float lp_scale(const unsigned *maxbits)
{
    if (!maxbits) return 1.0f;
    const float m = __uint_as_float(*maxbits);
    if (!(m > 0.0f) || !(m <= FLT_MAX)) return 1.0f;
    int e;
    frexpf(m, &e);
    return ldexpf(1.0f, 14 - e);
}

template <typename V>
__global__ __launch_bounds__(256) void gemm_lp_kernel(const V *__restrict__ La,
                                                      const V *__restrict__ Ub, int Kp,
                                                      float *__restrict__ C, long ldc, int nr,
                                                      int nc, const unsigned *__restrict__ maxes,
                                                      int tx, int tiles)
{
  // workgroup b takes the 128x128 tiles b, b + grid, ... (column tile fastest)
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = (t / tx) * 128 + (wave >> 1) * 64;
    const int j0 = (t % tx) * 128 + (wave & 1) * 64;
    const int lr = lane & 15, lk = lane >> 4;
    float4_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
    // the C tile is requested up front so its HBM latency overlaps the MFMAs
    float cpre[4][4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = i0 + a * 16 + 4 * lk + q;
            const float *cr = C + (long)(row < nr ? row : 0) * ldc;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int col = j0 + b * 16 + lr;
                cpre[a][b][q] = (row < nr && col < nc) ? cr[col] : 0.0f;
            }
        }
    // operands are indexed in 8-element vectors: a row holds Kp / 8
    const int kv = Kp / 8;
    const V *pa = La + (long)(i0 + lr) * kv + lk;
    const V *pb = Ub + (long)(j0 + lr) * kv + lk;
    const long sa = 16L * kv;   // next 16-row fragment
    V fa[4], fb[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) fa[a] = pa[a * sa];
#pragma unroll
    for (int b = 0; b < 4; ++b) fb[b] = pb[b * sa];
    for (int k = 0; k < kv; k += 4) {
        V na[4], nb[4];
        const bool more = k + 4 < kv;
        if (more) {
#pragma unroll
            for (int a = 0; a < 4; ++a) na[a] = pa[a * sa + k + 4];
#pragma unroll
            for (int b = 0; b < 4; ++b) nb[b] = pb[b * sa + k + 4];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = mfma_lp<V>(fa[a], fb[b], acc[a][b]);
        if (more) {
#pragma unroll
            for (int a = 0; a < 4; ++a) { fa[a] = na[a]; fb[a] = nb[a]; }
        }
    }
    float inv = 1.0f;
    if (maxes) inv = 1.0f / (lp_scale(maxes) * lp_scale(maxes + 1));
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = i0 + a * 16 + 4 * lk + q;
            if (row < nr) {
                float *cr = C + (long)row * ldc;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int col = j0 + b * 16 + lr;
                    if (col < nc) cr[col] = fmaf(acc[a][b][q], inv, cpre[a][b][q]);
                }
            }
        }
  }
}

// max |x| over an nr x nc row-major block into *out (float bits, >= 0)
__global__ __launch_bounds__(256) void absmax_rm_kernel(const float *__restrict__ X, long ldx,
                                                        int nr, int nc, unsigned *out)
{
    float m = 0.0f;
    const long total = (long)nr * nc;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
        const int i = (int)(q / nc), j = (int)(q % nc);
        m = fmaxf(m, fabsf(X[(long)i * ldx + j]));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

template <typename E> FLX_DEV E to_lp(float v);
template <> __device__ __forceinline__ __bf16 to_lp<__bf16>(float v) { return (__bf16)v; }
template <> __device__ __forceinline__ _Float16 to_lp<_Float16>(float v) { return (_Float16)v; }

// La[i][k] = -L[i][k] * s  (rows i < nr_pad, k < Kp; zero outside nr x kk)
template <typename E>
__global__ __launch_bounds__(256) void pack_l_kernel(const float *__restrict__ L, long ldl, int nr,
                                                     int kk, E *__restrict__ La, int nr_pad,
                                                     int Kp, const unsigned *__restrict__ maxes)
{
    const float s = -lp_scale(maxes);
    const long total = (long)nr_pad * Kp;
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < total; q += (long)gridDim.x * 256) {
        const int i = (int)(q / Kp), k = (int)(q % Kp);
        const float v = (i < nr && k < kk) ? L[(long)i * ldl + k] * s : 0.0f;
        La[q] = to_lp<E>(v);
    }
}

// Ub[j][k] = U[k][j] * s  (j < nc_pad, k < Kp; zero outside kk x nc);
// 32 x 32 tiles transposed through LDS.
template <typename E>
__global__ __launch_bounds__(256) void pack_u_kernel(const float *__restrict__ U, long ldu, int kk,
                                                     int nc, E *__restrict__ Ub, int nc_pad,
                                                     int Kp, const unsigned *__restrict__ maxes)
{
    __shared__ float tile[32][33];
    const float s = lp_scale(maxes);
    const int k0 = blockIdx.y * 32, j0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int q = ty; q < 32; q += 8) {
        const int k = k0 + q, j = j0 + tx;
        tile[q][tx] = (k < kk && j < nc) ? U[(long)k * ldu + j] * s : 0.0f;
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const int j = j0 + q, k = k0 + tx;
        if (j < nc_pad && k < Kp) Ub[(long)j * Kp + k] = to_lp<E>(tile[tx][q]);
    }
}

__global__ void zero_u32_kernel(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

size_t lp_pack_elems(int n, int nb)
{
    const int Kp = ((nb + 31) / 32) * 32;
    const int np = ((n + 127) / 128) * 128;
    return (size_t)np * Kp;
}

// C[r0.., c0..] -= W[r0.., k0..k0+kk) * W[k0..k0+kk, c0..) in precision prec
static hipError_t lp_update(hipStream_t s, const MixWork &mw, int prec, float *W, long ldw,
                            int r0, int nr, int c0, int nc, int k0, int kk, int max_groups = 0)
{
    if (nr <= 0 || nc <= 0 || kk <= 0) return hipSuccess;
    ProfScope ps(mw.prof, s, PROF_GEMM, 2.0 * nr * nc * kk,
                 4.0 * (2.0 * nr * nc + (double)nr * kk + (double)kk * nc));
    if (prec == MIX_FP32)
        return gemm_f32(s, 0, 0, nc, nr, kk, -1.0f, W + (long)k0 * ldw + c0, ldw,
                        W + (long)r0 * ldw + k0, ldw, 1.0f, W + (long)r0 * ldw + c0, ldw, 0,
                        max_groups);
    const int Kp = ((kk + 31) / 32) * 32;
    const int nr_pad = ((nr + 127) / 128) * 128, nc_pad = ((nc + 127) / 128) * 128;
    unsigned *mx = nullptr;
    const float *L = W + (long)r0 * ldw + k0;
    const float *U = W + (long)k0 * ldw + c0;
    if (prec == MIX_FP16) {
        mx = mw.lpmax;
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, mx, 2);
        hipLaunchKernelGGL(absmax_rm_kernel, dim3(512), dim3(256), 0, s, L, ldw, nr, kk, mx);
        hipLaunchKernelGGL(absmax_rm_kernel, dim3(512), dim3(256), 0, s, U, ldw, kk, nc, mx + 1);
    }
    const long la_total = (long)nr_pad * Kp;
    const int gl = (int)((la_total + 255) / 256 < 8192 ? (la_total + 255) / 256 : 8192);
    dim3 gu((nc_pad + 31) / 32, (Kp + 31) / 32);
    const int gtx = nc_pad / 128, gtiles = gtx * (nr_pad / 128);
    dim3 gg(max_groups > 0 && max_groups < gtiles ? max_groups : gtiles);
    if (prec == MIX_BF16) {
        __bf16 *La = reinterpret_cast<__bf16 *>(mw.la), *Ub = reinterpret_cast<__bf16 *>(mw.ub);
        hipLaunchKernelGGL(pack_l_kernel<__bf16>, dim3(gl), dim3(256), 0, s, L, ldw, nr, kk, La,
                           nr_pad, Kp, (const unsigned *)nullptr);
        hipLaunchKernelGGL(pack_u_kernel<__bf16>, gu, dim3(256), 0, s, U, ldw, kk, nc, Ub, nc_pad,
                           Kp, (const unsigned *)nullptr);
        hipLaunchKernelGGL(gemm_lp_kernel<bf16x8_t>, gg, dim3(256), 0, s,
                           reinterpret_cast<const bf16x8_t *>(La),
                           reinterpret_cast<const bf16x8_t *>(Ub), Kp,
                           W + (long)r0 * ldw + c0, ldw, nr, nc, (const unsigned *)nullptr, gtx,
                           gtiles);
    } else {
        _Float16 *La = reinterpret_cast<_Float16 *>(mw.la), *Ub = reinterpret_cast<_Float16 *>(mw.ub);
        hipLaunchKernelGGL(pack_l_kernel<_Float16>, dim3(gl), dim3(256), 0, s, L, ldw, nr, kk,
                           La, nr_pad, Kp, (const unsigned *)mx);
        hipLaunchKernelGGL(pack_u_kernel<_Float16>, gu, dim3(256), 0, s, U, ldw, kk, nc, Ub,
                           nc_pad, Kp, (const unsigned *)(mx + 1));
        hipLaunchKernelGGL(gemm_lp_kernel<f16x8_t>, gg, dim3(256), 0, s,
                           reinterpret_cast<const f16x8_t *>(La),
                           reinterpret_cast<const f16x8_t *>(Ub), Kp,
                           W + (long)r0 * ldw + c0, ldw, nr, nc, (const unsigned *)mx, gtx,
                           gtiles);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Driver: factor the n x n row-major FP32 W in place; perm[i] = original row
// now at position i; *zero = first zero-pivot column (0-based, INT_MAX if
// none; caller initialises).
// ---------------------------------------------------------------------------
hipError_t getrf_rm_mixed(hipStream_t s, const MixWork &mw, float *W, long ldw, int n,
                          int prec, int nb)
{
    if (n <= 0) return hipSuccess;
    const int ib = TS_B;
    if (nb < ib) nb = ib;
    if (nb > 256) nb = 256;
    nb = (nb / ib) * ib;
    // lookahead: interchanges inside the outer block first, the rest of the
    // columns once per block; the next block's columns are updated (in FP32)
    // before the rest of the trailing update, which runs in prec on the side
    // stream beside the next block's panels
    const bool look = mw.inv && mw.mv && mw.side && mw.ev_a && mw.ev_r;
    // the fused panel step moves only the block's columns, as the lookahead does
    const bool fused = look && mw.stage && mw.stage_perm;
    FLX_CHECK(iota(s, mw.perm, n));
    bool rest_pending = false;
    for (int J = 0; J < n; J += nb) {
        const int jb = (n - J < nb) ? n - J : nb;
        if (look) FLX_CHECK(inv_perm(s, mw.perm, mw.inv, n, mw.mv));
        for (int j = J; j < J + jb; j += ib) {
            const int w = (J + jb - j < ib) ? J + jb - j : ib;
            const int mp = n - j;
            const TsTree tree = tslu_tree(mp);
            if (fused) {
                // pivots, then the whole in-block step in one launch
                float *Wb = W + (long)j * ldw + J;
                {
                    ProfScope ps(mw.prof, s, PROF_PANEL, 0.0, 4.0 * mp * w);
                    hipLaunchKernelGGL(tslu_kernel, dim3(tree.count[0]), dim3(TS_R), 0, s,
                                       (const float *)(W + (long)j * ldw + j), ldw, mp, w, j,
                                       tree, mw.cand, mw.ccount, mw.counters, mw.u11, mw.win,
                                       mw.list, mw.zero, (const float *)Wb, jb, mw.stage,
                                       (const int *)mw.perm, mw.stage_perm);
                }
                ProfScope ps(mw.prof, s, PROF_TRSM_RM, 2.0 * (mp - w) * (double)w * (J + jb - j - w),
                             8.0 * mp * (double)(J + jb - j));
                hipLaunchKernelGGL(panel_apply_kernel, dim3((mp + AP_ROWS - 1) / AP_ROWS),
                                   dim3(256), 0, s, Wb, ldw, mp, j - J, w, jb,
                                   (const float *)mw.u11, (const int *)mw.list, j,
                                   (const float *)mw.stage, (const int *)mw.stage_perm, mw.perm);
                FLX_CHECK(hipGetLastError());
                continue;
            }
            {
                ProfScope ps(mw.prof, s, PROF_PANEL, 0.0, 8.0 * mp * w);
                hipLaunchKernelGGL(tslu_kernel, dim3(tree.count[0]), dim3(TS_R), 0, s,
                                   (const float *)(W + (long)j * ldw + j), ldw, mp, w, j, tree,
                                   mw.cand, mw.ccount, mw.counters, mw.u11, mw.win, mw.list,
                                   mw.zero, (const float *)nullptr, 0, (float *)nullptr,
                                   (const int *)nullptr, (int *)nullptr);
                hipLaunchKernelGGL(tslu_finish_kernel, dim3((mp + 255) / 256), dim3(256), 0, s,
                                   W + (long)j * ldw + j, ldw, mp, w, (const float *)mw.u11,
                                   (const int *)mw.win, w < mp ? w : mp);
                FLX_CHECK(hipGetLastError());
            }
            if (look)
                FLX_CHECK(rowperm_rm<float>(s, mw.prof, W + J, ldw, n, jb, jb, jb, mw.list, 1,
                                            mw.perm, nullptr));
            else
                FLX_CHECK(rowperm_rm<float>(s, mw.prof, W, ldw, n, n, n, n, mw.list, 1, mw.perm,
                                            nullptr));
            const int cr = J + jb - (j + w);
            if (cr > 0) {
                FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)j * ldw + j, ldw,
                                             W + (long)j * ldw + j + w, ldw, w, cr));
                FLX_CHECK(lp_update(s, mw, MIX_FP32, W, ldw, j + w, n - j - w, j + w, cr, j, w));
            }
        }
        const int nr = n - (J + jb);
        if (!look) {
            if (nr > 0) {
                FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)J * ldw + J, ldw,
                                             W + (long)J * ldw + J + jb, ldw, jb, nr));
                FLX_CHECK(lp_update(s, mw, prec, W, ldw, J + jb, nr, J + jb, nr, J, jb));
            }
            continue;
        }
        // the previous block's rest update reads columns this block moves
        if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, mw.ev_r, 0));
        rest_pending = false;
        FLX_CHECK(defer_moves<float>(s, mw.prof, W, ldw, n, n, J, jb, mw.perm, mw.inv, mw.mv,
                                     nullptr));
        if (nr > 0) {
            const int nl = nr < nb ? nr : nb;
            FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)J * ldw + J, ldw,
                                         W + (long)J * ldw + J + jb, ldw, jb, nl));
            FLX_CHECK(lp_update(s, mw, MIX_FP32, W, ldw, J + jb, nr, J + jb, nl, J, jb));
            const int c0 = J + jb + nl;
            if (n > c0) {
                FLX_CHECK(hipEventRecord(mw.ev_a, s));
                FLX_CHECK(hipStreamWaitEvent(mw.side, mw.ev_a, 0));
                FLX_CHECK(trsm_rm_llu<float>(mw.side, mw.prof, W + (long)J * ldw + J, ldw,
                                             W + (long)J * ldw + c0, ldw, jb, n - c0,
                                             mw.side_groups));
                FLX_CHECK(lp_update(mw.side, mw, prec, W, ldw, J + jb, nr, c0, n - c0, J, jb,
                                    mw.side_groups));
                FLX_CHECK(hipEventRecord(mw.ev_r, mw.side));
                rest_pending = true;
            }
        }
    }
    if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, mw.ev_r, 0));
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// LU without pivoting (for U^T A V with random butterflies, rbt.hip).
// Panel of w <= 32 columns over the rows [j, n): every workgroup factors the
// w x w diagonal block in LDS (one wave, a row per lane), then solves its
// rows below, x U11 = a, one row per thread with the row in registers;
// workgroup 0 writes the block back.  No exchange between workgroups.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void genp_panel_kernel(float *__restrict__ P, long ldw, int mp,
                                                         int w, int *__restrict__ zero, int base)
{
    __shared__ float s_d[32][33];
    __shared__ float s_r[32];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wd = w < mp ? w : mp;              // order of the diagonal block
    for (int e = t; e < 32 * 32; e += 256) {
        const int r = e >> 5, c = e & 31;
        s_d[r][c] = (r < wd && c < w) ? P[(long)r * ldw + c] : (r == c ? 1.0f : 0.0f);
    }
    __syncthreads();
    if (wave == 0) {
        // right-looking elimination; lane r keeps row r of the block
        for (int k = 0; k < wd; ++k) {
            const float piv = s_d[k][k];
            if (lane > k && lane < wd) {
                const float l = s_d[lane][k] / piv;
                s_d[lane][k] = l;
                for (int c = k + 1; c < w; ++c) s_d[lane][c] = fmaf(-l, s_d[k][c], s_d[lane][c]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (lane < 32) s_r[lane] = 1.0f / s_d[lane][lane];
        if (blockIdx.x == 0) {
            const unsigned long long zb = __ballot(lane < wd && s_d[lane][lane] == 0.0f);
            if (lane == 0 && zb) atomicMin(zero, base + (int)__ffsll(zb) - 1);
        }
    }
    __syncthreads();
    const int i = wd + blockIdx.x * 256 + t;     // a row below the block
    if (i < mp) {
        float x[32];
        float *row = P + (long)i * ldw;
        if (w == 32) {
#pragma unroll
            for (int c = 0; c < 32; c += 4) {
                const float4_t v = *(const float4_t *)(row + c);
                x[c] = v[0]; x[c + 1] = v[1]; x[c + 2] = v[2]; x[c + 3] = v[3];
            }
        } else {
#pragma unroll
            for (int c = 0; c < 32; ++c) x[c] = c < w ? row[c] : 0.0f;
        }
        // x U11 = a: x_k = (a_k - sum_{c<k} x_c U_ck) / U_kk, right-looking
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const float xk = x[k] * s_r[k];
            x[k] = xk;
#pragma unroll
            for (int c = k + 1; c < 32; ++c) x[c] = fmaf(-xk, s_d[k][c], x[c]);
        }
        if (w == 32) {
#pragma unroll
            for (int c = 0; c < 32; c += 4) {
                float4_t v;
                v[0] = x[c]; v[1] = x[c + 1]; v[2] = x[c + 2]; v[3] = x[c + 3];
                *(float4_t *)(row + c) = v;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (c < w) row[c] = x[c];
        }
    }
    if (blockIdx.x == 0)
        for (int e = t; e < 32 * 32; e += 256) {
            const int r = e >> 5, c = e & 31;
            if (r < wd && c < w) P[(long)r * ldw + c] = s_d[r][c];
        }
}

hipError_t getrf_rm_genp(hipStream_t s, const MixWork &mw, float *W, long ldw, int n,
                         int prec, int nb)
{
    if (n <= 0) return hipSuccess;
    constexpr int ib = 32;
    if (nb < ib) nb = ib;
    if (nb > 256) nb = 256;
    nb = (nb / ib) * ib;
    // lookahead: the next block's columns are updated (in FP32) first, the
    // rest of the trailing update runs in prec on the side stream beside the
    // next block's panels
    const bool look = mw.side && mw.ev_a && mw.ev_r;
    bool rest_pending = false;
    for (int J = 0; J < n; J += nb) {
        const int jb = (n - J < nb) ? n - J : nb;
        for (int j = J; j < J + jb; j += ib) {
            const int w = (J + jb - j < ib) ? J + jb - j : ib;
            const int mp = n - j;
            const int wd = w < mp ? w : mp;
            {
                ProfScope ps(mw.prof, s, PROF_PANEL, 0.0, 8.0 * mp * w);
                const int rows = mp - wd;
                hipLaunchKernelGGL(genp_panel_kernel, dim3(rows > 0 ? (rows + 255) / 256 : 1),
                                   dim3(256), 0, s, W + (long)j * ldw + j, ldw, mp, w, mw.zero,
                                   j);
                FLX_CHECK(hipGetLastError());
            }
            const int cr = J + jb - (j + w);
            if (cr > 0) {
                FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)j * ldw + j, ldw,
                                             W + (long)j * ldw + j + w, ldw, w, cr));
                FLX_CHECK(lp_update(s, mw, MIX_FP32, W, ldw, j + w, n - j - w, j + w, cr, j, w));
            }
        }
        const int nr = n - (J + jb);
        if (nr <= 0) continue;
        if (!look) {
            FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)J * ldw + J, ldw,
                                         W + (long)J * ldw + J + jb, ldw, jb, nr));
            FLX_CHECK(lp_update(s, mw, prec, W, ldw, J + jb, nr, J + jb, nr, J, jb));
            continue;
        }
        // the previous block's rest update reaches the next block's columns
        if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, mw.ev_r, 0));
        rest_pending = false;
        const int nl = nr < nb ? nr : nb;
        FLX_CHECK(trsm_rm_llu<float>(s, mw.prof, W + (long)J * ldw + J, ldw,
                                     W + (long)J * ldw + J + jb, ldw, jb, nl));
        FLX_CHECK(lp_update(s, mw, MIX_FP32, W, ldw, J + jb, nr, J + jb, nl, J, jb));
        const int c0 = J + jb + nl;
        if (n > c0) {
            FLX_CHECK(hipEventRecord(mw.ev_a, s));
            FLX_CHECK(hipStreamWaitEvent(mw.side, mw.ev_a, 0));
            FLX_CHECK(trsm_rm_llu<float>(mw.side, mw.prof, W + (long)J * ldw + J, ldw,
                                         W + (long)J * ldw + c0, ldw, jb, n - c0,
                                         mw.side_groups));
            FLX_CHECK(lp_update(mw.side, mw, prec, W, ldw, J + jb, nr, c0, n - c0, J, jb,
                                mw.side_groups));
            FLX_CHECK(hipEventRecord(mw.ev_r, mw.side));
            rest_pending = true;
        }
    }
    if (rest_pending) FLX_CHECK(hipStreamWaitEvent(s, mw.ev_r, 0));
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// FP64 column-major A -> FP32 row-major W, W[i][j] = A(i,j) * dr[i] * dc[j]
// (scales optional).  flag[0] is set when a finite value exceeds FP32's
// range (the overflow test of LAPACK's dlag2s).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void conv_d2s_rm_kernel(const double *__restrict__ A, long lda,
                                                          int n, const double *__restrict__ dr,
                                                          const double *__restrict__ dc,
                                                          float *__restrict__ W, long ldw,
                                                          int *flag)
{
    __shared__ float t[32][33];
    const int bi = blockIdx.x * 32, bj = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    bool ovf = false;
    for (int q = ty; q < 32; q += 8) {
        const int i = bi + tx, j = bj + q;
        if (i < n && j < n) {
            double v = A[(long)j * lda + i];
            if (dr) v *= dr[i];
            if (dc) v *= dc[j];
            ovf |= fabs(v) > (double)FLT_MAX;
            t[q][tx] = (float)v;
        }
    }
    __syncthreads();
    for (int q = ty; q < 32; q += 8) {
        const int i = bi + q, j = bj + tx;
        if (i < n && j < n) W[(long)i * ldw + j] = t[tx][q];
    }
    if (ovf) atomicOr(flag, 1);
}

hipError_t conv_d2s_rm(hipStream_t s, const double *A, long lda, int n, const double *dr,
                       const double *dc, float *W, long ldw, int *flag)
{
    if (n <= 0) return hipSuccess;
    ProfScope ps(tl_prof, s, PROF_TRANSPOSE, 0.0, 12.0 * n * n);
    dim3 g((n + 31) / 32, (n + 31) / 32);
    hipLaunchKernelGGL(conv_d2s_rm_kernel, g, dim3(256), 0, s, A, lda, n, dr, dc, W, ldw, flag);
    return hipGetLastError();
}

} // namespace flx
