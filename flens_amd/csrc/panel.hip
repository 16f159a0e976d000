// panel.hip -- cooperative unblocked LU of a tall row-major panel (FP64).
//
// Replaces the unblocked panel flens::lapack::tf2 (flens/lapack/ge/tf2.tcc:
// 56-136): per column iamax (first maximum), row swap, scaling by the
// reciprocal of the pivot (division below the safe minimum), rank-1 update.
//
// One workgroup holds PANEL_ROWS_PER_WG consecutive panel rows in registers
// (256 threads x 2 rows, fully unrolled column loop so every register index
// is static); taller panels use up to PANEL_GROUPS_MAX cooperating
// workgroups.  Rows come in and go out through LDS, so every global access
// of a wave covers whole rows.  Per column:
//   1. workgroup-local pivot search (DPP wave reductions, one barrier),
//      first max on ties;
//   2. the wave that owns the workgroup's candidate posts it -- a header
//      (position, row) and the row's 32 elements -- as 65 sequence-tagged
//      64-bit words, one store per lane;
//   3. every workgroup reads all posts (each thread at most a few words,
//      stale words re-polled together), stages their payloads in LDS (one
//      barrier), and every wave picks the same winner and takes the pivot
//      row into scalar registers (readlane), then scales and applies the
//      rank-1 update to its own rows.
// An agent-scope store-to-load round trip between two workgroups costs about
// 0.9 us (same XCD) to 1.2 us (different XCDs) on MI355X, so a column costs
// at least that much: the design keeps one round trip and two barriers per
// column.  The posts are double-buffered by column parity: a workgroup can
// only overwrite a slot after every workgroup has posted the next column,
// which each does only after it finished reading the previous one.  Every
// wait is bounded; a failed wait sets the error flag and the kernel still
// exits.
#include <climits>
#include <cfloat>
#include <type_traits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

constexpr int PANEL_ROWS_MAX = LU_N_MAX;   // panel height limit
constexpr int PANEL_WMAX = 64;
constexpr int PANEL_W = 32;             // panel width of the cooperative kernel
// moved rows one workgroup can report (pivot rows + rows they displace)
constexpr int PANEL_LIST_PER_WG = 2 * PANEL_W;
static_assert(1 + 2 * PANEL_LIST_PER_WG <= PANEL_LIST_SLICE, "list slice too small");
constexpr uint32_t SPIN_LIMIT = 1u << 20;  // ~1 s of polling per word

// Exchange slots, in 64-bit words, per column parity: one slot of EX_SLOT
// words per workgroup -- words [0, 64) the candidate row (two words per
// element, low half first), [64, 66) the reciprocal of its leading element
// (low, high), word 66 the header (position << 16 | row, or EX_NONE when
// the workgroup has no live row).  Positions and rows are below
// LU_N_MAX = 2^15.
constexpr int EX_GMAX = PANEL_ROWS_MAX / 512;   // room for the smallest workgroup
constexpr int EX_ROWW = 2 * PANEL_W;            // row words of one post
constexpr int EX_INV = EX_ROWW;                 // reciprocal of element 0
constexpr int EX_HDR = EX_ROWW + 2;             // header
constexpr int EX_POST = EX_ROWW + 3;            // words one workgroup posts per column
constexpr size_t EX_SLOT = 128;
constexpr size_t EX_PAR = (size_t)EX_GMAX * EX_SLOT;
constexpr uint32_t EX_NONE = 0xffffffffu;
static_assert(PANEL_ROWS_MAX <= 32768, "positions and rows are packed in 16 bits");

size_t lu_exch_words() { return 2 * EX_PAR; }

typedef double dvecW __attribute__((ext_vector_type(PANEL_W)));


// Pivot candidates are compared as (key, pos): key = bits(|v|) + 1 for a
// live row (the bit pattern of a non-negative double orders like the value,
// +1 keeps |v| = 0 above "no candidate" = 0), larger key wins, equal keys
// go to the lower position -- the first maximum of
// cxxblas/level1/iamax.tcc:42-58 in the interchanged order.
struct Cand {
    uint64_t key;
    int pos, rid;
};
// The key of a live row at position pos in column j.  The reference's iamax
// (cxxblas/level1/iamax.tcc:42-58) starts from the first entry's |x| and
// replaces it only on a strictly larger one, so a NaN is chosen only at the
// first live position (where it then stays: every key below ~0) and is
// skipped anywhere else (key 0 loses to every live number, |v| = 0 included).
FLX_DEV uint64_t pivot_key(double v, int pos, int j)
{
    const double a = fabs(v);
    const uint64_t k = (uint64_t)__double_as_longlong(a) + 1;
    const uint64_t nan_key = pos == j ? ~0ull : 0ull;
    return __builtin_isnan(a) ? nan_key : k;
}
FLX_DEV bool cand_gt(uint64_t ka, int pa, uint64_t kb, int pb)
{
    return ka > kb || (ka == kb && pa < pb);
}
// wave-wide best candidate; uniform result
FLX_DEV Cand wave_best(uint64_t key, int pos, int rid)
{
    const uint32_t hi = (uint32_t)(key >> 32), lo = (uint32_t)key;
    const uint32_t mhi = wave_umax(hi);
    const uint32_t mlo = wave_umax(hi == mhi ? lo : 0u);
    const bool top = hi == mhi && lo == mlo;
    const uint32_t mpos = wave_umin(top ? (uint32_t)pos : 0xffffffffu);
    const unsigned long long who = __ballot(top && (uint32_t)pos == mpos);
    Cand c;
    c.key = ((uint64_t)mhi << 32) | mlo;
    c.pos = (int)mpos;
    c.rid = who ? __builtin_amdgcn_readlane(rid, (int)(__ffsll(who) - 1)) : -1;
    return c;
}

// Largest 32-bit value over the first SPAN lanes (4, 16 or 64), uniform:
// xor-1 and xor-2 quad permutations, then row rotations by 4 and 8 give
// every lane the maximum of its 16-lane row; four readlanes join the rows.
template <int SPAN>
FLX_DEV uint32_t span_umax(uint32_t v)
{
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false));   // quad_perm 1,0,3,2
    v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false));   // quad_perm 2,3,0,1
    if constexpr (SPAN > 4) {
        v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xf, 0xf, false)); // row_ror 4
        v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xf, 0xf, false)); // row_ror 8
    }
    uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
    if constexpr (SPAN > 16) {
        m = max(m, (uint32_t)__builtin_amdgcn_readlane((int)v, 16));
        m = max(m, (uint32_t)__builtin_amdgcn_readlane((int)v, 32));
        m = max(m, (uint32_t)__builtin_amdgcn_readlane((int)v, 48));
    }
    return m;
}
// wave_best for candidates held by the first SPAN lanes (keys 0 beyond):
// when one lane alone holds the largest high key word -- all but exact or
// near-exact ties -- it is the winner; otherwise the exact wave_best.
template <int SPAN>
FLX_DEV Cand span_best(uint64_t key, int pos, int rid)
{
    const uint32_t hi = (uint32_t)(key >> 32);
    const uint32_t m = span_umax<SPAN>(hi);
    const unsigned long long top = __ballot(hi == m);
    if (__popcll(top) == 1) {
        const int L = (int)__ffsll(top) - 1;
        Cand c;
        c.key = ((uint64_t)m << 32) | (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, L);
        c.pos = __builtin_amdgcn_readlane(pos, L);
        c.rid = __builtin_amdgcn_readlane(rid, L);
        return c;
    }
    return wave_best(key, pos, rid);
}

// v as a wave-uniform (scalar) value; v must be the same in every lane
FLX_DEV double uniform_f64(double v)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// the double whose low and high halves are the two 32-bit LDS words at p
// (8-byte aligned)
FLX_DEV double lds_f64(const uint32_t *p)
{
    double d;
    __builtin_memcpy(&d, p, sizeof d);
    return d;
}

// the two 32-bit halves of a double, as a uniform value from lane src
FLX_DEV double read_lane_f64(double v, int src)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int NT, int RPT, int GMAX>
__global__ __launch_bounds__(NT) void panel_kernel(
    double *__restrict__ P, long ldw, int mp, int w, uint64_t *__restrict__ ex,
    uint32_t gen, double *__restrict__ lsc, int *__restrict__ piv, int piv_base,
    int *__restrict__ list, int *__restrict__ info, int info_base, int *__restrict__ err,
    int defer_l)
{
    constexpr int W = PANEL_W;
    constexpr int NWAVE = NT / 64;
    constexpr int ROWS = NT * RPT;                 // panel rows per workgroup
    static_assert(GMAX * ROWS <= PANEL_ROWS_MAX && GMAX <= EX_GMAX && GMAX <= 64, "groups");
    static_assert(W == 32 && EX_ROWW == 64 && NWAVE <= 4, "one posted word per lane");
    // words one workgroup reads per column at most, and per thread
    constexpr int WORDS_MAX = EX_POST * GMAX;
    constexpr int PER = (WORDS_MAX + NT - 1) / NT;
    constexpr int LS = EX_POST + 1;                // LDS stride of a post (8-byte aligned)
    constexpr int GSPAN = GMAX <= 4 ? 4 : GMAX <= 16 ? 16 : 64;
    // row staging through LDS, IO_R rows at a time: in, rows of W doubles
    // at stride IO_LD (16-byte aligned pairs), IO_PER pairs per thread per
    // chunk; out, rows of 2 W doubles (the multipliers, then the registers
    // from the row's start column) at the odd stride OUT_LD
    constexpr int IO_R = 128, IO_LD = W + 2, OUT_LD = 2 * W + 1;
    constexpr int IO_CHUNKS = RPT * NT / IO_R;
    constexpr int IO_PER = IO_R * (W / 2) / NT;
    static_assert(NT % IO_R == 0 && IO_R * (W / 2) % NT == 0, "row staging");

    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mn = mp < w ? mp : w;

    // posts of one column, by column parity: element c of a post's row at
    // [2c] (low half) and [2c + 1] (high half), its reciprocal at EX_INV,
    // the header at EX_HDR
    __shared__ __align__(16) uint32_t s_words[2][LS * GMAX];
    __shared__ __align__(16) double s_u[NWAVE][2][W];   // per wave: shifted pivot rows
    __shared__ __align__(16) double s_own[W + 2];        // the candidate row, multiplier, 1/pivot
    __shared__ uint64_t s_wkey[NWAVE];
    __shared__ int s_wpos[NWAVE], s_wrid[NWAVE];
    __shared__ int s_piv[W];
    __shared__ int s_cnt;
    __shared__ int s_fail;
    __shared__ __align__(16) double s_io[IO_R * OUT_LD];   // rows in / out, IO_R at a time
    __shared__ int s_ipos[IO_R];

    critical_path_priority();
    // Thread tid owns original panel rows g*ROWS + tid + k*NT.  Rows never
    // move during the panel: pos[k] is the row's current position under the
    // interchanges so far (LAPACK's row index), done[k] marks pivot rows.
    // A live row is kept rotated: element e is column j + e at column j;
    // its multipliers (the finished columns, L) go to the scratch lsc
    // (column-major, leading dimension mp) as they are made.  A pivot row
    // is never touched again after its column: element e is column pos + e.
    //
    // The rank-1 update of column j is split.  Right after the pivot is
    // known only element 0 of the live rows is updated (column j + 1, the
    // next pivot search's input); elements 1..W-2 follow one step later,
    // while the next column's posts are in flight.  In between a live row
    // is "pending": elements 2..W-1 still hold the old values, lm[k] its
    // multiplier.  A candidate row is posted in that form (element 0, old
    // elements 2..W-1, multiplier), and readers finish it with the previous
    // pivot row, which every workgroup holds: the same operation on the
    // same operands as the owner's own update, so every copy is
    // bit-identical.  The live part of a row takes fma(-l, u, a) whatever
    // l is (as the reference's rank-1 update of the rows below the pivot,
    // tf2.tcc:130-132, does for a non-finite multiplier too).
    int rid[RPT], pos[RPT];
    bool done[RPT], any[RPT];
    dvecW row[RPT];
    double lm[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        rid[k] = g * ROWS + tid + k * NT;
        pos[k] = rid[k];
        done[k] = rid[k] >= mp;
        any[k] = __ballot(rid[k] < mp) != 0;   // the wave has a row in slot k
        lm[k] = 0.0;
    }
    if (tid == 0) { s_fail = 0; s_cnt = 0; }
    int first_zero = 0;
    // The rows arrive through LDS: sixteen lanes load one row as 16-byte
    // pairs (a wave reads four whole rows per load, not 64 scattered ones),
    // every chunk's loads issued at once; then each thread takes its row
    // from LDS.
    double2 v[IO_CHUNKS][IO_PER];
#pragma unroll
    for (int q = 0; q < IO_CHUNKS; ++q) {
        const int k = q / (NT / IO_R), h = q % (NT / IO_R);
        const int rbase = g * ROWS + k * NT + h * IO_R;
#pragma unroll
        for (int i = 0; i < IO_PER; ++i) {
            const int p = tid + NT * i, r = p >> 4, c2 = (p & 15) * 2, gr = rbase + r;
            v[q][i] = double2(0.0, 0.0);
            if (gr < mp && c2 < w) {
                const double *src = P + (long)gr * ldw + c2;
                if (c2 + 1 < w) v[q][i] = *(const double2 *)src;
                else v[q][i].x = src[0];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < IO_CHUNKS; ++q) {
        const int k = q / (NT / IO_R), h = q % (NT / IO_R);
#pragma unroll
        for (int i = 0; i < IO_PER; ++i) {
            const int p = tid + NT * i;
            *(double2 *)&s_io[(p >> 4) * IO_LD + (p & 15) * 2] = v[q][i];
        }
        __syncthreads();
        if (tid / IO_R == h) {
#pragma unroll
            for (int c = 0; c < W; c += 2) {
                const double2 e = *(const double2 *)&s_io[(tid % IO_R) * IO_LD + c];
                row[k][c] = e.x;
                row[k][c + 1] = e.y;
            }
        }
        __syncthreads();
    }

    // The pivot row of column jj (the winner of group wg, posted pending
    // when jj > 0): element c in lane c, finished with the pivot row of
    // column jj - 1 (u_prev); writes this wave's shifted copy u (finished
    // columns 0).
    auto pivot_row = [&](int jj, int wg, const double *u_prev, double *u) {
        const uint32_t *post = &s_words[jj & 1][wg * LS];
        double e = 0.0;
        if (lane < W) {
            const double p = lds_f64(post + 2 * lane);
            e = p;
            if (jj > 0 && lane > 0) e = fma(-lds_f64(post + 2 * (W - 1)), u_prev[lane], p);
        }
        const int lv = W - jj;
        if (lane > 0 && lane < W) u[lane - 1] = lane < lv ? e : 0.0;
        if (lane == 0) u[W - 1] = 0.0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    // the deferred part of a column's update on the live rows, with its
    // pivot row u
    auto finish_update = [&](const double *u) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!any[k]) continue;
            if (!done[k]) {
                const double m = -lm[k];
#pragma unroll
                for (int c0 = 1; c0 < W - 1; c0 += 8) {
                    double uu[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) uu[q] = c0 + q < W - 1 ? u[c0 + q] : 0.0;
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        if (c0 + q < W - 1) row[k][c0 + q] = fma(m, uu[q], row[k][c0 + q + 1]);
                }
            }
        }
    };

    int wg_prev = 0;
    for (int j = 0; j < mn; ++j) {
        const int live = W - j;                  // live elements of a rotated row
        const int par = j & 1;
        // ---- 1. workgroup-local pivot candidate ------------------------------
        uint64_t key = 0;
        int kp = INT_MAX, kr = -1;
        double kv = 1.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const uint64_t kk = done[k] ? 0 : pivot_key(row[k][0], pos[k], j);
            if (cand_gt(kk, done[k] ? INT_MAX : pos[k], key, kp)) {
                key = kk; kp = pos[k]; kr = rid[k]; kv = row[k][0];
            }
        }
        // every lane's candidate reciprocal, ready when its row is posted
        const double kinv = 1.0 / kv;
        const Cand wb = span_best<64>(key, kp, kr);
        if (lane == 0) { s_wkey[wave] = wb.key; s_wpos[wave] = wb.pos; s_wrid[wave] = wb.rid; }
        __syncthreads();
        {
            const bool q = lane < NWAVE;
            const Cand c = span_best<4>(q ? s_wkey[lane] : 0ull, q ? s_wpos[lane] : INT_MAX,
                                        q ? s_wrid[lane] : -1);
            const int br = c.key == 0 ? -1 : c.rid;   // -1: no live row in this workgroup
            // ---- 2. the candidate row in posted form, one word per lane -----
            const int otid = br >= 0 ? (br - g * ROWS) % NT : 0;
            if (wave == (otid >> 6)) {
                if (lane == (otid & 63) && br >= 0) {
#pragma unroll
                    for (int k = 0; k < RPT; ++k)
                        if (rid[k] == br) {
                            *(dvecW *)s_own = row[k];
                            s_own[W] = lm[k];
                        }
                    s_own[W + 1] = kinv;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // lane l: half l & 1 of posted element e = l >> 1 (pending
                // form: element 0, old elements 2..W-1, the multiplier);
                // lanes 0, 1: the reciprocal; lane 2: the header
                const int e = lane >> 1;
                const int src = j == 0 ? e : (e == 0 ? 0 : (e < W - 1 ? e + 1 : W));
                const uint64_t bits = (uint64_t)__double_as_longlong(s_own[src]);
                const uint32_t half = (lane & 1) ? (uint32_t)(bits >> 32) : (uint32_t)bits;
                const uint64_t ibits = (uint64_t)__double_as_longlong(s_own[W + 1]);
                const uint32_t hdr = br >= 0 ? (((uint32_t)c.pos << 16) | (uint32_t)br) : EX_NONE;
                const uint32_t extra = lane == 0 ? (uint32_t)ibits
                                     : lane == 1 ? (uint32_t)(ibits >> 32) : hdr;
                if (G == 1) {
                    s_words[par][lane] = half;
                    if (lane < 3) s_words[par][EX_INV + lane] = extra;
                } else {
                    const uint64_t tag = (uint64_t)((gen << 8) | (uint32_t)(j + 1)) << 32;
                    uint64_t *slot = ex + (size_t)par * EX_PAR + (size_t)g * EX_SLOT;
                    st_agent(slot + lane, tag | half);
                    if (lane < 3) st_agent(slot + EX_INV + lane, tag | extra);
                }
            }
        }
        // ---- 3. the previous column's update, while the posts travel -------
        if (j > 0) {
            pivot_row(j - 1, wg_prev, s_u[wave][par], s_u[wave][par ^ 1]);
            finish_update(s_u[wave][par ^ 1]);
        }
        // ---- 4. gather every workgroup's post ------------------------------
        if (G > 1) {
            const uint32_t seq = (gen << 8) | (uint32_t)(j + 1);
            const uint64_t *base = ex + (size_t)par * EX_PAR;
            const int total = EX_POST * G;
            // every word this thread reads is requested at once; stale ones
            // are requested again together until every tag is the column's
            uint64_t v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = tid + k * NT;
                const int q = idx < total ? idx : 0;
                v[k] = idx < total ? ld_agent(base + (size_t)(q / EX_POST) * EX_SLOT + q % EX_POST)
                                   : ((uint64_t)seq << 32);
            }
            uint32_t spins = 0;
            for (;;) {
                bool stale = false;
#pragma unroll
                for (int k = 0; k < PER; ++k) stale = stale || seq_of(v[k]) != seq;
                if (!stale) break;
                // give up for good once any wait of this workgroup failed
                if (*(volatile int *)&s_fail || ++spins > SPIN_LIMIT) {
                    s_fail = 1;
                    break;
                }
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const int idx = tid + k * NT;
                    const int q = idx < total ? idx : 0;
                    if (seq_of(v[k]) != seq)
                        v[k] = ld_agent(base + (size_t)(q / EX_POST) * EX_SLOT + q % EX_POST);
                }
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = tid + k * NT;
                if (idx < total) s_words[par][(idx / EX_POST) * LS + idx % EX_POST] = (uint32_t)v[k];
            }
        }
        __syncthreads();
        // ---- 5. every wave picks the winner (lane q evaluates workgroup q) -
        // (each lane also reads its group's reciprocal, element 1 and
        // multiplier, so the winner's are a readlane away)
        const int qg = lane < G ? lane : 0;
        const uint32_t *qpost = &s_words[par][qg * LS];
        const uint32_t qh = lane < G ? qpost[EX_HDR] : EX_NONE;
        const double qv = lds_f64(qpost);
        const double qinv = lds_f64(qpost + EX_INV);
        const double qp1 = lds_f64(qpost + 2);
        const double qlw = lds_f64(qpost + 2 * (W - 1));
        const int qp = qh != EX_NONE ? (int)(qh >> 16) : INT_MAX;
        const uint64_t qk = qh != EX_NONE ? pivot_key(qv, qp, j) : 0ull;
        const Cand wc = span_best<GSPAN>(qk, qp, lane);   // rid slot carries the group
        int wg = wc.rid;
        const int wl = wg >= 0 && wg < 64 ? wg : 0;
        const uint32_t wh = (uint32_t)__builtin_amdgcn_readlane((int)qh, wl);
        int wpos = wc.pos;
        int wrid = (int)(wh & 0xffffu);
        // a winner outside the panel can only come from a failed wait: keep
        // every later index in bounds and report it
        const bool bad = wc.key == 0 || wg < 0 || wg >= G || wh == EX_NONE || wpos < j ||
                         wpos >= mp || wrid >= mp;
        if (bad) {
            if (lane == 0) s_fail = 1;
            wg = 0;
            wpos = j;
            wrid = -1;
        }
        wg_prev = wg;
        // the pivot, its reciprocal and the winner's element 1 (finished
        // with the previous pivot row as the owner would), as scalars
        const double pv = read_lane_f64(qv, wg);
        const double inv = read_lane_f64(qinv, wg);
        const double p1 = read_lane_f64(qp1, wg);
        const double lw = read_lane_f64(qlw, wg);
        const double up1 = uniform_f64(s_u[wave][par ^ 1][1]);
        const double u0 = 1 < live ? (j > 0 ? fma(-lw, up1, p1) : p1) : 0.0;

        // ---- 6. interchange bookkeeping, multipliers, element 0 ------------
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (rid[k] == wrid) { pos[k] = j; done[k] = true; }
            else if (pos[k] == j) pos[k] = wpos;
        }
        const bool zero = pv == 0.0;
        const bool tiny = fabs(pv) < DBL_MIN;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (!any[k]) continue;
            if (!done[k]) {
                // the multiplier (tf2.tcc:112-121: times the reciprocal,
                // divided below the safe minimum; a zero pivot scales nothing)
                double l = row[k][0];
                if (!zero) {
                    if (tiny) l = l / pv;
                    else l = l * inv;
                }
                lm[k] = l;
                lsc[(long)j * mp + rid[k]] = l;
                row[k][0] = fma(-l, u0, row[k][1]);
            }
        }
        if (tid == 0) {
            s_piv[j] = wpos;
            if (zero && first_zero == 0) first_zero = j + 1;
        }
        // s_wkey / s_own / s_words[par] are rewritten only after later
        // barriers, which every thread reaches after its reads above
    }
    if (mn > 0) {
        const int jl = mn - 1, pl = jl & 1;
        pivot_row(jl, wg_prev, s_u[wave][pl ^ 1], s_u[wave][pl]);
        finish_update(s_u[wave][pl]);
    }

    // write every row to its final position: columns below `start` are
    // its multipliers (from lsc), element e of the registers is column
    // start + e (a pivot row: start = its column; a row left below the
    // panel's pivots: start = mn); record the moved rows
    int *mylist = list + (size_t)g * PANEL_LIST_SLICE;
    if (defer_l) {
        // Only the pivot rows (final positions below mn) are written here;
        // every other row of the panel is its mn multipliers, which the
        // caller's next kernel copies from lsc (column-major, by original
        // row) into the row's final position.  A pivot row's multipliers
        // are requested at once, then stored with the row's registers.
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (rid[k] < mp && done[k]) {
                double pm[W];
#pragma unroll
                for (int c = 0; c < W; ++c) pm[c] = lsc[(long)c * mp + rid[k]];
                double *dst = P + (long)pos[k] * ldw;
                const int start = pos[k];
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < start) dst[c] = pm[c];
#pragma unroll
                for (int e = 0; e < W; ++e)
                    if (start + e < w) dst[start + e] = row[k][e];
            }
        }
    } else {
        // the multipliers are loaded first, all at once (unconditional, in
        // bounds: lsc holds W columns of mp rows), so the epilogue waits for
        // one memory latency instead of one per column
        double mv[RPT][W];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const long r = rid[k] < mp ? rid[k] : 0;
#pragma unroll
            for (int c = 0; c < W; ++c) mv[k][c] = lsc[(long)c * mp + r];
        }
        // the final rows leave through LDS as they came in: each thread stages
        // its row, sixteen lanes store one row as 16-byte pairs
#pragma unroll
        for (int q = 0; q < IO_CHUNKS; ++q) {
            const int k = q / (NT / IO_R), h = q % (NT / IO_R);
            if (tid / IO_R == h) {
                // multipliers at [0, W), then the registers at [start, start + W)
                // over them: unconditional stores, the later ones win
                // (start <= W: a pivot row's column, mn, or 0 for a slot
                // beyond the panel, whose row is never stored)
                double *st = s_io + (tid % IO_R) * OUT_LD;
                const int start = rid[k] >= mp ? 0 : done[k] ? pos[k] : mn;
#pragma unroll
                for (int c = 0; c < W; ++c) st[c] = mv[k][c];
#pragma unroll
                for (int e = 0; e < W; ++e) st[start + e] = row[k][e];
                s_ipos[tid % IO_R] = rid[k] < mp ? pos[k] : -1;
            }
            __syncthreads();
            // every LDS read first (unconditional, in the staged block), then
            // the stores: one LDS latency per chunk, not one per pair
            int dps[IO_PER];
            double2 pv[IO_PER];
#pragma unroll
            for (int i = 0; i < IO_PER; ++i) {
                const int p = tid + NT * i, r = p >> 4, c2 = (p & 15) * 2;
                dps[i] = s_ipos[r];
                pv[i] = double2(s_io[r * OUT_LD + c2], s_io[r * OUT_LD + c2 + 1]);
            }
#pragma unroll
            for (int i = 0; i < IO_PER; ++i) {
                const int p = tid + NT * i, c2 = (p & 15) * 2;
                if (dps[i] >= 0 && c2 < w) {
                    double *dst = P + (long)dps[i] * ldw + c2;
                    if (c2 + 1 < w) *(double2 *)dst = pv[i];
                    else dst[0] = pv[i].x;
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        if (rid[k] < mp) {
            if (pos[k] != rid[k]) {
                const int e = atomicAdd(&s_cnt, 1);
                if (e < PANEL_LIST_PER_WG) {
                    mylist[1 + 2 * e] = piv_base + pos[k];
                    mylist[2 + 2 * e] = piv_base + rid[k];
                }
            }
        }
    }
    __syncthreads();
    if (g == 0)
        for (int q = tid; q < mn; q += NT) piv[q] = piv_base + s_piv[q];
    if (tid == 0) {
        mylist[0] = s_cnt < PANEL_LIST_PER_WG ? s_cnt : PANEL_LIST_PER_WG;
        if (s_cnt > PANEL_LIST_PER_WG) s_fail = 1;
        if (g == 0 && first_zero) atomicMin(info, info_base + first_zero);
        if (s_fail) atomicExch(err, 1);
    }
}

__global__ void zero_u64_kernel(uint64_t *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_panel(hipStream_t s, LuWork &wk, double *P, long ldw,
                               int mp, int w, int *piv, int piv_base, int *info,
                               int ib, int defer_l)
{
    if (mp > PANEL_ROWS_MAX || !wk.lsc) return hipErrorInvalidValue;
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
    // 256 threads (one wave per SIMD) x 2 rows per thread: 512 rows per
    // workgroup, at most LU_N_MAX / 512 cooperating workgroups
    (void)ib;
    constexpr int NT = 256, RPT = 2;
    static_assert(NT * RPT == PANEL_ROWS_PER_WG, "panel geometry");
    // the words a workgroup gathers per column scale with the group count
    const int G = panel_groups(mp);
    if (G <= 8)
        hipLaunchKernelGGL((panel_kernel<NT, RPT, 8>), dim3(G), dim3(NT), 0, s, P, ldw, mp, w,
                           wk.exch, wk.gen, wk.lsc, piv, piv_base, wk.list, info, piv_base, wk.flags, defer_l);
    else if (G <= 16)
        hipLaunchKernelGGL((panel_kernel<NT, RPT, 16>), dim3(G), dim3(NT), 0, s, P, ldw, mp, w,
                           wk.exch, wk.gen, wk.lsc, piv, piv_base, wk.list, info, piv_base, wk.flags, defer_l);
    else if (G <= 32)
        hipLaunchKernelGGL((panel_kernel<NT, RPT, 32>), dim3(G), dim3(NT), 0, s, P, ldw, mp, w,
                           wk.exch, wk.gen, wk.lsc, piv, piv_base, wk.list, info, piv_base, wk.flags, defer_l);
    else
        hipLaunchKernelGGL((panel_kernel<NT, RPT, LU_N_MAX / PANEL_ROWS_PER_WG>), dim3(G),
                           dim3(NT), 0, s, P, ldw, mp, w, wk.exch, wk.gen, wk.lsc, piv,
                           piv_base, wk.list, info, piv_base, wk.flags, defer_l);
    return hipGetLastError();
}

} // namespace flx
