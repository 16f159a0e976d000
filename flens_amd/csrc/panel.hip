// panel.hip -- cooperative unblocked LU of a tall row-major panel (FP64).
//
// Replaces the unblocked panel flens::lapack::tf2 (flens/lapack/ge/tf2.tcc:
// 56-136): per column iamax (first maximum), row swap, scaling by the
// reciprocal of the pivot (division below the safe minimum), rank-1 update.
//
// One workgroup holds 512 consecutive panel rows in registers (256 threads x
// 2 rows, fully unrolled column loop so every register index is static);
// taller panels use up to 32 cooperating workgroups.  Per column:
//   1. workgroup-local pivot search (shuffles + LDS), first max on ties;
//   2. every workgroup posts its best candidate (|v|, row, the row's data)
//      and the owner of row j posts row j, as sequence-tagged 64-bit words;
//   3. every workgroup reads all posts (one agent-scope round trip, no
//      fence, no counter), picks the same winner, swaps, scales and applies
//      the rank-1 update to its own rows.
// The posts are double-buffered by column parity: a workgroup can only
// overwrite a slot after every workgroup has posted the next column, which
// each does only after it finished reading the previous one.  Every wait is
// bounded; a failed wait sets the error flag and the kernel still exits.
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

// Exchange slots, in 64-bit words, per column parity, sized for the smallest
// workgroup (256 threads -> at most 64 workgroups):
//   headers [64][4] | candidate rows [64][2*WMAX] | diagonal row [2*WMAX]
constexpr int EX_GMAX = PANEL_ROWS_MAX / 256;
constexpr size_t EX_HDR = 0;
constexpr size_t EX_ROWS = EX_GMAX * 4;
constexpr size_t EX_DIAG = EX_ROWS + (size_t)EX_GMAX * 2 * PANEL_WMAX;
constexpr size_t EX_PAR = EX_DIAG + 2 * PANEL_WMAX;

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
    if (a != a) return pos == j ? ~0ull : 0ull;
    return (uint64_t)__double_as_longlong(a) + 1;
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

template <int NT, int RPT, int GMAX>
__global__ __launch_bounds__(NT) void panel_kernel(
    double *__restrict__ P, long ldw, int mp, int w, uint64_t *__restrict__ ex,
    uint32_t gen, int *__restrict__ piv, int piv_base, int *__restrict__ list,
    int *__restrict__ info, int info_base, int *__restrict__ err)
{
    constexpr int W = PANEL_W;
    constexpr int NWAVE = NT / 64;
    constexpr int ROWS = NT * RPT;                 // panel rows per workgroup
    static_assert(GMAX <= PANEL_ROWS_MAX / ROWS, "groups");
    // words one workgroup reads per column at most: G headers (4 words) and
    // the live part of G candidate rows (2 words per element)
    constexpr int WORDS_MAX = 4 * GMAX + 2 * W * GMAX;
    constexpr int PER = (WORDS_MAX + NT - 1) / NT;

    const int G = gridDim.x;
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int mn = mp < w ? mp : w;

    __shared__ uint64_t s_raw[WORDS_MAX];     // posted words as first read
    __shared__ uint32_t s_words[WORDS_MAX];   // low halves of the posted words
    __shared__ double s_own[W];                // this workgroup's candidate row
    __shared__ double s_u[W];                  // pivot row, shifted, finished cols = 0
    __shared__ uint64_t s_wkey[NWAVE];
    __shared__ int s_wpos[NWAVE], s_wrid[NWAVE];
    __shared__ int s_win[3];                   // pos, rid, source workgroup
    __shared__ double s_pv;
    __shared__ int s_ufin;                     // 1: every entry of s_u is finite
    __shared__ int s_piv[W];
    __shared__ int s_cnt;
    __shared__ int s_fail;

    critical_path_priority();
    // Thread tid owns original panel rows g*ROWS + tid + k*NT.  Rows never
    // move during the panel: pos[k] is the row's current position under the
    // interchanges so far (LAPACK's row index), done[k] marks pivot rows.
    // Each row is kept rotated: element 0 is always the current column j,
    // elements W-j.. hold the finished columns 0..j-1.
    int rid[RPT], pos[RPT];
    bool done[RPT];
    dvecW row[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        rid[k] = g * ROWS + tid + k * NT;
        pos[k] = rid[k];
        done[k] = rid[k] >= mp;
        const double *src = P + (long)(rid[k] < mp ? rid[k] : 0) * ldw;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            double v = 0.0;
            if (rid[k] < mp && c < w) v = src[c];
            row[k][c] = v;
        }
    }
    if (tid == 0) { s_fail = 0; s_cnt = 0; }
    int first_zero = 0;
    __syncthreads();

    for (int j = 0; j < mn; ++j) {
        const int live = W - j;                  // live elements of a rotated row
        // ---- 1. local pivot candidate ------------------------------------
        uint64_t key = 0;
        int kp = INT_MAX, kr = -1;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const uint64_t kk = done[k] ? 0 : pivot_key(row[k][0], pos[k], j);
            if (cand_gt(kk, done[k] ? INT_MAX : pos[k], key, kp)) { key = kk; kp = pos[k]; kr = rid[k]; }
        }
        const Cand wb = wave_best(key, kp, kr);
        if (lane == 0) { s_wkey[wave] = wb.key; s_wpos[wave] = wb.pos; s_wrid[wave] = wb.rid; }
        __syncthreads();
        uint64_t bk = s_wkey[0];
        int bp = s_wpos[0], br = s_wrid[0];
#pragma unroll
        for (int q = 1; q < NWAVE; ++q)
            if (cand_gt(s_wkey[q], s_wpos[q], bk, bp)) { bk = s_wkey[q]; bp = s_wpos[q]; br = s_wrid[q]; }

        int wpos = bp, wrid = br;
        if (G == 1) {
            // single workgroup: the winner's owner fills the shifted pivot row
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (rid[k] == br) {
                    s_pv = row[k][0];
                    bool fin = true;
#pragma unroll
                    for (int c = 0; c < W - 1; ++c) {
                        const double uc = (c + 1 < live) ? row[k][c + 1] : 0.0;
                        fin = fin && __builtin_isfinite(uc);
                        s_u[c] = uc;
                    }
                    s_u[W - 1] = 0.0;
                    s_ufin = fin ? 1 : 0;
                }
            __syncthreads();
        } else {
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (rid[k] == br) {
#pragma unroll
                    for (int c = 0; c < W; ++c) s_own[c] = row[k][c];
                }
            __syncthreads();
            // ---- 2./3. all-gather: header + live part of the candidate row -
            const uint32_t seq = (gen << 8) | (uint32_t)(j + 1);
            uint64_t *base = ex + (size_t)(j & 1) * EX_PAR;
            const uint64_t tag = (uint64_t)seq << 32;
            if (tid < 2 * live) {
                const int c = tid >> 1;
                uint64_t b = (uint64_t)__double_as_longlong(s_own[c]);
                st_agent(base + EX_ROWS + (size_t)g * 2 * PANEL_WMAX + tid,
                         tag | ((tid & 1) ? (b & 0xffffffffull) : (b >> 32)));
            } else if (tid >= NT - 4) {
                const int q = tid - (NT - 4);
                uint64_t pay = q == 0 ? (bk >> 32)
                             : q == 1 ? (bk & 0xffffffffull)
                             : q == 2 ? (uint64_t)(uint32_t)bp : (uint64_t)(uint32_t)br;
                st_agent(base + EX_HDR + g * 4 + q, tag | pay);
            }
            // words: [0, 4G) headers, then G rows of 2*live words each
            const int nhdr = 4 * G, rw = 2 * live, total = nhdr + G * rw;
            auto word_addr = [&](int idx) -> size_t {
                if (idx < nhdr) return EX_HDR + idx;
                const int q = idx - nhdr;
                return EX_ROWS + (size_t)(q / rw) * 2 * PANEL_WMAX + q % rw;
            };
            // every word this thread reads is requested at once (one round
            // trip) and staged in LDS; one runtime loop re-polls stale tags
            uint64_t v[PER];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = tid + k * NT;
                v[k] = idx < total ? ld_agent(base + word_addr(idx)) : tag;
            }
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const int idx = tid + k * NT;
                if (idx < total) s_raw[idx] = v[k];
            }
            for (int idx = tid; idx < total; idx += NT) {
                uint64_t x = s_raw[idx];
                uint32_t spins = 0;
                while (seq_of(x) != seq) {
                    // give up for good once any wait of this workgroup failed
                    if (*(volatile int *)&s_fail || ++spins > SPIN_LIMIT) {
                        s_fail = 1;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    x = ld_agent(base + word_addr(idx));
                }
                s_words[idx] = (uint32_t)(x & 0xffffffffull);
            }
            __syncthreads();
            if (wave == 0) {
                uint64_t qk = 0;
                int qp = INT_MAX;
                if (lane < G) {
                    qk = ((uint64_t)s_words[4 * lane] << 32) | s_words[4 * lane + 1];
                    qp = (int)s_words[4 * lane + 2];
                }
                const Cand wc = wave_best(qk, qp, lane);   // rid slot carries the group
                const int wg = wc.rid;
                int wr = (wg >= 0 && wg < G) ? (int)s_words[4 * wg + 3] : -1;
                int wp = wc.pos;
                // a winner outside the panel can only come from a failed wait:
                // keep every later index in bounds and report it
                const bool bad = wp < j || wp >= mp || wr < 0 || wr >= mp || wg < 0 || wg >= G;
                if (lane == 0) {
                    if (bad) s_fail = 1;
                    s_win[0] = bad ? j : wp;
                    s_win[1] = bad ? -1 : wr;
                }
                // the winner's row, shifted by one, finished columns zeroed
                const int src = nhdr + (bad ? 0 : wg) * rw;
                double uc = 0.0;
                if (lane < W) {
                    if (lane < live)
                        uc = __longlong_as_double((long long)(((uint64_t)s_words[src + 2 * lane] << 32) |
                                                              (uint64_t)s_words[src + 2 * lane + 1]));
                    if (lane == 0) s_pv = uc;
                    else s_u[lane - 1] = uc;
                    if (lane == W - 1) s_u[W - 1] = 0.0;
                }
                // the pivot itself (lane 0) is not part of u
                const bool nonfin = lane > 0 && lane < W && !__builtin_isfinite(uc);
                const unsigned long long bad_u = __ballot(nonfin);
                if (lane == 0) s_ufin = bad_u ? 0 : 1;
            }
            __syncthreads();
            wpos = s_win[0];
            wrid = s_win[1];
        }
        const double pv = s_pv;
        const bool ufin = s_ufin != 0;

        // ---- interchange bookkeeping, scale, rank-1 update + rotate --------
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            if (rid[k] == wrid) { pos[k] = j; done[k] = true; }
            else if (pos[k] == j) pos[k] = wpos;
        }
        const bool zero = pv == 0.0;
        const double inv = 1.0 / pv;
        const bool tiny = fabs(pv) < DBL_MIN;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            // multiplier of a live row; finished rows just rotate (l = 0)
            double l = row[k][0];
            if (!zero) l = tiny ? l / pv : l * inv;
            const double lm = done[k] ? 0.0 : l;
            const double last = done[k] ? row[k][0] : l;
            // the reference updates only the rows below the pivot
            // (tf2.tcc:130-132): a finished row must not see 0 * Inf / NaN
            // from a pivot row that holds them
            const bool plain = done[k] ? ufin : __builtin_isfinite(lm);
            if (plain) {
                // new[c] = row[c+1] - l*u[c+1]; finished columns see u = 0
#pragma unroll
                for (int c = 0; c < W - 1; ++c) row[k][c] = fma(-lm, s_u[c], row[k][c + 1]);
            } else {
                // a non-finite multiplier must not touch the finished
                // columns; a finished row only rotates
                const int upd = done[k] ? 0 : live - 1;
#pragma unroll
                for (int c = 0; c < W - 1; ++c)
                    row[k][c] = (c < upd) ? fma(-lm, s_u[c], row[k][c + 1]) : row[k][c + 1];
            }
            row[k][W - 1] = last;
        }
        if (tid == 0) {
            s_piv[j] = wpos;
            if (zero && first_zero == 0) first_zero = j + 1;
        }
        // s_u / s_own are rewritten next column only after its first barrier,
        // which every thread reaches after finishing the update above
    }
    // undo the remaining rotation: element c holds column c again
    for (int j = mn; j < W; ++j) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const double first = row[k][0];
#pragma unroll
            for (int c = 0; c < W - 1; ++c) row[k][c] = row[k][c + 1];
            row[k][W - 1] = first;
        }
    }

    // write every row to its final position; record the moved ones
    int *mylist = list + (size_t)g * PANEL_LIST_SLICE;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        if (rid[k] < mp) {
            double *dst = P + (long)pos[k] * ldw;
#pragma unroll
            for (int c = 0; c < W; ++c)
                if (c < w) dst[c] = row[k][c];
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
                               int ib)
{
    if (mp > PANEL_ROWS_MAX) return hipErrorInvalidValue;
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
    // workgroup, at most PANEL_GROUPS_MAX cooperating workgroups
    (void)ib;
    constexpr int NT = 256, RPT = 2;
    static_assert(NT * RPT == PANEL_ROWS_PER_WG, "panel geometry");
    // the words a workgroup gathers per column scale with the group count:
    // panels of at most half the largest height use the smaller variant
    const int G = panel_groups(mp);
    if (G <= PANEL_GROUPS_MAX / 2)
        hipLaunchKernelGGL((panel_kernel<NT, RPT, PANEL_GROUPS_MAX / 2>), dim3(G), dim3(NT), 0, s,
                           P, ldw, mp, w, wk.exch, wk.gen, piv, piv_base, wk.list, info, piv_base,
                           wk.flags);
    else
        hipLaunchKernelGGL((panel_kernel<NT, RPT, PANEL_GROUPS_MAX>), dim3(G), dim3(NT), 0, s, P,
                           ldw, mp, w, wk.exch, wk.gen, piv, piv_base, wk.list, info, piv_base,
                           wk.flags);
    return hipGetLastError();
}

} // namespace flx
