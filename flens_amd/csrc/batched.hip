// batched.hip -- thousands of independent, uneven LU solves and GEMMs in one
// persistent launch each.
//
// Solves: every system (n <= 512, nrhs <= 8, column-major, ld = n) is handled
// by one 256-thread workgroup from start to end: blocked right-looking LU
// with partial pivoting (the algorithm of flens::lapack::trf, flens/lapack/
// ge/trf.tcc:59-152, with tf2's pivot rule, tf2.tcc:101-126: first maximum,
// INFO = first exact zero) -- panel of 32 columns held in registers, MFMA
// trailing update -- then the forward / backward substitution of trs
// (flens/lapack/ge/trs.tcc:59-109).  Workgroups take systems from an atomic
// queue ordered by decreasing cost, so the launch needs no host round trip
// per system and balances itself.
//   mode 0: FP64 LU in place + solve (B overwritten by X);
//   mode 1: FP32 LU of a converted copy, FP64 refinement (residuals on the
//           FP64 A, FP32 corrections), FP64 LU + solve when refinement does
//           not reach the tolerance within itmax steps (A overwritten then).
// GEMMs: C_i <- beta_i C_i + alpha_i A_i B_i (FLENS generic semantics: beta
// scales C first, so beta = 0 computes 0*C) over a flat list of output
// tiles of all problems, one MFMA tile per workgroup pass.
#include <cfloat>
#include <climits>
#include "flx_common.h"
#include "flx_kernels.h"

namespace flx {

constexpr int BT = 256;           // threads per workgroup
constexpr int BPW = 32;           // panel width
constexpr int BMAX = 512;         // largest n
constexpr int BSLOT = BMAX / BT;  // panel rows per thread
constexpr int BRHS = 8;           // largest nrhs

// ---- small helpers ----------------------------------------------------------
FLX_DEV double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
FLX_DEV float readlane_t(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
FLX_DEV double readlane_t(double v, int lane) { return readlane_d(v, lane); }

// |v| ordered as unsigned bits: (hi, lo) of a double, (bits, 0) of a float
FLX_DEV void abs_key(double v, uint32_t &hi, uint32_t &lo)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
    hi = (uint32_t)(b >> 32);
    lo = (uint32_t)b;
}
FLX_DEV void abs_key(float v, uint32_t &hi, uint32_t &lo)
{
    hi = __float_as_uint(fabsf(v));
    lo = 0u;
}

FLX_DEV double4_t mfma_t(double a, double b, double4_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
FLX_DEV float4_t mfma_t(float a, float b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Acc4;
template <> struct Acc4<double> {
    typedef double4_t type;
    FLX_DEV int row(int lk, int q) { return lk + 4 * q; }
};
template <> struct Acc4<float> {
    typedef float4_t type;
    FLX_DEV int row(int lk, int q) { return 4 * lk + q; }
};

// the batched solver's precision ladders (BatchProb::cfg bits 0-1)
constexpr int LADDER_FP64 = 0, LADDER_FP32 = 1, LADDER_BF16 = 2, LADDER_FP16 = 3;

// x rounded to nearest-even BF16 (RND 1) or FP16 (RND 2), kept as a float
template <int RND, typename T>
FLX_DEV T round_op(T x)
{
    if constexpr (RND == 1) {
        uint32_t u = __float_as_uint(x);
        if ((u & 0x7f800000u) != 0x7f800000u) {
            u += 0x7fffu + ((u >> 16) & 1u);
            u &= 0xffff0000u;
        }
        return __uint_as_float(u);
    } else if constexpr (RND == 2) {
        return (float)(_Float16)x;
    } else {
        return x;
    }
}

// shared memory of one workgroup (the solve buffers double as scratch)
struct BShared {
    double cand[2][4][BPW];        // per-wave pivot candidate rows (as double)
    double cpiv[2][4];             // their pivot entries
    uint32_t khi[2][4], klo[2][4];
    int kpos[2][4];
    double l11[BPW][BPW + 1];
    int perm[BMAX];                // original row now at each position
    double gsum[4][BRHS];          // block sums (wg_sum8)
    int mv_dst[2 * BPW], mv_src[2 * BPW];
    int nmv;
    int info;
    double y[BMAX * BRHS];         // solve right-hand sides (T-typed view)
    double xs[BMAX * BRHS];        // refinement iterate (FP64), between factorisations
    unsigned long long red[4 * BRHS];
    double rs[BRHS], bn[BRHS];
    int idx, idx2;
};

// ---- one panel of wg_getrf: rows k..n-1 of columns k..k+w-1, SL rows per
// thread (1 while the panel has at most BT rows, else 2), factored with
// partial pivoting and written back at the final positions; L11 to LDS and
// the row moves for the other columns to sh.mv_*.
template <typename T, int SL>
FLX_DEV void wg_panel(T *__restrict__ A, int n, int k, int w, int *__restrict__ ipiv,
                      BShared &sh)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T r[SL][BPW];
    int pos[SL], start[SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        start[s] = k + t + BT * s;
        pos[s] = start[s];
        const bool ok = start[s] < n;
#pragma unroll
        for (int c = 0; c < BPW; ++c)
            r[s][c] = (ok && c < w) ? A[(long)(k + c) * n + start[s]] : T(0);
    }
    // The column loop stays rolled: it runs once per panel of every
    // system, and unrolled (32 columns x SL rows) it overflows the
    // instruction cache.  Column j of a row is read by a select over the
    // row's registers and written back inside the update, which runs
    // over all BPW columns with the pivot row's entries at and left of j
    // posted as zeros -- every register index stays static (a dynamic
    // one would move the rows to scratch memory).
#pragma unroll 1
    for (int j = 0; j < w; ++j) {
        const int par = j & 1;
        const int kj = k + j;
        T vj[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            T v = r[s][0];
#pragma unroll
            for (int c = 1; c < BPW; ++c) v = (c == j) ? r[s][c] : v;
            vj[s] = v;
        }
        // best candidate of this thread: largest |v|, then lowest position
        uint32_t bhi = 0, blo = 0;
        int bpos = INT_MAX, bs = -1;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            if (pos[s] >= kj && pos[s] < n) {
                uint32_t hi, lo;
                abs_key(vj[s], hi, lo);
                if (vj[s] != vj[s]) {
                    // iamax's NaN rule (cxxblas/level1/iamax.tcc:42-58): a NaN
                    // wins only at the first live position, elsewhere it ranks
                    // as |0| (the first position, lowest, wins every tie)
                    const bool first = pos[s] == kj;
                    hi = first ? 0xfffffffeu : 0u;
                    lo = first ? 0xffffffffu : 0u;
                }
                const bool better = bs < 0 || hi > bhi || (hi == bhi && lo > blo) ||
                                    (hi == bhi && lo == blo && pos[s] < bpos);
                if (better) { bhi = hi; blo = lo; bpos = pos[s]; bs = s; }
            }
        }
        // wave: max hi, then max lo among those, then min position
        const uint32_t valid = bs >= 0 ? 1u : 0u;
        const uint32_t mhi = wave_umax(valid ? bhi + 1u : 0u);
        const bool c1 = valid && bhi + 1u == mhi;
        const uint32_t mlo = wave_umax(c1 ? blo : 0u);
        const bool c2 = c1 && blo == mlo;
        const uint32_t mpos = wave_umin(c2 ? (uint32_t)bpos : 0xffffffffu);
        if (c2 && (uint32_t)bpos == mpos) {
            // the slot's row by blending with exact 0/1 weights; entries at
            // and left of column j posted as zero (the update's mask), the
            // pivot itself separately
            T wgt[SL];
#pragma unroll
            for (int s = 0; s < SL; ++s) wgt[s] = (bs == s) ? T(1) : T(0);
            T pv = vj[0] * wgt[0];
#pragma unroll
            for (int s = 1; s < SL; ++s) pv = pv + vj[s] * wgt[s];
            sh.cpiv[par][wave] = (double)pv;
#pragma unroll
            for (int c = 0; c < BPW; ++c) {
                T v = r[0][c] * wgt[0];
#pragma unroll
                for (int s = 1; s < SL; ++s) v = v + r[s][c] * wgt[s];
                sh.cand[par][wave][c] = (c > j) ? (double)v : 0.0;
            }
        }
        if (lane == 0) {
            sh.khi[par][wave] = mhi;
            sh.klo[par][wave] = mlo;
            sh.kpos[par][wave] = (int)mpos;
        }
        __syncthreads();
        int bw = 0;
#pragma unroll
        for (int q = 1; q < 4; ++q) {
            const uint32_t h = sh.khi[par][q], l = sh.klo[par][q];
            const int p = sh.kpos[par][q];
            const uint32_t hb = sh.khi[par][bw], lb = sh.klo[par][bw];
            const int pb = sh.kpos[par][bw];
            if (h > hb || (h == hb && (l > lb || (l == lb && p < pb)))) bw = q;
        }
        const int pp = sh.kpos[par][bw];   // pivot row position
        const T piv = (T)sh.cpiv[par][bw];
        if (t == 0) {
            ipiv[kj] = pp + 1;
            const int a = sh.perm[kj];
            sh.perm[kj] = sh.perm[pp];
            sh.perm[pp] = a;
            if (piv == T(0) && sh.info == 0) sh.info = kj + 1;
        }
        T pm[BPW];
#pragma unroll
        for (int c = 0; c < BPW; ++c) pm[c] = (T)sh.cand[par][bw][c];
        // scaling as tf2 (tf2.tcc:116-124): by the reciprocal of the pivot,
        // dividing when |pivot| is below the smallest normal number
        constexpr T TINY = sizeof(T) == 8 ? (T)DBL_MIN : (T)FLT_MIN;
        const bool tiny = fabs(piv) < TINY;
        const T inv = (piv != T(0)) ? T(1) / piv : T(0);
        // swap positions kj <-> pp, eliminate below kj
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            if (pos[s] == pp) pos[s] = kj;
            else if (pos[s] == kj) pos[s] = pp;
            if (pos[s] > kj && pos[s] < n) {
                const T l = (piv == T(0)) ? T(0) : tiny ? vj[s] / piv : vj[s] * inv;
#pragma unroll
                for (int c = 0; c < BPW; ++c) r[s][c] = (c == j) ? l : r[s][c] - l * pm[c];
            }
        }
    }
    __syncthreads();
    // ---- write the panel back at the final positions; L11 to LDS
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        if (start[s] < n) {
#pragma unroll
            for (int c = 0; c < BPW; ++c)
                if (c < w) A[(long)(k + c) * n + pos[s]] = r[s][c];
            if (pos[s] < k + w)
#pragma unroll
                for (int c = 0; c < BPW; ++c) sh.l11[pos[s] - k][c] = (double)r[s][c];
        }
    }
    // ---- the panel's row moves, for the columns outside it
    if (t == 0) sh.nmv = 0;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SL; ++s) {
        if (start[s] < n && pos[s] != start[s]) {
            const int e = atomicAdd(&sh.nmv, 1);
            if (e < 2 * BPW) { sh.mv_dst[e] = pos[s]; sh.mv_src[e] = start[s]; }
        }
    }
}

// ---- trailing update of one panel step of wg_getrf (the columns right of
// panel k..k+w-1): the trailing columns in strips of ST 16-column tiles,
// one strip per wave at a time: row moves, U12 = X A12 and A22 -= L21 U12
// on 16x16 MFMA tiles (each L21 fragment serves the strip's tiles).  The K
// index of MFMA step q on lane group lk is kx(q) = 16 (q / 4) + (row of
// accumulator entry q % 4), so the two U12 accumulator tiles of a column
// tile are the update's B fragments as they stand.  RND rounds the GEMM's
// operands (L21, U12) to BF16 (1) or FP16 (2) before the FP32 MFMA: the
// products of two such values are exact in FP32, so the update is the one
// a BF16 / FP16 MFMA with FP32 accumulation computes; the factors keep
// their FP32 values.  ST (the tile configuration) trades L21 fragment
// reuse (2) against more strips, and so more busy waves, on small
// trailing matrices (1).
template <typename T, int RND, int ST>
FLX_DEV void wg_update(T *__restrict__ A, int n, int k, int w, int nmv, BShared &sh)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c1 = k + w, m2 = n - c1;
        const int lr = lane & 15, lk = lane >> 4;
        constexpr int KS = BPW / 4;
        auto kx = [&](int q) { return 16 * (q / 4) + Acc4<T>::row(lk, q % 4); };
        // X fragments: tile 0 (rows 0-15) needs K < 16 only
        T x0[4], x1[KS];
#pragma unroll
        for (int q = 0; q < KS; ++q) {
            const int kk = 4 * q + lk;
            if (q < 4) x0[q] = kk < lr ? (T)sh.l11[lr][kk] : (kk == lr ? T(1) : T(0));
            x1[q] = kk < 16 + lr ? (T)sh.l11[16 + lr][kk] : (kk == 16 + lr ? T(1) : T(0));
        }
        const bool act = lane < nmv;
        const int e = act ? lane : 0;
        const int src = sh.mv_src[e], dst = sh.mv_dst[e];
        const int nt = (m2 + 15) / 16, nst = (nt + ST - 1) / ST;
        for (int st = wave; st < nst; st += 4) {
            const int c0 = c1 + 16 * ST * st;
            // row moves of the strip's columns, 16 at a time
            if (nmv > 0) {
#pragma unroll
                for (int h = 0; h < ST; ++h) {
                    T v[16];
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const int cc = c0 + 16 * h + c;
                        v[c] = (act && cc < n) ? A[(long)cc * n + src] : T(0);
                    }
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const int cc = c0 + 16 * h + c;
                        if (act && cc < n) A[(long)cc * n + dst] = v[c];
                    }
                }
                // the strip's other lanes read these rows next
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            }
            bool cok[ST];
            T *cp[ST];
            T bu[ST][KS];
#pragma unroll
            for (int h = 0; h < ST; ++h) {
                const int col = c0 + 16 * h + lr;
                cok[h] = col < n;
                cp[h] = A + (long)(cok[h] ? col : c0) * n;
                // U12 (32 x 16) = X A12
                T g[KS];
#pragma unroll
                for (int q = 0; q < KS; ++q) g[q] = cok[h] ? cp[h][k + 4 * q + lk] : T(0);
                typename Acc4<T>::type u0 = {T(0), T(0), T(0), T(0)}, u1 = u0;
#pragma unroll
                for (int q = 0; q < 4; ++q) u0 = mfma_t(x0[q], g[q], u0);
#pragma unroll
                for (int q = 0; q < KS; ++q) u1 = mfma_t(x1[q], g[q], u1);
#pragma unroll
                for (int q = 0; q < 4; ++q) { bu[h][q] = u0[q]; bu[h][4 + q] = u1[q]; }
                if (cok[h]) {
#pragma unroll
                    for (int q = 0; q < KS; ++q) cp[h][k + kx(q)] = bu[h][q];
                }
#pragma unroll
                for (int q = 0; q < KS; ++q) bu[h][q] = round_op<RND>(bu[h][q]);
            }
            // A22 -= L21 U12, 16-row blocks, the next block's L21
            // fragments and C entries requested before this block's MFMAs
            T a[KS], cv[ST][4];
            auto fetch = [&](int ti) {
                const int ar = c1 + 16 * ti + lr;
                const bool rok = ti < nt && ar < n;
#pragma unroll
                for (int q = 0; q < KS; ++q)
                    a[q] = round_op<RND>(rok ? A[(long)(k + kx(q)) * n + ar] : T(0));
#pragma unroll
                for (int h = 0; h < ST; ++h)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = c1 + 16 * ti + Acc4<T>::row(lk, q);
                        cv[h][q] = (ti < nt && cok[h] && row < n) ? cp[h][row] : T(0);
                    }
            };
            fetch(0);
            for (int ti = 0; ti < nt; ++ti) {
                typename Acc4<T>::type acc[ST];
                T c[ST][4];
#pragma unroll
                for (int h = 0; h < ST; ++h) {
                    acc[h] = {T(0), T(0), T(0), T(0)};
#pragma unroll
                    for (int q = 0; q < KS; ++q) acc[h] = mfma_t(a[q], bu[h][q], acc[h]);
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[h][q] = cv[h][q];
                }
                fetch(ti + 1);
#pragma unroll
                for (int h = 0; h < ST; ++h)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = c1 + 16 * ti + Acc4<T>::row(lk, q);
                        if (cok[h] && row < n) cp[h][row] = c[h][q] - acc[h][q];
                    }
            }
        }
}

// ---- LU of one system: A (n x n, ld n, column-major) in place -----------------
// ipiv (n ints, 1-based, LAPACK convention); sh.perm receives the row order;
// returns INFO (first exact zero pivot, 1-based, 0 if none).
template <typename T, int RND = 0>
FLX_DEV int wg_getrf(T *__restrict__ A, int n, int *__restrict__ ipiv, BShared &sh, int tile)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int i = t; i < n; i += BT) sh.perm[i] = i;
    if (t == 0) sh.info = 0;
    __syncthreads();
    for (int k = 0; k < n; k += BPW) {
        const int w = min(BPW, n - k);
        // ---- panel (one row per thread while it has at most BT rows)
        if (n - k <= BT) wg_panel<T, 1>(A, n, k, w, ipiv, sh);
        else wg_panel<T, BSLOT>(A, n, k, w, ipiv, sh);
        __syncthreads();
        const int nmv = min(sh.nmv, 2 * BPW);
        const int c1 = k + w, m2 = n - c1;      // first trailing column, trailing size
        // row position after this panel, by original row: the later
        // panels' interchanges on this panel's columns are applied once, at
        // the end (sh.y is free until the solve: n ints per panel)
        {
            int *const after = reinterpret_cast<int *>(sh.y) + (k / BPW) * n;
            for (int i = t; i < n; i += BT) after[sh.perm[i]] = i;
        }
        // ---- phase A: wave 0 replaces L11 (unit lower) by its inverse X,
        // row by row (X[i][j] = -(L[i][j] + sum_{j<q<i} L[i][q] X[q][j]))
        if (wave == 0 && m2 > 0) {
            const int j = lane & (BPW - 1);
            for (int i = 1; i < BPW; ++i) {
                const double li = sh.l11[i][j];
                double acc = li;
                for (int q = 1; q < i; ++q) {
                    const double xq = sh.l11[q][j];
                    const double lq = readlane_d(li, q);
                    if (j < q) acc = fma(lq, xq, acc);
                }
                if (lane < i) sh.l11[i][lane] = -acc;
            }
        }
        __syncthreads();
        if (m2 > 0) {
            if (tile) wg_update<T, RND, 1>(A, n, k, w, nmv, sh);
            else wg_update<T, RND, 2>(A, n, k, w, nmv, sh);
        }
        __syncthreads();
    }
    // ---- the later panels' interchanges on the factored columns (the
    // reference applies them panel by panel, flens/lapack/ge/trf.tcc:126):
    // row r >= k + BPW of a column in panel k takes the entry of the row now
    // at r as it stood after that panel.  Four columns per wave per pass,
    // every load of a pass before its stores.
    if (n > BPW) {
        const int *after = reinterpret_cast<const int *>(sh.y);
        constexpr int GC = 4, RL = BMAX / 64;
        const int jend = ((n - 1) / BPW) * BPW;   // first column of the last panel
        for (int j0 = GC * wave; j0 < jend; j0 += 4 * GC) {
            T v[GC][RL];
#pragma unroll
            for (int g = 0; g < GC; ++g) {
                const int j = j0 + g, pj = j / BPW;
#pragma unroll
                for (int u = 0; u < RL; ++u) {
                    const int r = (pj + 1) * BPW + lane + 64 * u;
                    v[g][u] = (j < n && r < n) ? A[(long)j * n + after[pj * n + sh.perm[r]]] : T(0);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
            for (int g = 0; g < GC; ++g) {
                const int j = j0 + g, pj = j / BPW;
#pragma unroll
                for (int u = 0; u < RL; ++u) {
                    const int r = (pj + 1) * BPW + lane + 64 * u;
                    if (j < n && r < n) A[(long)j * n + r] = v[g][u];
                }
            }
        }
        __syncthreads();
    }
    return sh.info;
}

// ---- solve with the factors: Y (n x nrhs, rows already permuted, in sh.y
// as T) <- U^{-1} L^{-1} Y.  64-row blocks: each wave substitutes two
// right-hand sides within the block (one row per lane, the solved row
// broadcast by readlane), then all threads fold the block into the rows
// below (above, for U).
template <typename T>
FLX_DEV void wg_trs(const T *__restrict__ A, int n, int nrhs, BShared &sh)
{
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    T *y = reinterpret_cast<T *>(sh.y);   // y[i * BRHS + c]
    const int nbk = (n + 63) / 64;
    constexpr int QC = 8;                 // block columns per batch of loads
    for (int pass = 0; pass < 2; ++pass) {
        const bool lower = pass == 0;
        for (int bb = 0; bb < nbk; ++bb) {
            const int blk = lower ? bb : nbk - 1 - bb;
            const int r0 = blk * 64, rows = min(64, n - r0);
            // wave w substitutes right-hand sides RW w .. RW w + RW - 1
            constexpr int RW = BRHS / 4;
            if (RW * wave < nrhs) {
                const int i = r0 + lane, cw = RW * wave;
                const bool rok = lane < rows;
                T v[RW];
#pragma unroll
                for (int c = 0; c < RW; ++c) v[c] = (rok && cw + c < nrhs) ? y[i * BRHS + cw + c] : T(0);
                // step s handles block column q = s (L, forwards) or
                // q = 63 - s (U, backwards); this lane's entry of column q:
                // below the diagonal for L, on and above it for U
                auto col = [&](int s) { return lower ? s : 63 - s; };
                auto ld = [&](int q) -> T {
                    const bool ok = rok && q < rows && (lower ? lane > q : lane <= q);
                    return ok ? A[(long)(r0 + q) * n + i] : T(0);
                };
                auto step = [&](int q, T a) {
                    if (q >= rows) return;          // past the block (uniform)
                    if (!lower) {
                        const T d = readlane_t(a, q);
                        if (lane == q) {
#pragma unroll
                            for (int c = 0; c < RW; ++c) v[c] = v[c] / d;
                        }
                    }
                    const T m = (lower || lane < q) ? a : T(0);
#pragma unroll
                    for (int c = 0; c < RW; ++c) {
                        const T xq = readlane_t(v[c], q);
                        v[c] = v[c] - m * xq;
                    }
                };
                // two batches in flight: the next batch's loads are issued
                // before the current batch's steps
                T b0[QC], b1[QC];
#pragma unroll
                for (int u = 0; u < QC; ++u) b0[u] = ld(col(u));
                for (int s0 = 0; s0 < 64; s0 += 2 * QC) {
#pragma unroll
                    for (int u = 0; u < QC; ++u) b1[u] = ld(col(s0 + QC + u));
#pragma unroll
                    for (int u = 0; u < QC; ++u) step(col(s0 + u), b0[u]);
#pragma unroll
                    for (int u = 0; u < QC; ++u)
                        b0[u] = (s0 + 2 * QC < 64) ? ld(col(s0 + 2 * QC + u)) : T(0);
#pragma unroll
                    for (int u = 0; u < QC; ++u) step(col(s0 + QC + u), b1[u]);
                }
                if (rok)
#pragma unroll
                    for (int c = 0; c < RW; ++c)
                        if (cw + c < nrhs) y[i * BRHS + cw + c] = v[c];
            }
            __syncthreads();
            // fold into the remaining rows: below (L) or above (U), the
            // block's 64 column entries of a row loaded 16 at a time
            const int f0 = lower ? r0 + rows : 0, f1 = lower ? n : r0;
            for (int i = f0 + t; i < f1; i += BT) {
                T acc[BRHS];
#pragma unroll
                for (int c = 0; c < BRHS; ++c) acc[c] = (c < nrhs) ? y[i * BRHS + c] : T(0);
                for (int q0 = 0; q0 < rows; q0 += 16) {
                    T av[16];
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        av[u] = (q0 + u < rows) ? A[(long)(r0 + q0 + u) * n + i] : T(0);
#pragma unroll
                    for (int u = 0; u < 16; ++u) {
                        if (q0 + u < rows) {
#pragma unroll
                            for (int c = 0; c < BRHS; ++c)
                                acc[c] = acc[c] - av[u] * y[(r0 + q0 + u) * BRHS + c];
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < BRHS; ++c)
                    if (c < nrhs) y[i * BRHS + c] = acc[c];
            }
            __syncthreads();
        }
    }
}

// y = (T) (src[perm[i]] * scale_c), src column-major n x nrhs (ld n)
// (rows scaled by rowscale[perm[i]] when given)
template <typename T>
FLX_DEV void wg_load_rhs(const double *__restrict__ src, int n, int nrhs, const double *scale,
                         BShared &sh, const double *rowscale = nullptr)
{
    T *y = reinterpret_cast<T *>(sh.y);
    for (int q0 = threadIdx.x; q0 < n * BRHS; q0 += 8 * BT) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + BT * u, i = q / BRHS, c = q % BRHS;
            const int pi = q < n * BRHS ? sh.perm[i] : 0;
            v[u] = (q < n * BRHS && c < nrhs) ? src[(long)c * n + pi] * (rowscale ? rowscale[pi] : 1.0) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + BT * u, c = q % BRHS;
            if (q < n * BRHS) y[q] = (c < nrhs) ? (T)(v[u] * (scale ? scale[c] : 1.0)) : T(0);
        }
    }
    __syncthreads();
}

struct BatchProb {
    long a;      // offset of A (and of its FP32 copy) in elements
    long b;      // offset of B / X / R in elements
    int n, nrhs;
    int piv;     // offset of ipiv
    int id;      // caller's index (info / iter outputs)
    int cfg;     // ladder | (tile << 2) (BatchCfg)
    int pad;
};

// block-wide max of a non-negative double per right-hand side into sh.red
FLX_DEV void wg_colmax_reset(BShared &sh)
{
    if (threadIdx.x < 4 * BRHS) sh.red[threadIdx.x] = 0ull;
    __syncthreads();
}

// selector features of one system (see the feature definitions below)
FLX_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// The selector network transposed for the kernel (a copy in device memory
// made once per plan), so that lane j reads consecutive words: W1T[i][j],
// b1[j], W2T[k][j], b2[j], W3[o][j], b3[o] (mlp.hip keeps W1[j][i],
// W2[j][k]).
constexpr int SEL_W1T = 0, SEL_B1 = SEL_W1T + MLP_IN * MLP_HIDDEN, SEL_W2T = SEL_B1 + MLP_HIDDEN,
              SEL_B2 = SEL_W2T + MLP_HIDDEN * MLP_HIDDEN, SEL_W3 = SEL_B2 + MLP_HIDDEN,
              SEL_B3 = SEL_W3 + MLP_OUT * MLP_HIDDEN, SEL_WORDS = SEL_B3 + MLP_OUT;
static_assert(SEL_WORDS == MLP_PARAMS, "selector layout");
__global__ __launch_bounds__(256) void stage_selector_kernel(const float *__restrict__ P,
                                                             float *__restrict__ S)
{
    constexpr int H = MLP_HIDDEN;
    constexpr int oB1 = H * MLP_IN, oW2 = oB1 + H, oB2 = oW2 + H * H;
    for (int q = blockIdx.x * 256 + threadIdx.x; q < MLP_PARAMS; q += gridDim.x * 256) {
        int dst = q;
        if (q < oB1) { const int j = q / MLP_IN, i = q % MLP_IN; dst = SEL_W1T + i * H + j; }
        else if (q >= oW2 && q < oB2) { const int j = (q - oW2) / H, k = (q - oW2) % H; dst = SEL_W2T + k * H + j; }
        S[dst] = P[q];
    }
}

// In-kernel inference by wave 0 from the transposed network: lane j computes
// hidden unit j; returns the arg-max class (uniform in the wave).
FLX_DEV int wg_select(const float *__restrict__ S, const float (&f)[MLP_IN])
{
    constexpr int H = MLP_HIDDEN;
    const int lane = threadIdx.x & 63;
    float a = S[SEL_B1 + lane];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) a = fmaf(S[SEL_W1T + i * H + lane], f[i], a);
    const float h1 = fmaxf(a, 0.0f);
    float b = S[SEL_B2 + lane];
    for (int k = 0; k < H; ++k) b = fmaf(S[SEL_W2T + k * H + lane], readlane_t(h1, k), b);
    const float h2 = fmaxf(b, 0.0f);
    int best = 0;
    float bz = -INFINITY;
    for (int o = 0; o < MLP_OUT; ++o) {
        float z = h2 * S[SEL_W3 + o * H + lane];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) z += __shfl_xor(z, d, 64);
        z += S[SEL_B3 + o];
        if (z > bz) { bz = z; best = o; }
    }
    return best;
}

// the network's class for the whole workgroup (wave 0 evaluates it)
FLX_DEV int wg_select_all(const float *__restrict__ P, const float (&f)[MLP_IN], BShared &sh)
{
    if (threadIdx.x < 64) {
        const int c = wg_select(P, f);
        if (threadIdx.x == 0) sh.idx2 = c;
    }
    __syncthreads();
    const int c = sh.idx2;
    __syncthreads();
    return c;
}

// block sums of BRHS values per thread; every thread receives the sums
FLX_DEV void wg_sum8(double (&v)[BRHS], BShared &sh)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < BRHS; ++c)
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) v[c] += __shfl_xor(v[c], d, 64);
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < BRHS; ++c) sh.gsum[wave][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < BRHS; ++c) v[c] = sh.gsum[0][c] + sh.gsum[1][c] + sh.gsum[2][c] + sh.gsum[3][c];
    __syncthreads();
}

FLX_DEV double red_max(const BShared &sh, int slot)
{
    return __longlong_as_double((long long)sh.red[slot]);
}

FLX_DEV double wave_max_d(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}
FLX_DEV double wave_min_d(double v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = fmin(v, __shfl_xor(v, d, 64));
    return v;
}

// The low-precision copy w of a: FP32 values (RND 0), FP32 values rounded to
// BF16 (1), or rows scaled by rsc (mu / row max) and rounded to FP16 (2).
// Returns 1 (uniform) when an entry does not fit the format.
template <int RND>
FLX_DEV int wg_convert(const double *__restrict__ a, int n, float *__restrict__ w,
                       const double *__restrict__ rsc)
{
    const int t = threadIdx.x;
    int ovf = 0;
    double sc[2] = {1.0, 1.0};
    if (RND == 2) {
#pragma unroll
        for (int h = 0; h < 2; ++h) sc[h] = (t + BT * h < n) ? rsc[t + BT * h] : 1.0;
    }
    for (int j0 = 0; j0 < n; j0 += 16) {
        double v[2][16];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = t + BT * h;
                v[h][u] = (i < n && j0 + u < n) ? a[(long)(j0 + u) * n + i] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = t + BT * h, j = j0 + u;
                if (i < n && j < n) {
                    const double x = v[h][u] * sc[h];
                    ovf |= !(fabs(x) <= (double)FLT_MAX);
                    w[(long)j * n + i] = round_op<RND>((float)x);
                }
            }
    }
    return __syncthreads_or(ovf);
}

// y (sh.y, float) <- M^{-1} src for the low-precision factor w, src column-
// major n x nrhs (ld n) in FP64: rows permuted and scaled by rsc (FP16
// ladder) and by colscale[c] (into FP32 range), then the two triangular
// solves.
FLX_DEV void wg_precond(const float *__restrict__ w, int n, int nrhs, const double *src,
                        const double *colscale, const double *rsc, BShared &sh)
{
    wg_load_rhs<float>(src, n, nrhs, colscale, sh, rsc);
    wg_trs<float>(w, n, nrhs, sh);
}

// Lower bound of ||A^{-1}||_inf from two preconditioned solves (e = 1, then
// the signs of A^{-1} e), with the low-precision factor; e and the signs go
// through the scratch column g (n doubles).
FLX_DEV double wg_inv_norm_est(const float *__restrict__ w, int n, const double *rsc,
                               double *__restrict__ g, BShared &sh)
{
    const int t = threadIdx.x;
    for (int i = t; i < n; i += BT) g[i] = 1.0;
    __syncthreads();
    wg_precond(w, n, 1, g, nullptr, rsc, sh);
    const float *y = reinterpret_cast<const float *>(sh.y);
    wg_colmax_reset(sh);
    // the solves return x in the order of the unknowns; its signs form the
    // second right-hand side (row-permuted again on loading)
    for (int i = t; i < n; i += BT) {
        const float v = y[i * BRHS];
        atomicMax(&sh.red[0], (unsigned long long)__double_as_longlong(fabs((double)v)));
        g[i] = (v >= 0.0f) ? 1.0 : -1.0;
    }
    __syncthreads();
    wg_precond(w, n, 1, g, nullptr, rsc, sh);
    for (int i = t; i < n; i += BT)
        atomicMax(&sh.red[0], (unsigned long long)__double_as_longlong(fabs((double)y[i * BRHS])));
    __syncthreads();
    const double r = red_max(sh, 0);
    __syncthreads();
    return r;
}

// r = b - A x (x = sh.xs) into dst (column-major n x nrhs, ld n), with
// sh.red[c] = ||r_c||_inf and sh.red[2 BRHS + c] = ||x_c||_inf
FLX_DEV void wg_residual(const double *__restrict__ a, const double *__restrict__ bx, int n,
                         int nrhs, double *__restrict__ dst, BShared &sh)
{
    const int t = threadIdx.x;
    wg_colmax_reset(sh);
    for (int i = t; i < n; i += BT) {
        double acc[BRHS];
#pragma unroll
        for (int c = 0; c < BRHS; ++c) acc[c] = (c < nrhs) ? bx[(long)c * n + i] : 0.0;
        for (int j0 = 0; j0 < n; j0 += 16) {
            double av[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) av[u] = (j0 + u < n) ? a[(long)(j0 + u) * n + i] : 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (j0 + u < n) {
#pragma unroll
                    for (int c = 0; c < BRHS; ++c)
                        acc[c] = fma(-av[u], sh.xs[(j0 + u) * BRHS + c], acc[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < BRHS; ++c) {
            if (c < nrhs) {
                dst[(long)c * n + i] = acc[c];
                atomicMax(&sh.red[c], (unsigned long long)__double_as_longlong(fabs(acc[c])));
                atomicMax(&sh.red[2 * BRHS + c],
                          (unsigned long long)__double_as_longlong(fabs(sh.xs[i * BRHS + c])));
            }
        }
    }
    __syncthreads();
}

// dst (column-major n x nrhs, ld n) = A v, v at rows i*BRHS + c
FLX_DEV void wg_matvec(const double *__restrict__ a, int n, int nrhs, const double *__restrict__ v,
                       double *__restrict__ dst)
{
    for (int i = threadIdx.x; i < n; i += BT) {
        double acc[BRHS];
#pragma unroll
        for (int c = 0; c < BRHS; ++c) acc[c] = 0.0;
        for (int j0 = 0; j0 < n; j0 += 16) {
            double av[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) av[u] = (j0 + u < n) ? a[(long)(j0 + u) * n + i] : 0.0;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                if (j0 + u < n) {
#pragma unroll
                    for (int c = 0; c < BRHS; ++c)
                        if (c < nrhs) acc[c] = fma(av[u], v[(j0 + u) * BRHS + c], acc[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < BRHS; ++c)
            if (c < nrhs) dst[(long)c * n + i] = acc[c];
    }
    __syncthreads();
}

// ||r_c|| <= sqrt(n) 2^-53 (||A|| ||x_c|| + ||b_c||) for every column (the
// norms of wg_residual and sh.bn); sets sh.rs[c] = 1 / ||r_c|| (FP32 range)
FLX_DEV bool wg_converged(int n, int nrhs, double anrm, BShared &sh)
{
    const double tol = sqrt((double)n) * 1.1102230246251565e-16;
    bool all = true;
#pragma unroll
    for (int c = 0; c < BRHS; ++c) {
        const double rn = red_max(sh, c), xn = red_max(sh, 2 * BRHS + c);
        if (c < nrhs) all = all && (rn <= tol * (anrm * xn + sh.bn[c]));
    }
    if (threadIdx.x < BRHS) {
        const double rn = red_max(sh, threadIdx.x);
        sh.rs[threadIdx.x] = (rn > 0.0 && rn <= DBL_MAX) ? 1.0 / rn : 1.0;
    }
    __syncthreads();
    return all;
}

// x0 = M^{-1} b into sh.xs (FP64)
FLX_DEV void wg_initial_solve(const float *__restrict__ w, const double *__restrict__ bx, int n,
                              int nrhs, const double *rsc, BShared &sh)
{
    wg_precond(w, n, nrhs, bx, nullptr, rsc, sh);
    for (int q = threadIdx.x; q < n * BRHS; q += BT)
        sh.xs[q] = (double)reinterpret_cast<const float *>(sh.y)[q];
    __syncthreads();
}

// x (sh.xs) to B on convergence
FLX_DEV void wg_store_x(double *__restrict__ bx, int n, int nrhs, BShared &sh)
{
    for (int q = threadIdx.x; q < n * nrhs; q += BT) {
        const int c = q / n, i = q % n;
        bx[q] = sh.xs[i * BRHS + c];
    }
}

// Classic refinement (dsgesv's loop, dsgesv.f:247-310): x += M^{-1} r with
// FP64 residuals, at most itmax corrections.  Returns the number of
// corrections, or -itmax - 1 when the tolerance was not reached.
FLX_DEV int wg_refine_classic(const double *__restrict__ a, double *__restrict__ bx,
                              const float *__restrict__ w, double *__restrict__ r, int n, int nrhs,
                              double anrm, int itmax, BShared &sh)
{
    wg_initial_solve(w, bx, n, nrhs, nullptr, sh);
    for (int it = 0; it <= itmax; ++it) {
        wg_residual(a, bx, n, nrhs, r, sh);
        if (wg_converged(n, nrhs, anrm, sh)) { wg_store_x(bx, n, nrhs, sh); return it; }
        if (it == itmax) break;
        // x += M^{-1} r  (r scaled by 1/||r|| into FP32 range and back)
        wg_precond(w, n, nrhs, r, sh.rs, nullptr, sh);
        for (int q = threadIdx.x; q < n * BRHS; q += BT) {
            const int c = q % BRHS;
            sh.xs[q] += (double)reinterpret_cast<const float *>(sh.y)[q] / sh.rs[c];
        }
        __syncthreads();
    }
    return -itmax - 1;
}

// Per-workgroup scratch of the low-precision ladders (doubles): the Krylov
// basis V (GM + 1 vectors, row i / column c at i * BRHS + c), the operator's
// product U (column-major, ld n), the FP16 ladder's row scales and one
// vector for the condition estimate.
constexpr int GM = 8;                         // GMRES steps per refinement step
constexpr long GV = 0;
constexpr long GU = GV + (long)(GM + 1) * BMAX * BRHS;
constexpr long GRSC = GU + (long)BMAX * BRHS;
constexpr long GVEC = GRSC + BMAX;
constexpr long GSTRIDE = GVEC + BMAX;

// GMRES-IR (Carson and Higham, SIAM J. Sci. Comput. 40(2), 2018): the outer
// loop is classic refinement with FP64 residuals; each correction solves
// M^{-1} A d = M^{-1} r by GMRES (at most GM steps, modified Gram-Schmidt,
// Givens rotations, FP64 vectors) with the low-precision factor applied in
// FP32.  The operator is scaled by 1/||A||_inf and the right-hand side by
// 1/||r_c||_inf so that every preconditioner input stays in FP32 range.
// Returns the number of corrections or -itmax - 1; *inner += GMRES steps.
FLX_DEV int wg_refine_gmres(const double *__restrict__ a, double *__restrict__ bx,
                            const float *__restrict__ w, int n, int nrhs, double anrm,
                            const double *rsc, double *__restrict__ gws, int itmax, int &inner,
                            BShared &sh)
{
    const int t = threadIdx.x;
    double *V = gws + GV, *U = gws + GU;
    double *HB = &sh.l11[0][0];               // free once the factor is done
    auto H = [&](int c, int i, int j) -> double & { return HB[(c * (GM + 1) + i) * GM + j]; };
    double *CS = HB + BRHS * (GM + 1) * GM, *SN = CS + BRHS * GM, *GV_ = SN + BRHS * GM;
    double *RES = GV_ + BRHS * (GM + 1), *RINV = RES + BRHS, *SCA = RINV + BRHS, *BETA = SCA + BRHS;
    static_assert(BRHS * ((GM + 1) * GM + 2 * GM + (GM + 1) + 4) <= 33 * 32, "GMRES state");
    auto RES0 = [&](int c) { return BETA[c]; };
    const double sA = anrm > 0.0 && anrm <= DBL_MAX ? 1.0 / anrm : 1.0;
    wg_initial_solve(w, bx, n, nrhs, rsc, sh);
    for (int it = 0; it <= itmax; ++it) {
        wg_residual(a, bx, n, nrhs, U, sh);
        if (wg_converged(n, nrhs, anrm, sh)) { wg_store_x(bx, n, nrhs, sh); return it; }
        if (it == itmax) break;
        // z0 = M^{-1} (r / ||r||): rows t, t + BT of every column in registers
        wg_precond(w, n, nrhs, U, sh.rs, rsc, sh);
        double z[2][BRHS], nz[BRHS];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int c = 0; c < BRHS; ++c) {
                const int i = t + BT * h;
                z[h][c] = (i < n && c < nrhs) ? (double)reinterpret_cast<const float *>(sh.y)[i * BRHS + c] : 0.0;
            }
#pragma unroll
        for (int c = 0; c < BRHS; ++c) nz[c] = z[0][c] * z[0][c] + z[1][c] * z[1][c];
        wg_sum8(nz, sh);
        if (t < BRHS) {
            const double beta = sqrt(nz[t]);
            RINV[t] = sh.rs[t];
            SCA[t] = sA;
            BETA[t] = beta;
            GV_[t * (GM + 1)] = beta;
            for (int i = 1; i <= GM; ++i) GV_[t * (GM + 1) + i] = 0.0;
            RES[t] = beta;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = t + BT * h;
            if (i < n)
#pragma unroll
                for (int c = 0; c < BRHS; ++c)
                    V[i * BRHS + c] = nz[c] > 0.0 ? z[h][c] / sqrt(nz[c]) : 0.0;
        }
        __syncthreads();
        int steps = 0;
        for (int j = 0; j < GM; ++j) {
            // w = M^{-1} (A v_j / ||A||)
            wg_matvec(a, n, nrhs, V + (long)j * BMAX * BRHS, U);
            wg_precond(w, n, nrhs, U, SCA, rsc, sh);
            double wv[2][BRHS];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int c = 0; c < BRHS; ++c) {
                    const int i = t + BT * h;
                    wv[h][c] = (i < n && c < nrhs) ? (double)reinterpret_cast<const float *>(sh.y)[i * BRHS + c] : 0.0;
                }
            // modified Gram-Schmidt against v_0 .. v_j
            for (int q = 0; q <= j; ++q) {
                const double *vq = V + (long)q * BMAX * BRHS;
                double d[BRHS];
#pragma unroll
                for (int c = 0; c < BRHS; ++c) d[c] = 0.0;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int i = t + BT * h;
                    if (i < n)
#pragma unroll
                        for (int c = 0; c < BRHS; ++c) d[c] = fma(wv[h][c], vq[i * BRHS + c], d[c]);
                }
                wg_sum8(d, sh);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int i = t + BT * h;
                    if (i < n)
#pragma unroll
                        for (int c = 0; c < BRHS; ++c) wv[h][c] = fma(-d[c], vq[i * BRHS + c], wv[h][c]);
                }
                if (t < BRHS) H(t, q, j) = d[t];
            }
            double nw[BRHS];
#pragma unroll
            for (int c = 0; c < BRHS; ++c) nw[c] = wv[0][c] * wv[0][c] + wv[1][c] * wv[1][c];
            wg_sum8(nw, sh);
            double *vn = V + (long)(j + 1) * BMAX * BRHS;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = t + BT * h;
                if (i < n)
#pragma unroll
                    for (int c = 0; c < BRHS; ++c)
                        vn[i * BRHS + c] = nw[c] > 0.0 ? wv[h][c] / sqrt(nw[c]) : 0.0;
            }
            // Givens rotations of column j of H (one thread per right-hand side)
            if (t < BRHS) {
                const int c = t;
                H(c, j + 1, j) = sqrt(nw[c]);
                for (int q = 0; q < j; ++q) {
                    const double h0 = H(c, q, j), h1 = H(c, q + 1, j);
                    H(c, q, j) = CS[c * GM + q] * h0 + SN[c * GM + q] * h1;
                    H(c, q + 1, j) = -SN[c * GM + q] * h0 + CS[c * GM + q] * h1;
                }
                const double h0 = H(c, j, j), h1 = H(c, j + 1, j);
                const double rr = hypot(h0, h1);
                const double cj = rr > 0.0 ? h0 / rr : 1.0, sj = rr > 0.0 ? h1 / rr : 0.0;
                CS[c * GM + j] = cj;
                SN[c * GM + j] = sj;
                H(c, j, j) = rr;
                H(c, j + 1, j) = 0.0;
                const double gj = GV_[c * (GM + 1) + j];
                GV_[c * (GM + 1) + j] = cj * gj;
                GV_[c * (GM + 1) + j + 1] = -sj * gj;
                RES[c] = fabs(-sj * gj);
            }
            __syncthreads();
            steps = j + 1;
            // every column's GMRES residual below 1e-10 of its start, or a
            // breakdown (w in the span of the basis)
            bool done = true;
#pragma unroll
            for (int c = 0; c < BRHS; ++c)
                if (c < nrhs)
                    done = done && (RES[c] <= 1e-10 * RES0(c) || !(nw[c] > 0.0));
            __syncthreads();
            if (done) break;
        }
        inner += steps;
        // y = H^{-1} g (upper triangular, steps x steps), per column
        if (t < BRHS) {
            const int c = t;
            for (int i = steps - 1; i >= 0; --i) {
                double s = GV_[c * (GM + 1) + i];
                for (int q = i + 1; q < steps; ++q) s -= H(c, i, q) * GV_[c * (GM + 1) + q];
                const double d = H(c, i, i);
                GV_[c * (GM + 1) + i] = d != 0.0 ? s / d : 0.0;
            }
        }
        __syncthreads();
        // x += (||r_c|| / ||A||) V y  (undoing the two scalings)
        for (int q = t; q < n * BRHS; q += BT) {
            const int i = q / BRHS, c = q % BRHS;
            if (c < nrhs) {
                double d = 0.0;
                for (int k = 0; k < steps; ++k) d = fma(V[(long)k * BMAX * BRHS + i * BRHS + c], GV_[c * (GM + 1) + k], d);
                sh.xs[q] += d * sA / RINV[c];
            }
        }
        __syncthreads();
    }
    return -itmax - 1;
}

// FP16 ladder: rsc_i = mu / max_j |a_ij| (mu = 0.1 x the largest FP16
// value, Higham, Pranesh and Zounon, SIAM J. Sci. Comput. 41(4), 2019), so
// every row of the scaled matrix has its largest entry at mu
FLX_DEV void wg_fp16_row_scales(const double *__restrict__ a, int n, double *__restrict__ rsc)
{
    constexpr double MU = 0.1 * 65504.0;
    const int t = threadIdx.x;
    double mx[2] = {0.0, 0.0};
    for (int j0 = 0; j0 < n; j0 += 16) {
        double v[2][16];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int i = t + BT * h;
                v[h][u] = (i < n && j0 + u < n) ? a[(long)(j0 + u) * n + i] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 16; ++u) mx[h] = fmax(mx[h], fabs(v[h][u]));
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (t + BT * h < n) rsc[t + BT * h] = (mx[h] > 0.0 && mx[h] <= DBL_MAX) ? MU / mx[h] : 1.0;
    __syncthreads();
}

// the selector's features before a factor is paid: log2 n / 9, nrhs / 8,
// the condition feature (KAPPA_UNKNOWN until a low-precision factor gives
// log10 kappa_inf / 16), log10 (diagonal dominance) / 4, log10 (max / min
// row norm) / 12
constexpr float KAPPA_UNKNOWN = -0.25f;
FLX_DEV void make_features(float (&f)[MLP_IN], int n, int nrhs, double anrm, double rmin,
                           double ddom)
{
    f[0] = __log2f((float)n) / 9.0f;
    f[1] = (float)nrhs / 8.0f;
    f[2] = KAPPA_UNKNOWN;
    f[3] = clampf(__log10f((float)fmin(fmax(ddom, 1e-30), 1e30)), -4.0f, 4.0f) / 4.0f;
    const double ratio = anrm / fmax(rmin, 1e-300);
    f[4] = clampf(__log10f((float)fmin(fmax(ratio, 1.0), 1e30)), 0.0f, 12.0f) / 12.0f;
    if (!(ratio <= DBL_MAX)) f[4] = 1.0f;
}
FLX_DEV float kappa_feature(double kap)
{
    if (!(kap <= DBL_MAX)) return 20.0f / 16.0f;
    return clampf(__log10f((float)fmin(fmax(kap, 1.0), 1e30)), 0.0f, 20.0f) / 16.0f;
}

// Row statistics of every system ahead of the solve kernels (one workgroup
// per system, few registers, so that many workgroups per CU stream the
// matrices at HBM rate): stats[3 id + {0, 1, 2}] = ||A||_inf, the smallest
// row 1-norm, the smallest diagonal-dominance ratio.  With the selector
// (selT) wave 0 then evaluates the network on the system's pre-factor
// features and writes its class to cls[id] (and choice[id]).
__global__ __launch_bounds__(BT) void row_stats_kernel(const BatchProb *__restrict__ probs,
                                                       int nprob, const double *__restrict__ A,
                                                       double *__restrict__ stats,
                                                       const float *__restrict__ selT,
                                                       int *__restrict__ cls,
                                                       int *__restrict__ choice)
{
    __shared__ double part[4][3];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const BatchProb p = probs[blockIdx.x];
    const int n = p.n;
    if (n <= 0 || n > BMAX) return;
    const double *a = A + p.a;
    double rs[2] = {0.0, 0.0}, dg[2] = {0.0, 0.0};
    for (int j0 = 0; j0 < n; j0 += 8) {
        double v[2][8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = t + BT * h;
                v[h][u] = (i < n && j0 + u < n) ? a[(long)(j0 + u) * n + i] : 0.0;
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const double av = fabs(v[h][u]);
                rs[h] += av;
                if (j0 + u == t + BT * h) dg[h] = av;
            }
    }
    double an = 0.0, amin = DBL_MAX, dmin = DBL_MAX;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (t + BT * h < n) {
            an = fmax(an, rs[h] != rs[h] ? DBL_MAX : rs[h]);
            amin = fmin(amin, rs[h]);
            dmin = fmin(dmin, dg[h] / fmax(rs[h] - dg[h], 1e-300));
        }
    }
    an = wave_max_d(an);
    amin = wave_min_d(amin);
    dmin = wave_min_d(dmin);
    if (lane == 0) { part[wave][0] = an; part[wave][1] = amin; part[wave][2] = dmin; }
    __syncthreads();
    const double anrm = fmax(fmax(part[0][0], part[1][0]), fmax(part[2][0], part[3][0]));
    const double rmin = fmin(fmin(part[0][1], part[1][1]), fmin(part[2][1], part[3][1]));
    const double ddom = fmin(fmin(part[0][2], part[1][2]), fmin(part[2][2], part[3][2]));
    if (t == 0) {
        stats[3L * p.id + 0] = anrm;
        stats[3L * p.id + 1] = rmin;
        stats[3L * p.id + 2] = ddom;
    }
    if (selT && wave == 0) {
        float f[MLP_IN];
        make_features(f, n, p.nrhs, anrm, rmin, ddom);
        const int c = wg_select(selT, f);
        if (lane == 0) {
            cls[p.id] = c;
            if (choice) choice[p.id] = c;
        }
    }
}

// rule-based ladder: FP32 factor, classic refinement while the condition
// estimate stays below this bound, else FP64 at once
constexpr double RULE_KAPPA_MAX = 1e6;

// POLICY 0: every system runs the ladder and tile of its table entry;
// 1: the selector network picks them before any factor (features without
// the condition estimate) and, after a low-precision factor, sends the
// system to FP64 when the estimate makes it pick the FP64 ladder;
// 2: the rule-based ladder.  A failed or skipped refinement ends in the
// FP64 LU (iters < 0).
template <int POLICY, bool LOWP>
__global__ __launch_bounds__(BT, 2) void batched_gesv_kernel(const BatchProb *__restrict__ probs,
                                                           int nprob, int *__restrict__ counter,
                                                           double *__restrict__ A,
                                                           double *__restrict__ B,
                                                           float *__restrict__ W32,
                                                           double *__restrict__ R,
                                                           double *__restrict__ G,
                                                           int *__restrict__ ipiv,
                                                           int *__restrict__ info,
                                                           int *__restrict__ iters, int itmax,
                                                           BatchOut out)
{
    __shared__ BShared sh;
    const int t = threadIdx.x;
    double *gws = G ? G + (long)blockIdx.x * GSTRIDE : nullptr;
    for (;;) {
        if (t == 0) sh.idx = atomicAdd(counter, 1);
        __syncthreads();
        const int idx = sh.idx;
        __syncthreads();
        if (idx >= nprob) break;
        const BatchProb p = probs[idx];
        const int n = p.n, nrhs = p.nrhs;
        const long long t_start = wall_clock64();
        if (n <= 0 || n > BMAX || nrhs < 0 || nrhs > BRHS) {
            if (t == 0) { info[p.id] = -1; if (iters) iters[p.id] = 0; }
            continue;
        }
        if (POLICY == 1) {
            // the learned dispatch runs two instances over the table: the
            // FP64-only one takes the systems the network sent to FP64
            const int c = out.cls[p.id];
            if ((((c & 3) == LADDER_FP64) != !LOWP)) continue;
            if (!LOWP) {
                const int tl = (c >> 2) & 1;
                double *a = A + p.a;
                double *bx = B + p.b;
                const long long t0 = t_start;
                const int inf = wg_getrf<double>(a, n, ipiv + p.piv, sh, tl);
                if (inf == 0 && nrhs > 0) {
                    wg_load_rhs<double>(bx, n, nrhs, nullptr, sh);
                    wg_trs<double>(a, n, nrhs, sh);
                    for (int q = t; q < n * nrhs; q += BT) {
                        const int cc = q / n, i = q % n;
                        bx[q] = sh.y[i * BRHS + cc];
                    }
                }
                if (t == 0) {
                    info[p.id] = inf;
                    if (iters) iters[p.id] = 0;
                    if (out.inner) out.inner[p.id] = 0;
                    if (out.used) out.used[p.id] = tl << 2;
                    if (out.tsys) out.tsys[p.id] = wall_clock64() - t0;
                }
                __syncthreads();
                continue;
            }
        }
        double *a = A + p.a;
        double *bx = B + p.b;
        int *pv = ipiv + p.piv;
        int ladder = p.cfg & 3, tile = (p.cfg >> 2) & 1;
        if (POLICY == 2) ladder = LADDER_FP32;
        if (!LOWP || !gws || !out.stats) ladder = LADDER_FP64;   // FP64 only
        float f[MLP_IN];
        double anrm = 0.0, rmin = 0.0, ddom = 0.0;
        const bool want_f = LOWP && (POLICY == 1 || out.feat) && out.stats;
        if (want_f || ladder != LADDER_FP64) {
            anrm = out.stats[3L * p.id];
            rmin = out.stats[3L * p.id + 1];
            ddom = out.stats[3L * p.id + 2];
            make_features(f, n, nrhs, anrm, rmin, ddom);
        }
        if (POLICY == 1 && want_f) {
            const int c = out.cls[p.id];
            ladder = c & 3;
            tile = (c >> 2) & 1;
        }
        int it_code = 0, inner = 0;
        bool need64 = ladder == LADDER_FP64;
        if constexpr (LOWP) if (!need64) {
            float *w = W32 + p.a;
            const double *rsc = nullptr;
            int ovf;
            if (ladder == LADDER_FP16) {
                wg_fp16_row_scales(a, n, gws + GRSC);
                rsc = gws + GRSC;
                ovf = wg_convert<2>(a, n, w, rsc);
            } else if (ladder == LADDER_BF16) {
                ovf = wg_convert<1>(a, n, w, nullptr);
            } else {
                ovf = wg_convert<0>(a, n, w, nullptr);
            }
            int inf32 = 0;
            if (!ovf) {
                if (ladder == LADDER_BF16) inf32 = wg_getrf<float, 1>(w, n, pv, sh, tile);
                else if (ladder == LADDER_FP16) inf32 = wg_getrf<float, 2>(w, n, pv, sh, tile);
                else inf32 = wg_getrf<float, 0>(w, n, pv, sh, tile);
            }
            if (ovf) it_code = -2;
            else if (inf32) it_code = -3;
            if (!ovf && !inf32) {
                // ||b_c||_inf
                wg_colmax_reset(sh);
                for (int q = t; q < n * nrhs; q += BT)
                    atomicMax(&sh.red[BRHS + q / n],
                              (unsigned long long)__double_as_longlong(fabs(bx[q])));
                __syncthreads();
                if (t < BRHS) sh.bn[t] = red_max(sh, BRHS + t);
                __syncthreads();
                bool skip = false;
                if (POLICY != 0 || out.feat) {
                    const double ainv = wg_inv_norm_est(w, n, rsc, gws + GVEC, sh);
                    f[2] = kappa_feature(anrm * ainv);
                    if (POLICY == 1) skip = (wg_select_all(out.sel, f, sh) & 3) == LADDER_FP64;
                    if (POLICY == 2) skip = !(anrm * ainv <= RULE_KAPPA_MAX);
                }
                if (skip) it_code = -4;
                else if (ladder == LADDER_FP32)
                    it_code = wg_refine_classic(a, bx, w, R + p.b, n, nrhs, anrm, itmax, sh);
                else
                    it_code = wg_refine_gmres(a, bx, w, n, nrhs, anrm, rsc, gws, itmax, inner, sh);
            }
            need64 = it_code < 0;
            __syncthreads();
        }
        if (want_f && out.feat && t < MLP_IN) {
            float v = f[0];
#pragma unroll
            for (int q = 1; q < MLP_IN; ++q) v = (t == q) ? f[q] : v;
            out.feat[(long)p.id * MLP_IN + t] = v;
        }
        int inf = 0;
        if (need64) {
            inf = wg_getrf<double>(a, n, pv, sh, tile);
            if (inf == 0 && nrhs > 0) {
                wg_load_rhs<double>(bx, n, nrhs, nullptr, sh);
                wg_trs<double>(a, n, nrhs, sh);
                for (int q = t; q < n * nrhs; q += BT) {
                    const int c = q / n, i = q % n;
                    bx[q] = sh.y[i * BRHS + c];
                }
            }
        }
        if (t == 0) {
            info[p.id] = inf;
            if (iters) iters[p.id] = it_code;
            if (out.inner) out.inner[p.id] = inner;
            if (out.used) out.used[p.id] = ladder | (tile << 2);
            if (out.tsys) out.tsys[p.id] = wall_clock64() - t_start;
        }
        __syncthreads();
    }
}

// ---- batched GEMM over a flat tile list ---------------------------------------
struct GemmProb {
    long a, b, c;        // offsets (elements); column-major, ld = rows
    int m, n, k;
    int tile0;           // first tile index of this problem
    int cfg;             // tile shape: 0 32x32, 1 64x64, 2 128x64
    int id;              // the caller's index (per-problem timing)
    double alpha, beta;
};

// C tile (TM x TN) of problem p, tile `local` of its own tiles (m-fastest):
// 4 waves of (TM/2 x TN/2), 16x16x4 FP64 MFMAs with operands read from L2
// (no LDS).
template <int TM, int TN>
FLX_DEV void bg_tile(const GemmProb &p, int local, const double *__restrict__ A,
                     const double *__restrict__ Bm, double *__restrict__ Cm)
{
    constexpr int WM = TM / 2, WN = TN / 2, FI = WM / 16, FJ = WN / 16;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    {
        const int tm = (p.m + TM - 1) / TM;
        const int i0 = (local % tm) * TM + (wave & 1) * WM;
        const int j0 = (local / tm) * TN + (wave >> 1) * WN;
        const double *a = A + p.a, *b = Bm + p.b;
        double *c = Cm + p.c;
        double4_t acc[FI][FJ];
#pragma unroll
        for (int x = 0; x < FI; ++x)
#pragma unroll
            for (int y = 0; y < FJ; ++y) acc[x][y] = double4_t{0.0, 0.0, 0.0, 0.0};
        // K in batches of KU MFMA steps (4 each), two batches in flight: the
        // next batch's fragments are requested before this batch's MFMAs
        constexpr int KU = 4;
        double fa[2][KU][FI], fb[2][KU][FJ];
        auto fetch = [&](int buf, int kb) {
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                const int kq = kb + 4 * u + lk;
#pragma unroll
                for (int x = 0; x < FI; ++x) {
                    const int r = i0 + x * 16 + lr;
                    fa[buf][u][x] = (r < p.m && kq < p.k) ? a[(long)kq * p.m + r] : 0.0;
                }
#pragma unroll
                for (int y = 0; y < FJ; ++y) {
                    const int cc = j0 + y * 16 + lr;
                    fb[buf][u][y] = (cc < p.n && kq < p.k) ? b[(long)cc * p.k + kq] : 0.0;
                }
            }
        };
        auto steps = [&](int buf, int kb) {
#pragma unroll
            for (int u = 0; u < KU; ++u) {
                if (kb + 4 * u < p.k) {            // uniform
#pragma unroll
                    for (int x = 0; x < FI; ++x)
#pragma unroll
                        for (int y = 0; y < FJ; ++y)
                            acc[x][y] = mfma_t(fa[buf][u][x], fb[buf][u][y], acc[x][y]);
                }
            }
        };
        fetch(0, 0);
        for (int kk = 0; kk < p.k; kk += 8 * KU) {
            fetch(1, kk + 4 * KU);
            steps(0, kk);
            fetch(0, kk + 8 * KU);
            steps(1, kk + 4 * KU);
        }
        // C: every entry of the tile loaded before any is written
        double cv[FI][FJ][4];
#pragma unroll
        for (int x = 0; x < FI; ++x)
#pragma unroll
            for (int y = 0; y < FJ; ++y) {
                const int col = j0 + y * 16 + lr;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = i0 + x * 16 + lk + 4 * q;
                    cv[x][y][q] = (row < p.m && col < p.n) ? c[(long)col * p.m + row] : 0.0;
                }
            }
#pragma unroll
        for (int x = 0; x < FI; ++x)
#pragma unroll
            for (int y = 0; y < FJ; ++y) {
                const int col = j0 + y * 16 + lr;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = i0 + x * 16 + lk + 4 * q;
                    if (row < p.m && col < p.n)
                        c[(long)col * p.m + row] = p.beta * cv[x][y][q] + p.alpha * acc[x][y][q];
                }
            }
    }
}

// Persistent workgroups take tiles from a flat list over all problems (an
// atomic counter); each tile runs with its problem's shape.  tticks (when
// given) accumulates the in-kernel time of every problem's tiles (100 MHz
// ticks), which the autotuner compares per size bin.
__global__ __launch_bounds__(256) void batched_gemm_kernel(const GemmProb *__restrict__ probs,
                                                           int nprob, int ntiles,
                                                           int *__restrict__ counter,
                                                           const double *__restrict__ A,
                                                           const double *__restrict__ Bm,
                                                           double *__restrict__ Cm,
                                                           unsigned long long *__restrict__ tticks)
{
    __shared__ int s_tile;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) s_tile = atomicAdd(counter, 1);
        __syncthreads();
        const int tile = s_tile;
        __syncthreads();
        if (tile >= ntiles) break;
        // the problem owning this tile: last p with tile0 <= tile
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (probs[mid].tile0 <= tile) lo = mid; else hi = mid - 1;
        }
        const GemmProb p = probs[lo];
        const long long t0 = wall_clock64();
        if (p.cfg == 0) bg_tile<32, 32>(p, tile - p.tile0, A, Bm, Cm);
        else if (p.cfg == 1) bg_tile<64, 64>(p, tile - p.tile0, A, Bm, Cm);
        else bg_tile<128, 64>(p, tile - p.tile0, A, Bm, Cm);
        if (tticks && t == 0) atomicAdd(&tticks[p.id], (unsigned long long)(wall_clock64() - t0));
    }
}

// ---- launchers ------------------------------------------------------------------
__global__ void zero_int_kernel(int *p) { *p = 0; }

size_t batched_prob_bytes() { return sizeof(BatchProb); }
size_t batched_gemm_prob_bytes() { return sizeof(GemmProb); }

hipError_t batched_gesv(hipStream_t s, int policy, const void *probs, int nprob, int *counter,
                        double *A, double *B, float *W32, double *R, double *G, int *ipiv,
                        int *info, int *iters, int itmax, int grid, const BatchOut &out)
{
    if (nprob <= 0) return hipSuccess;
    if (policy == 1 && !out.sel) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
    BatchOut o = out;
    if (policy == 1) {
        if (!out.selT || !out.cls || !out.stats || !G) return hipErrorInvalidValue;
        hipLaunchKernelGGL(stage_selector_kernel, dim3(8), dim3(256), 0, s, out.sel, out.selT);
        o.sel = out.selT;
    }
    if (out.stats)
        hipLaunchKernelGGL(row_stats_kernel, dim3(nprob), dim3(BT), 0, s, (const BatchProb *)probs,
                           nprob, (const double *)A, out.stats,
                           (const float *)(policy == 1 ? out.selT : nullptr), out.cls, out.choice);
#define FLX_BG(m_, l_)                                                                      \
    hipLaunchKernelGGL((batched_gesv_kernel<m_, l_>), dim3(grid), dim3(BT), 0, s,         \
                       (const BatchProb *)probs, nprob, counter, A, B, W32, R, G, ipiv,    \
                       info, iters, itmax, o)
    // without scratch (an FP64-only plan) the kernel holds no low-precision code
    if (policy == 0 && !G) FLX_BG(0, false);
    else if (policy == 0) FLX_BG(0, true);
    else if (policy == 1) {
        // the systems sent to FP64 first, then the low-precision ones
        FLX_BG(1, false);
        hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
        FLX_BG(1, true);
    } else FLX_BG(2, true);
#undef FLX_BG
    return hipGetLastError();
}

size_t batched_scratch_doubles(int grid) { return (size_t)grid * GSTRIDE; }

hipError_t batched_gemm(hipStream_t s, const void *probs, int nprob, int ntiles, int *counter,
                        const double *A, const double *B, double *C, int grid,
                        unsigned long long *tticks)
{
    if (nprob <= 0 || ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(batched_gemm_kernel, dim3(grid), dim3(256), 0, s, (const GemmProb *)probs,
                       nprob, ntiles, counter, A, B, C, tticks);
    return hipGetLastError();
}

int batched_gemm_tile(int cfg, int *tm, int *tn)
{
    static const int TMS[3] = {32, 64, 128}, TNS[3] = {32, 64, 64};
    if (cfg < 0 || cfg > 2) return 0;
    *tm = TMS[cfg];
    *tn = TNS[cfg];
    return 1;
}

} // namespace flx
