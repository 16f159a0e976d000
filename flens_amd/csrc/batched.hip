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

// shared memory of one workgroup (the solve buffers double as scratch)
struct BShared {
    double cand[2][4][BPW];        // per-wave pivot candidate rows (as double)
    double cpiv[2][4];             // their pivot entries
    uint32_t khi[2][4], klo[2][4];
    int kpos[2][4];
    double l11[BPW][BPW + 1];
    int perm[BMAX];                // original row now at each position
    int ptmp[BMAX];
    int mv_dst[2 * BPW], mv_src[2 * BPW];
    int nmv;
    int info;
    double y[BMAX * BRHS];         // solve right-hand sides (T-typed view)
    double xs[BMAX * BRHS];        // refinement iterate (FP64), between factorisations
    unsigned long long red[4 * BRHS];
    double rs[BRHS], bn[BRHS];
    int idx;
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

// ---- LU of one system: A (n x n, ld n, column-major) in place -----------------
// ipiv (n ints, 1-based, LAPACK convention); sh.perm receives the row order;
// returns INFO (first exact zero pivot, 1-based, 0 if none).
template <typename T>
FLX_DEV int wg_getrf(T *__restrict__ A, int n, int *__restrict__ ipiv, BShared &sh)
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
        // ---- phase B: the trailing columns in strips of two 16-column
        // tiles, one strip per wave at a time: row moves, U12 = X A12 and
        // A22 -= L21 U12 on 16x16 MFMA tiles (each L21 fragment serves both
        // tiles).  The K index of MFMA step q on lane group lk is
        // kx(q) = 16 (q / 4) + (row of accumulator entry q % 4), so the two
        // U12 accumulator tiles of a column tile are the update's B
        // fragments as they stand.
        if (m2 > 0) {
            const int lr = lane & 15, lk = lane >> 4;
            constexpr int KS = BPW / 4, ST = 2;
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
                }
                // A22 -= L21 U12, 16-row blocks, the next block's L21
                // fragments and C entries requested before this block's MFMAs
                T a[KS], cv[ST][4];
                auto fetch = [&](int ti) {
                    const int ar = c1 + 16 * ti + lr;
                    const bool rok = ti < nt && ar < n;
#pragma unroll
                    for (int q = 0; q < KS; ++q)
                        a[q] = rok ? A[(long)(k + kx(q)) * n + ar] : T(0);
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
template <typename T>
FLX_DEV void wg_load_rhs(const double *__restrict__ src, int n, int nrhs, const double *scale,
                         BShared &sh)
{
    T *y = reinterpret_cast<T *>(sh.y);
    for (int q0 = threadIdx.x; q0 < n * BRHS; q0 += 8 * BT) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + BT * u, i = q / BRHS, c = q % BRHS;
            v[u] = (q < n * BRHS && c < nrhs) ? src[(long)c * n + sh.perm[i]] : 0.0;
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
};

// block-wide max of a non-negative double per right-hand side into sh.red
FLX_DEV void wg_colmax_reset(BShared &sh)
{
    if (threadIdx.x < 4 * BRHS) sh.red[threadIdx.x] = 0ull;
    __syncthreads();
}

// selector features of one system (see the feature definitions below)
FLX_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// In-kernel inference of the selector network (mlp.hip layout) by wave 0:
// lane j computes hidden unit j; returns the arg-max class (uniform).
FLX_DEV int wg_select(const float *__restrict__ P, const float (&f)[MLP_IN])
{
    constexpr int H = MLP_HIDDEN;
    constexpr int oB1 = H * MLP_IN, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H,
                  oB3 = oW3 + MLP_OUT * H;
    const int lane = threadIdx.x & 63;
    float a = P[oB1 + lane];
#pragma unroll
    for (int i = 0; i < MLP_IN; ++i) a = fmaf(P[lane * MLP_IN + i], f[i], a);
    const float h1 = fmaxf(a, 0.0f);
    float b = P[oB2 + lane];
    for (int k = 0; k < H; ++k) b = fmaf(P[oW2 + lane * H + k], readlane_t(h1, k), b);
    const float h2 = fmaxf(b, 0.0f);
    int best = 0;
    float bz = -INFINITY;
    for (int o = 0; o < MLP_OUT; ++o) {
        float z = h2 * P[oW3 + o * H + lane];
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) z += __shfl_xor(z, d, 64);
        z += P[oB3 + o];
        if (z > bz) { bz = z; best = o; }
    }
    return best;
}

template <int MODE>
__global__ __launch_bounds__(BT, 2) void batched_gesv_kernel(const BatchProb *__restrict__ probs,
                                                           int nprob, int *__restrict__ counter,
                                                           double *__restrict__ A,
                                                           double *__restrict__ B,
                                                           float *__restrict__ W32,
                                                           double *__restrict__ R,
                                                           int *__restrict__ ipiv,
                                                           int *__restrict__ info,
                                                           int *__restrict__ iters, int itmax,
                                                           BatchOut out)
{
    __shared__ BShared sh;
    __shared__ int s_choice;
    const int t = threadIdx.x;
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
        double *a = A + p.a;
        double *bx = B + p.b;
        int *pv = ipiv + p.piv;
        int it_code = 0;
        bool need64 = (MODE == 0);
        if (MODE >= 1) {
            float *w = W32 + p.a;
            double *r = R + p.b;
            // FP32 copy (overflow -> fall back, as dsgesv's ITER = -2) in the
            // same pass as the row sums of |A| (||A||_inf and the row
            // statistics of the features): rows t and t + BT (n <= 2 BT),
            // 16 columns per batch of loads
            wg_colmax_reset(sh);
            if (t < 2) sh.red[3 * BRHS + 1 + t] = (unsigned long long)__double_as_longlong(DBL_MAX);
            __syncthreads();
            int ovf = 0;
            {
                double rs[2] = {0.0, 0.0}, dg[2] = {0.0, 0.0};
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
                                const double av = fabs(v[h][u]);
                                ovf |= av > (double)FLT_MAX;
                                w[(long)j * n + i] = (float)v[h][u];
                                rs[h] += av;
                                if (j == i) dg[h] = av;
                            }
                        }
                }
                double an = 0.0, amin = DBL_MAX, dmin = DBL_MAX;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (t + BT * h < n) {
                        an = fmax(an, rs[h]);
                        amin = fmin(amin, rs[h]);
                        dmin = fmin(dmin, dg[h] / fmax(rs[h] - dg[h], 1e-300));
                    }
                }
                atomicMax(&sh.red[0], (unsigned long long)__double_as_longlong(an));
                atomicMin(&sh.red[3 * BRHS + 1], (unsigned long long)__double_as_longlong(amin));
                atomicMin(&sh.red[3 * BRHS + 2], (unsigned long long)__double_as_longlong(dmin));
            }
            ovf = __syncthreads_or(ovf);
            int inf32 = ovf ? 1 : wg_getrf<float>(w, n, pv, sh);
            if (ovf) it_code = -2;
            else if (inf32) it_code = -3;
            if (!ovf && !inf32) {
                // ||b_c||_inf
                for (int q = t; q < n * nrhs; q += BT)
                    atomicMax(&sh.red[BRHS + q / n],
                              (unsigned long long)__double_as_longlong(fabs(bx[q])));
                __syncthreads();
                const double anrm = __longlong_as_double((long long)sh.red[0]);
                const double rmin = __longlong_as_double((long long)sh.red[3 * BRHS + 1]);
                const double ddom = __longlong_as_double((long long)sh.red[3 * BRHS + 2]);
                if (t < BRHS) sh.bn[t] = __longlong_as_double((long long)sh.red[BRHS + t]);
                __syncthreads();
                bool skip_ir = false;
                if (MODE == 2 || out.feat) {
                    // ||A^{-1}||_inf >= ||A^{-1} e||_inf for e = 1 and e = sign(A^{-1} 1)
                    float *y = reinterpret_cast<float *>(sh.y);
                    for (int q = t; q < n * BRHS; q += BT) y[q] = (q % BRHS == 0) ? 1.0f : 0.0f;
                    __syncthreads();
                    wg_trs<float>(w, n, 1, sh);
                    wg_colmax_reset(sh);
                    for (int i = t; i < n; i += BT) {
                        const float v = y[i * BRHS];
                        atomicMax(&sh.red[0], (unsigned long long)__double_as_longlong(fabs((double)v)));
                        sh.xs[i] = (v >= 0.0f) ? 1.0 : -1.0;
                    }
                    __syncthreads();
                    for (int q = t; q < n * BRHS; q += BT)
                        y[q] = (q % BRHS == 0) ? (float)sh.xs[sh.perm[q / BRHS]] : 0.0f;
                    __syncthreads();
                    wg_trs<float>(w, n, 1, sh);
                    for (int i = t; i < n; i += BT)
                        atomicMax(&sh.red[0], (unsigned long long)__double_as_longlong(fabs((double)y[i * BRHS])));
                    __syncthreads();
                    const double ainv = __longlong_as_double((long long)sh.red[0]);
                    __syncthreads();
                    // features: log2 n / 9, nrhs / 8, log10 kappa_inf / 16,
                    // log10 (diagonal dominance) / 4, log10 (row-norm ratio) / 12
                    float f[MLP_IN];
                    f[0] = log2f((float)n) / 9.0f;
                    f[1] = (float)nrhs / 8.0f;
                    const double kap = anrm * ainv;
                    f[2] = clampf((float)log10(fmax(kap, 1.0)), 0.0f, 20.0f) / 16.0f;
                    f[3] = clampf((float)log10(fmax(ddom, 1e-300)), -4.0f, 4.0f) / 4.0f;
                    f[4] = clampf((float)log10(anrm / fmax(rmin, 1e-300)), 0.0f, 12.0f) / 12.0f;
                    if (!(kap <= DBL_MAX)) f[2] = 20.0f / 16.0f;
                    if (out.feat && t < MLP_IN) out.feat[(long)p.id * MLP_IN + t] = f[t];
                    if (MODE == 2) {
                        int c = 1;
                        if (t < 64) c = wg_select(out.sel, f);
                        if (t == 0) s_choice = c;
                        __syncthreads();
                        c = s_choice;
                        if (out.choice && t == 0) out.choice[p.id] = c;
                        skip_ir = (c == 0);
                    }
                }
                const double tol = sqrt((double)n) * 1.1102230246251565e-16;
                // x0 = M^{-1} b ; x kept in sh.xs (FP64)
                wg_load_rhs<float>(bx, n, nrhs, nullptr, sh);
                wg_trs<float>(w, n, nrhs, sh);
                for (int q = t; q < n * BRHS; q += BT)
                    sh.xs[q] = (double)reinterpret_cast<float *>(sh.y)[q];
                __syncthreads();
                bool conv = false;
                int it = 0;
                for (; it <= itmax && !skip_ir; ++it) {
                    // r = b - A x, and the norms
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
                                r[(long)c * n + i] = acc[c];
                                atomicMax(&sh.red[c], (unsigned long long)__double_as_longlong(fabs(acc[c])));
                                atomicMax(&sh.red[2 * BRHS + c],
                                          (unsigned long long)__double_as_longlong(fabs(sh.xs[i * BRHS + c])));
                            }
                        }
                    }
                    __syncthreads();
                    bool all = true;
#pragma unroll
                    for (int c = 0; c < BRHS; ++c) {
                        const double rn = __longlong_as_double((long long)sh.red[c]);
                        const double xn = __longlong_as_double((long long)sh.red[2 * BRHS + c]);
                        if (c < nrhs) all = all && (rn <= tol * (anrm * xn + sh.bn[c]));
                    }
                    if (t < BRHS) {
                        const double rn = __longlong_as_double((long long)sh.red[t]);
                        sh.rs[t] = (rn > 0.0 && rn <= DBL_MAX) ? 1.0 / rn : 1.0;
                    }
                    __syncthreads();
                    if (all) { conv = true; break; }
                    if (it == itmax) break;
                    // x += M^{-1} r  (r scaled by 1/||r|| into FP32 range and back)
                    wg_load_rhs<float>(r, n, nrhs, sh.rs, sh);
                    wg_trs<float>(w, n, nrhs, sh);
                    for (int q = t; q < n * BRHS; q += BT) {
                        const int c = q % BRHS;
                        sh.xs[q] += (double)reinterpret_cast<float *>(sh.y)[q] / sh.rs[c];
                    }
                    __syncthreads();
                }
                if (conv) {
                    for (int q = t; q < n * nrhs; q += BT) {
                        const int c = q / n, i = q % n;
                        bx[q] = sh.xs[i * BRHS + c];
                    }
                    it_code = it;
                } else {
                    it_code = skip_ir ? -4 : -itmax - 1;   // -4: the selector chose FP64
                }
            }
            need64 = it_code < 0;
            __syncthreads();
        }
        int inf = 0;
        if (need64) {
            inf = wg_getrf<double>(a, n, pv, sh);
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
    double alpha, beta;
};

// C tile (TM x TN) per 256-thread workgroup pass; 4 waves of (TM/2 x TN/2),
// 16x16x4 FP64 MFMAs with operands read from L2 (no LDS).
template <int TM, int TN>
__global__ __launch_bounds__(256) void batched_gemm_kernel(const GemmProb *__restrict__ probs,
                                                           int nprob, int ntiles,
                                                           int *__restrict__ counter,
                                                           const double *__restrict__ A,
                                                           const double *__restrict__ Bm,
                                                           double *__restrict__ Cm)
{
    constexpr int WM = TM / 2, WN = TN / 2, FI = WM / 16, FJ = WN / 16;
    __shared__ int s_tile;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
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
        const int tm = (p.m + TM - 1) / TM;
        const int local = tile - p.tile0;
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

// ---- launchers ------------------------------------------------------------------
__global__ void zero_int_kernel(int *p) { *p = 0; }

size_t batched_prob_bytes() { return sizeof(BatchProb); }
size_t batched_gemm_prob_bytes() { return sizeof(GemmProb); }

hipError_t batched_gesv(hipStream_t s, int mode, const void *probs, int nprob, int *counter,
                        double *A, double *B, float *W32, double *R, int *ipiv, int *info,
                        int *iters, int itmax, int grid, const BatchOut &out)
{
    if (nprob <= 0) return hipSuccess;
    if (mode == 2 && !out.sel) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
#define FLX_BG(m_)                                                                          \
    hipLaunchKernelGGL(batched_gesv_kernel<m_>, dim3(grid), dim3(BT), 0, s,               \
                       (const BatchProb *)probs, nprob, counter, A, B, W32, R, ipiv, info, \
                       iters, itmax, out)
    if (mode == 0) FLX_BG(0);
    else if (mode == 1) FLX_BG(1);
    else FLX_BG(2);
#undef FLX_BG
    return hipGetLastError();
}

hipError_t batched_gemm(hipStream_t s, int tile_cfg, const void *probs, int nprob, int ntiles,
                        int *counter, const double *A, const double *B, double *C, int grid)
{
    if (nprob <= 0 || ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(1), 0, s, counter);
    switch (tile_cfg) {
    case 0:
        hipLaunchKernelGGL((batched_gemm_kernel<32, 32>), dim3(grid), dim3(256), 0, s,
                           (const GemmProb *)probs, nprob, ntiles, counter, A, B, C);
        break;
    case 1:
        hipLaunchKernelGGL((batched_gemm_kernel<64, 64>), dim3(grid), dim3(256), 0, s,
                           (const GemmProb *)probs, nprob, ntiles, counter, A, B, C);
        break;
    default:
        hipLaunchKernelGGL((batched_gemm_kernel<128, 64>), dim3(grid), dim3(256), 0, s,
                           (const GemmProb *)probs, nprob, ntiles, counter, A, B, C);
        break;
    }
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
