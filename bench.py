"""Benchmark: dense LU solves on one MI355X.

Primary workload (`value`): plain FP64 LU with partial pivoting + solve,
n = 8192, A i.i.d. U[-1,1) (seed 2), nrhs = 16 with B i.i.d. U[-1,1)
(seed 3), blocked right-looking LU through the library's device entry
flx_dgesv (getrf with B carried along as extra columns, so its
interchanges and block updates leave L^-1 P B there, then the U solve of
getrs).  One step = restore A and B from resident
pristine copies (device-to-device) + one solve.  The result of the last step
must reach a normwise backward error <= n*eps64 or the run fails.

Also measured in the same run (unless --quick), each in its own JSON key:
  mixed_vs_fp64  mixed-precision LU + FP64 refinement against plain FP64 LU
                 at n = 16384, A = U diag(sigma) V^T with kappa in
                 {1e2, 1e6, 1e10}: times, backward / forward errors, speedups;
  batched        20,000 uneven solves (FP64, FP32 + refinement) and 20,000
                 uneven GEMMs, one persistent graph-captured launch each;
  selector       the learned precision selector against the same batch
                 solved without it.

Prints ONE JSON line.  `roofline` is for the dominant kernel of the primary
workload (the cooperative panel factorisation, HBM-bound: algorithmic bytes
16 * rows * width per launch over its average launch time), from HIP events
recorded around each of its launches on the stream it runs on, over a
second pass of the timed steps; `roofline_gemm` gives the trailing-update
GEMM the same way.  `peaks_measured` holds the on-device probes (MFMA
loops, streaming copy) that `frac_of_probe` divides by.  `nb_sweep` times
the same workload at nb in {64, 128, 256}.  `vs_baseline` carries the
metric's ratio: mixed+IR over plain FP64 solves/s on the n = 16384 job.
`cpu_baseline` is the oracle's restatement of the reference's generic
getrf + getrs on one core, projected from block steps sampled at several
depths with their full trailing width (cpu_trf_projection).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s & TFLOP/s at backward err <= n*eps64: mixed+IR vs plain FP64 LU"
PEAK_FP64_MFMA = 78.6     # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0         # GB/s
PEAK_MFMA = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "fp16": 2500.0}   # dense TFLOP/s


def lu_flops(n: int, nrhs: int) -> float:
    return 2.0 / 3.0 * n ** 3 - 0.5 * n ** 2 + 5.0 / 6.0 * n + nrhs * (2.0 * n * n - n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


_T0 = time.perf_counter()


def progress(msg: str):
    print(f"[bench {time.perf_counter() - _T0:6.1f}s] {msg}", file=sys.stderr, flush=True)


def cpu_trf_projection(n: int, nb: int, depths: int = 3, rows=(64, 256)):
    """Projected one-core time of the oracle's blocked getrf (a restatement
    of the reference's generic trf, flens/lapack/ge/trf.tcc:59-152) on an
    n x n matrix with block size nb.

    The block step at depth j (m = n - j rows, r = n - j - jb trailing
    columns) is the panel tf2 on m x jb plus the interchanges, the
    unit-lower trsm of the jb x r block row and the trailing GEMM.  The
    generic GEMM (gemm.tcc: column-major mapped to row-major, one gemv per
    row of C) re-reads the whole jb x r block row for every row of C, so
    its per-row cost depends on r and the cache that block fits in, and a
    sample must keep the full width r.  At a few depths j:
      P(m)   the panel alone, on the full m x jb (a[j:, j:j+jb]);
      U(ms)  a step on a[j:j+ms, j:] (all r trailing columns, ms rows) less
             the panel on its ms rows, for two ms: the trsm and GEMM of the
             step, linear in the rows, extrapolated to m.
    The sum over all n/nb steps interpolates P and U linearly in m.  The
    interchanges of the already-factored columns (laswp on cols0) are not
    sampled.  Returns (seconds, flops sampled, seconds sampled, text)."""
    import numpy as np
    from oracle import oracle as O
    a = O.uniform(n, n, 2)
    starts = list(range(0, n, nb))
    pick = sorted({starts[int(round(q * (len(starts) - 1)))] for q in np.linspace(0, 0.9, depths)})
    fl_s = dt_s = 0.0

    def run(i0, j0, r, c):
        nonlocal fl_s, dt_s
        t0 = time.perf_counter()
        O.dgetrf_step_at(a, i0, j0, r, c, nb)
        dt = time.perf_counter() - t0
        fl_s += O.getrf_steps_flops(r, c, nb, 1)
        dt_s += dt
        return dt

    ms, P, U = [], [], []
    for j in pick:
        progress(f"CPU sample nb={nb}: depth j={j}")
        m = n - j
        jb = min(nb, m)
        p_full = run(j, j, m, jb)
        u = []
        for extra in rows:
            r = min(jb + extra, m)
            t_step = run(j, j, r, n - j)
            t_pan = run(j, j, r, jb)
            u.append((r, max(0.0, t_step - t_pan)))
        (r1, u1), (r2, u2) = u
        h = max(0.0, (u2 - u1) / (r2 - r1)) if r2 > r1 else 0.0
        ms.append(m)
        P.append(p_full)
        U.append(max(0.0, u1 + (m - r1) * h))
    ms_a = np.array(ms[::-1], dtype=float)
    P_a, U_a = np.array(P[::-1]), np.array(U[::-1])
    total = 0.0
    for j in starts:
        m = n - j
        pm = float(np.interp(m, ms_a, P_a))
        # the update shrinks with r * m below the last sampled depth
        um = float(np.interp(m, ms_a, U_a)) if m >= ms_a[0] else U_a[0] * (m * max(0, m - nb)) / (ms_a[0] * max(1.0, ms_a[0] - nb))
        if m > ms_a[0] or j == starts[-1]:
            pass
        total += pm + (um if m > nb else 0.0)
    desc = (f"nb={nb}: at depths j in {pick}, the panel on the full (n-j) x nb and the "
            f"step's trsm + GEMM on {rows[0]} and {rows[1]} rows below the panel over all "
            f"n-j columns (lda={n}), interpolated over the {len(starts)} block steps")
    return total, fl_s, dt_s, desc


def cpu_trs_time(n: int, nrhs: int):
    """One-core time of the oracle's getrs (trs.tcc:84-92: laswp, unit-lower
    and upper triangular solves) for nrhs right-hand sides, from a solve with
    one right-hand side at the full n (each right-hand side column is solved
    independently by the same loops)."""
    from oracle import oracle as O
    import numpy as np
    a = O.uniform(n, n, 2)
    a[np.arange(n), np.arange(n)] += 4.0       # any nonsingular triangle: the time does not depend on values
    ipiv = np.arange(1, n + 1, dtype=np.int32)
    b = O.uniform(n, 1, 3)
    t0 = time.perf_counter()
    O.dgetrs(0, a, ipiv, b)
    return nrhs * (time.perf_counter() - t0)


def cpu_baseline(n: int, nrhs: int, nbs=(64, 128, 256), headline_nb: int = 128):
    """The oracle (restatement of the reference's generic trf + trs, one
    core) on a bounded sample of the C2 workload: block steps at several
    depths per nb (cpu_trf_projection) and one right-hand side of the solve,
    projected to the full n = 8192 trf + trs.  Returns the cpu_baseline
    object (for headline_nb) and the per-nb projections."""
    import threading
    per_nb, res, errs = {}, {}, []

    def one(nb):
        try:
            # the headline blocking sampled at three depths, the others at two
            # with smaller row samples (the nb = 256 steps are the costliest)
            res[nb] = (cpu_trf_projection(n, nb) if nb == headline_nb else
                       cpu_trf_projection(n, nb, depths=2, rows=(32, 128)))
        except Exception as e:      # re-raised below
            errs.append(e)
    # one thread per nb, each a single-core measurement (the oracle's C code
    # runs with the interpreter lock released); the solve is timed after them
    threads = [threading.Thread(target=one, args=(nb,), daemon=True) for nb in nbs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errs:
        raise errs[0]
    t_trs = cpu_trs_time(n, nrhs)
    fl_tot = dt_tot = 0.0
    for nb in nbs:
        t_trf, fl_s, dt_s, desc = res[nb]
        fl_tot += fl_s
        dt_tot += dt_s
        per_nb[nb] = {"trf_s": t_trf, "trs_s": t_trs, "solves_per_s": 1.0 / (t_trf + t_trs),
                      "sample": desc}
        progress(f"CPU projection nb={nb}: trf {t_trf:.0f} s + trs {t_trs:.1f} s "
                 f"(sampled {fl_s / 1e9:.2f} GFLOP in {dt_s:.1f} s)")
    h = per_nb[headline_nb]
    return {
        "value": h["solves_per_s"],
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "projected_trf_s": h["trf_s"],
        "projected_trs_s": t_trs,
        "sample": (f"{h['sample']}; trs from one right-hand side x {nrhs}; "
                   f"all nb together sampled {fl_tot / 1e9:.2f} GFLOP in {dt_tot:.1f} s, each nb "
                   f"on 1 core of {cpu_model()} ({os.cpu_count()} visible; the nb samples in "
                   f"parallel threads beside the GPU sections)"),
    }, per_nb


def mixed_section(n: int, kappas=(1e2, 1e6, 1e10), seed: int = 1234):
    """Mixed-precision LU + FP64 refinement against plain FP64 LU on
    A = U diag(sigma) V^T (Haar U, V; sigma geometric 1 -> 1/kappa), nrhs = 1,
    b = A x_true.  Mixed = FP32 factor with the solver's automatic ladder
    (condition estimate -> classic refinement, GMRES-IR or FP64); the dsgesv
    method (classic refinement, FP64 fallback) and FP16-factor GMRES-IR are
    timed beside it.  Times include everything (factor, estimate,
    refinement, fallback)."""
    import torch
    import flens_amd as F
    from flens_amd import testmat as T
    # plain FP64 at its best blocking for n = 16384 (nb = 256; nb = 128 is
    # the better choice at n = 8192)
    F.default_handle().set_lu_blocking(256, 32)
    t = time.perf_counter()
    mats = T.svd_matrix_device(n, list(kappas), seed)
    torch.cuda.synchronize()
    progress(f"mixed: generated n={n} test matrices in {time.perf_counter() - t:.1f} s")
    fl = 2.0 / 3.0 * n ** 3 + 2.0 * n * n
    ffl = 2.0 / 3.0 * n ** 3          # the factorisation's flops
    eps = 2.0 ** -53
    out = {"n": n, "nrhs": 1, "flops_per_solve": fl, "per_kappa": []}
    A2 = F.DeviceMatrix.empty(n, n)
    X = F.DeviceMatrix.empty(n, 1)

    def errors(A, Xt, B):
        a, x, b = A.storage.T, X.storage.T, B.storage.T
        r = b - a @ x
        eta = (r.abs().max() / (a.abs().sum(1).max() * x.abs().max() + b.abs().max())).item()
        fwd = ((X.storage - Xt.storage).abs().max() / Xt.storage.abs().max()).item()
        return eta, fwd

    tot64 = totmx = 0.0
    for kappa, A, Xt, B in mats:
        rec = {"kappa": kappa}
        best = 1e30
        for rep in range(2):
            A2.storage.copy_(A.storage)
            X.storage.copy_(B.storage)
            torch.cuda.synchronize()
            t = time.perf_counter()
            piv, info = F.lapack.sv(A2, X)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        eta, fwd = errors(A, Xt, B)
        rec["fp64"] = {"ms": 1e3 * best, "solves_per_s": 1.0 / best, "tflops": fl / best / 1e12,
                       "backward_error": eta, "forward_error": fwd}
        for name, kw in (("fp32_auto", dict(prec="fp32", refine="auto", itermax=30, gmres_m=40)),
                         ("fp32_classic", dict(prec="fp32", refine="classic", itermax=30)),
                         ("bf16_gmres", dict(prec="bf16", refine="gmres", itermax=3, gmres_m=40)),
                         ("fp16_gmres", dict(prec="fp16", refine="gmres", itermax=3, gmres_m=40))):
            bt, st = 1e30, None
            for rep in range(2):
                X.storage.zero_()
                A2.storage.copy_(A.storage)
                torch.cuda.synchronize()
                t = time.perf_counter()
                st = F.lapack.sv_mixed(A2, B, X, fallback=True, **kw)
                torch.cuda.synchronize()
                bt = min(bt, time.perf_counter() - t)
            eta, fwd = errors(A, Xt, B)
            rec[name] = {"ms": 1e3 * bt, "solves_per_s": 1.0 / bt, "tflops_fp64_equiv": fl / bt / 1e12,
                         "backward_error": eta, "forward_error": fwd, "iter": st["iter"],
                         "inner_iters": st["inner_iters"], "fallback": st["fallback"],
                         "factor_ms": st["factor_ms"], "refine_ms": st["refine_ms"],
                         "kappa_est": st["kappa_est"],
                         "factor_tflops": ffl / (st["factor_ms"] * 1e-3) / 1e12 if st["factor_ms"] > 0 else 0.0,
                         "factor_frac_of_mfma_peak": (ffl / (st["factor_ms"] * 1e-3) / 1e12 /
                                                      PEAK_MFMA[kw["prec"]]) if st["factor_ms"] > 0 else 0.0,
                         "meets_n_eps": eta <= n * eps,
                         "speedup_vs_fp64": rec["fp64"]["ms"] / (1e3 * bt)}
        tot64 += rec["fp64"]["ms"]
        totmx += rec["fp32_auto"]["ms"]
        out["per_kappa"].append(rec)
        progress(f"mixed: kappa={kappa:.0e} fp64 {rec['fp64']['ms']:.1f} ms, "
                 f"fp32 auto {rec['fp32_auto']['ms']:.1f} ms (iter {rec['fp32_auto']['iter']}, "
                 f"inner {rec['fp32_auto']['inner_iters']}), "
                 f"fp32+IR {rec['fp32_classic']['ms']:.1f} ms (iter {rec['fp32_classic']['iter']}), "
                 f"bf16+GMRES {rec['bf16_gmres']['ms']:.1f} ms (iter {rec['bf16_gmres']['iter']}), "
                 f"fp16+GMRES {rec['fp16_gmres']['ms']:.1f} ms (iter {rec['fp16_gmres']['iter']})")
    out["job"] = {"fp64_solves_per_s": len(kappas) / (tot64 * 1e-3),
                  "mixed_solves_per_s": len(kappas) / (totmx * 1e-3),
                  "speedup": tot64 / totmx}
    return out


def batched_section(count: int = 20000, seed: int = 4):
    """20,000 systems n ~ logU[8,512], A ~ N(0,1), nrhs ~ U{1..8} under each
    precision ladder (FP64; FP32 + classic refinement; BF16 and FP16-scaled
    LU + GMRES-IR; the rule-based ladder), with the LU tile autotuned per
    size bin, and 20,000 GEMMs M,N,K ~ logU[16,512], each
    batch one persistent graph-captured launch with an autotuned shape.
    Roofline per problem: max(flops / 78.6 TF, bytes / 8 TB/s), summed."""
    import numpy as np
    import torch
    import flens_amd as F
    rng = np.random.default_rng(seed)
    n = np.clip(np.exp(rng.uniform(np.log(8), np.log(512), count)).round(), 8, 512).astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    nd = n.astype(np.float64)
    na = int((n.astype(np.int64) ** 2).sum())
    nb = int((n.astype(np.int64) * nrhs).sum())
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    A = torch.randn(na, dtype=torch.float64, device="cuda", generator=g)
    B = torch.randn(nb, dtype=torch.float64, device="cuda", generator=g)
    A0, B0 = A.clone(), B.clone()
    ipiv = torch.zeros(int(n.sum()), dtype=torch.int32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    iters = torch.zeros(count, dtype=torch.int32, device="cuda")
    fl_i = 2.0 / 3.0 * nd ** 3 + 2.0 * nd ** 2 * nrhs
    by_i = 8.0 * (2.0 * nd ** 2 + 2.0 * nd * nrhs)
    roof = float(np.maximum(fl_i / (PEAK_FP64_MFMA * 1e12), by_i / (PEAK_HBM * 1e9)).sum())
    out = {"count": count, "flops": float(fl_i.sum()), "roofline_ms": 1e3 * roof}
    for mode, name in ((F.BATCH_FP64, "fp64"), (F.BATCH_FP32_IR, "fp32_ir"),
                       (F.BATCH_BF16_GMRES, "bf16_gmres"), (F.BATCH_FP16_GMRES, "fp16_gmres"),
                       (F.BATCH_RULE, "rule")):
        plan = F.batched.gesv_plan(n, nrhs, A, B, ipiv, info, iters, mode=mode, itmax=10, tune=True)
        ts = []
        for rep in range(3):
            A.copy_(A0)
            B.copy_(B0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            plan.execute()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        it = iters.cpu().numpy()
        dt = min(ts)
        out[name] = {"ms": 1e3 * dt, "solves_per_s": count / dt, "tflops": float(fl_i.sum()) / dt / 1e12,
                     "roofline_frac": roof / dt, "config": plan.config(),
                     "bin_tiles": plan.bin_tiles(),
                     "info_nonzero": int((info.cpu().numpy() != 0).sum()),
                     **({"refinement_failures": int((it < 0).sum()),
                         "mean_iters": float(it[it >= 0].mean()) if (it >= 0).any() else 0.0}
                        if mode != F.BATCH_FP64 else {})}
        plan.close()
        progress(f"batched {name}: {1e3 * dt:.1f} ms, {count / dt:.0f} solves/s")
    del A, B, A0, B0
    m, nn, k = (np.clip(np.exp(rng.uniform(np.log(16), np.log(512), count)).round(), 16, 512)
                .astype(np.int32) for _ in range(3))
    GA = torch.randn(int((m.astype(np.int64) * k).sum()), dtype=torch.float64, device="cuda", generator=g)
    GB = torch.randn(int((k.astype(np.int64) * nn).sum()), dtype=torch.float64, device="cuda", generator=g)
    GC = torch.randn(int((m.astype(np.int64) * nn).sum()), dtype=torch.float64, device="cuda", generator=g)
    gfl_i = 2.0 * m.astype(np.float64) * nn * k
    gby_i = 8.0 * (m.astype(np.float64) * k + k.astype(np.float64) * nn + 2.0 * m * nn)
    groof = float(np.maximum(gfl_i / (PEAK_FP64_MFMA * 1e12), gby_i / (PEAK_HBM * 1e9)).sum())
    plan = F.batched.gemm_plan(m, nn, k, np.full(count, 1.0), np.full(count, 0.5), GA, GB, GC, tune=True)
    ts = []
    for rep in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.execute()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t)
    dt = min(ts)
    out["gemm"] = {"count": count, "ms": 1e3 * dt, "tflops": float(gfl_i.sum()) / dt / 1e12,
                   "roofline_ms": 1e3 * groof, "roofline_frac": groof / dt, "config": plan.config(),
                   "bin_tiles": plan.bin_tiles()}
    plan.close()
    progress(f"batched gemm: {1e3 * dt:.1f} ms")
    return out


def gemm_section(reps: int = 20, seed: int = 7):
    """The DGEMM workload C = 1.5 A B - 0.5 C at M = N = K = 2048 (A, B, C
    U[-1,1), column-major, ld = 2048) and an LU-shaped trailing update
    C(7936 x 7936) -= A(7936 x 256) B(256 x 7936).  A 256 MiB device copy
    runs first: it is the calibration dispatch for HBM-traffic counters."""
    import torch
    import flens_amd as F
    h = F.default_handle()
    L = F.lib()
    cal = torch.empty(32 << 20, dtype=torch.float64, device="cuda")
    cal2 = torch.empty_like(cal)
    cal2.copy_(cal)
    out = {}
    for name, (m, n, k, alpha, beta) in {"dgemm2048": (2048, 2048, 2048, 1.5, -0.5),
                                         "trailing7936x256": (7936, 7936, 256, -1.0, 1.0)}.items():
        A, B, Cm = F.DeviceMatrix.empty(m, k), F.DeviceMatrix.empty(k, n), F.DeviceMatrix.empty(m, n)
        for i, X in enumerate((A, B, Cm)):
            assert L.flx_fill_uniform(h.ptr, F._vp(X.ptr), X.m, X.n, X.ld, seed + i, 0) == 0
        F.blas.mm(0, 0, alpha, A, B, beta, Cm)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            F.blas.mm(0, 0, alpha, A, B, beta, Cm)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        fl = 2.0 * m * n * k
        out[name] = {"m": m, "n": n, "k": k, "ms": 1e3 * dt, "tflops": fl / dt / 1e12,
                     "frac_fp64_mfma": fl / dt / 1e12 / PEAK_FP64_MFMA,
                     "algorithmic_GB": 8.0 * (m * k + k * n + 2.0 * m * n) / 1e9}
    progress(f"gemm: 2048^3 {out['dgemm2048']['ms']:.3f} ms "
             f"({out['dgemm2048']['tflops']:.1f} TF/s), trailing 7936^2x256 "
             f"{out['trailing7936x256']['ms']:.3f} ms ({out['trailing7936x256']['tflops']:.1f} TF/s)")
    return out


def aux_section(n: int = 8192, seed: int = 5):
    """Norms, equilibration, condition estimate, refinement and the expert
    driver at n = 8192 (one right-hand side), A = U[-1,1) with rows scaled
    over 1e-6..1e6 (so equilibration has work to do).  Matrix passes are
    priced against HBM: lan reads A once (8 n^2 bytes), equ twice; con and
    rfs are dominated by triangular solves and are reported as times."""
    import torch
    import flens_amd as F
    h = F.default_handle()
    L = F.lib()
    A0 = F.DeviceMatrix.empty(n, n)
    assert L.flx_fill_uniform(h.ptr, F._vp(A0.ptr), n, n, n, seed, 0) == 0
    A0.storage.mul_(torch.logspace(-6, 6, n, dtype=torch.float64, device="cuda")[None, :])
    B0 = F.DeviceMatrix.empty(n, 1)
    assert L.flx_fill_uniform(h.ptr, F._vp(B0.ptr), n, 1, n, seed + 1, 0) == 0
    out = {"n": n, "nrhs": 1}
    byt = 8.0 * n * n

    def timed(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t)
        return best, r

    for nm in "MOIF":
        t, v = timed(lambda: F.lapack.lan(nm, A0))
        passes = 2 if nm == "F" else 1
        out[f"lan_{nm}"] = {"ms": 1e3 * t, "value": v,
                            "GB_per_s": passes * byt / t / 1e9,
                            "hbm_frac": passes * byt / t / 1e9 / PEAK_HBM}
    t, e = timed(lambda: F.lapack.equ(A0))
    out["equ"] = {"ms": 1e3 * t, "GB_per_s": 2 * byt / t / 1e9,
                  "hbm_frac": 2 * byt / t / 1e9 / PEAK_HBM, "info": e[5]}
    LU = F.DeviceMatrix.empty(n, n)
    LU.storage.copy_(A0.storage)
    piv, info = F.lapack.trf(LU)
    torch.cuda.synchronize()
    anorm = F.lapack.lan("O", A0)
    t, rc = timed(lambda: F.lapack.con("O", LU, anorm))
    out["con"] = {"ms": 1e3 * t, "rcond": rc}
    X = F.DeviceMatrix.empty(n, 1)

    def rfs_once():
        X.storage.copy_(B0.storage)
        F.lapack.trs(0, LU, piv, X)
        return F.lapack.rfs(0, A0, LU, piv, B0, X)
    t, fb = timed(rfs_once)
    out["rfs"] = {"ms": 1e3 * t, "ferr": float(fb[0][0]), "berr": float(fb[1][0])}
    A1, B1, AF = F.DeviceMatrix.empty(n, n), F.DeviceMatrix.empty(n, 1), F.DeviceMatrix.empty(n, n)
    p1 = torch.zeros(n, dtype=torch.int32, device="cuda")
    r = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros(n, dtype=torch.float64, device="cuda")

    def svx_once():
        A1.storage.copy_(A0.storage)
        B1.storage.copy_(B0.storage)
        return F.lapack.svx("E", 0, A1, AF, p1, "N", r, c, B1, X)
    t, d = timed(svx_once, reps=2)
    out["svx_E"] = {"ms": 1e3 * t, "equed": d["equed"], "rcond": d["rcond"],
                    "ferr": float(d["ferr"][0]), "berr": float(d["berr"][0]), "info": d["info"]}
    # getri: inv(A) from the factor (LAPACK's count 4/3 n^3 flops)
    Ainv = F.DeviceMatrix.empty(n, n)

    def tri_once():
        Ainv.storage.copy_(LU.storage)
        return F.lapack.tri(Ainv, piv)
    t, info_t = timed(tri_once, reps=2)
    out["tri"] = {"ms": 1e3 * t, "info": info_t,
                  "tflops": 4.0 / 3.0 * n ** 3 / t / 1e12}
    progress(f"aux: lan O {out['lan_O']['ms']:.2f} ms, equ {out['equ']['ms']:.2f} ms, "
             f"con {out['con']['ms']:.1f} ms, rfs {out['rfs']['ms']:.1f} ms, "
             f"svx(E) {out['svx_E']['ms']:.1f} ms, tri {out['tri']['ms']:.1f} ms")
    return out


def selector_section(train_batches: int = 2, train_count: int = 12500, eval_count: int = 20000):
    """Train the selector on train_batches x train_count self-generated
    systems, each solved under all 8 classes (4 precision ladders x 2
    tiles), and compare the learned dispatch on a held-out batch with every
    fixed policy (summed in-kernel time and batch wall time)."""
    import numpy as np
    import torch
    from flens_amd import selector as S
    t = time.perf_counter()
    Xs, ys, ws = [], [], []
    for b in range(train_batches):
        n, nrhs, kind, A, B = S.generate(train_count, 101 + b)
        col = S.collect(n, nrhs, A, B)
        X, y, w = S.training_set(col)
        Xs.append(X); ys.append(y); ws.append(w)
        del A, B, col
        progress(f"selector: collected batch {b + 1}/{train_batches} ({train_count} systems x 8 classes)")
    X = torch.cat(Xs)
    params, hist = S.train(X, np.concatenate(ys), np.concatenate(ws), epochs=30)
    t_train = time.perf_counter() - t
    n, nrhs, kind, A, B = S.generate(eval_count, 303)
    pol = S.policies(n, nrhs, A, B, params)
    ln = pol["learned"]
    eta = S.backward_errors(n, nrhs, A, B, ln["X"])
    ticks = {k: float(v["ticks"].astype(np.float64).sum()) for k, v in pol.items()}
    wall = {k: v["ms"] for k, v in pol.items()}
    fixed = {k: v for k, v in ticks.items() if k != "learned"}
    best = min(fixed, key=fixed.get)
    res = {"systems_solved_for_training": train_batches * train_count * 8,
           "training_systems": train_batches * train_count, "samples": int(X.shape[0]),
           "eval_systems": eval_count, "loss_first": hist[0], "loss_last": hist[-1],
           "ticks": ticks, "batch_ms": wall, "best_fixed_policy": best,
           "learned_vs_best_fixed_ticks": ticks["learned"] / fixed[best],
           "learned_vs_best_fixed_ms": wall["learned"] / wall[best],
           "learned_ladders": np.bincount(ln["used"] & 3, minlength=4).tolist(),
           "learned_tiles": np.bincount(ln["used"] >> 2, minlength=2).tolist(),
           "fp64_tuned_bin_tiles": pol["fp64_tuned"]["bins"],
           "max_backward_error_over_n_eps": float(eta.max()),
           "train_seconds": t_train, "seconds": time.perf_counter() - t}
    progress(f"selector: learned / best fixed ({best}) = {res['learned_vs_best_fixed_ticks']:.3f} "
             f"(ticks), {res['learned_vs_best_fixed_ms']:.3f} (wall)")
    return res


# HBM bytes per launch of the trailing-update GEMMs in the default workload
# (n = 8192, nrhs = 16, nb = 128, ib = 32): rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE summed over the GEMM kernels of 7 factorisations (188 update
# launches each: the next block's first panel columns on the one-wave-per-
# block kernel, its other columns and the rest on the tiled kernel), per
# launch, scaled by what the same counters read for a transpose of known
# size in the same run (fetched = 2.0 x FETCH_SIZE, written = 1.0 x
# WRITE_SIZE): (2 x 42110.9 + 58748.7) KiB.
LU8192_GEMM_TRAFFIC_BYTES = (2.0 * 42110.9 + 58748.7) * 1024.0
# HBM bytes per panel launch in the default workload: rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE over 1792 launches, averaged (1292.55 and 1180.30
# KiB), scaled by what the same counters read for the 1 GiB streaming copy of
# the peak probe in the same run (fetched = 1.93618 x FETCH_SIZE, written =
# 0.976469 x WRITE_SIZE).
LU8192_PANEL_TRAFFIC_BYTES = (1.93618 * 1292.55 + 0.976469 * 1180.30) * 1024.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--nb", type=int, default=128)
    ap.add_argument("--ib", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--quick", action="store_true",
                    help="only the FP64 LU workload (no mixed / batched / selector sections)")
    ap.add_argument("--mixed-n", type=int, default=16384)
    ap.add_argument("--section", choices=["mixed", "batched", "selector", "aux", "gemm"],
                    help="run only this section and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.section:
        import torch
        assert torch.cuda.is_available(), "bench.py needs a GPU"
        fn = {"mixed": lambda: mixed_section(args.mixed_n), "batched": batched_section,
              "selector": selector_section, "aux": aux_section,
              "gemm": gemm_section}[args.section]
        print(json.dumps({args.section: fn()}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(the LU path runs on one GPU; no multi-GPU path is built)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import flens_amd as F

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, nrhs = args.n, args.nrhs
    h = F.default_handle()
    h.set_lu_blocking(args.nb, args.ib)
    L = F.lib()
    A0 = F.DeviceMatrix.empty(n, n)
    B0 = F.DeviceMatrix.empty(n, nrhs)
    A = F.DeviceMatrix.empty(n, n)
    B = F.DeviceMatrix.empty(n, nrhs)
    assert L.flx_fill_uniform(h.ptr, F._vp(A0.ptr), n, n, n, 2, 0) == 0
    assert L.flx_fill_uniform(h.ptr, F._vp(B0.ptr), n, nrhs, n, 3, 0) == 0

    def step():
        A.storage.copy_(A0.storage)
        B.storage.copy_(B0.storage)
        piv, info = F.lapack.sv(A, B)
        if info != 0:
            raise RuntimeError(f"dgesv info={info}")

    progress(f"inputs resident (n={n}, nrhs={nrhs}); warmup {args.warmup} steps")
    for _ in range(args.warmup):
        step()
    progress(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    h.sync()
    progress(f"timed region done: {1e3 * (t1 - t0) / args.steps:.2f} ms per step")
    elapsed = t1 - t0
    ms = 1e3 * elapsed / args.steps
    # The roofline kernels' launch times: the same K steps again with HIP
    # events around the panel and trailing-update GEMM launches only, on the
    # streams they run on.  The events cost time of their own, so they stay
    # out of the pass that gives `value`.
    h.profile_start(classes=["panel", "gemm"])
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    rprof = h.profile_stop()
    h.sync()
    progress(f"roofline pass (panel + GEMM events): {1e3 * (t3 - t2) / args.steps:.2f} ms per step")
    # per-kernel breakdown from separate steps
    prof_steps = max(1, min(args.steps, 5))
    progress(f"profiling {prof_steps} more steps (HIP events per launch); checking the solution")
    h.profile_start()
    for _ in range(prof_steps):
        step()
    torch.cuda.synchronize()
    prof = h.profile_stop()
    h.sync()

    # backward error of the last solve: ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)
    a0 = A0.storage.T
    b0 = B0.storage.T
    x = B.storage.T
    r = b0 - a0 @ x
    eta = (r.abs().max(0).values /
           (a0.abs().sum(1).max() * x.abs().max(0).values + b0.abs().max(0).values)).max().item()
    eps64 = 2.0 ** -53
    if not eta <= n * eps64:
        raise RuntimeError(f"backward error {eta:.3e} above n*eps64 = {n * eps64:.3e}")

    # The CPU baseline (one core, the oracle's C code: ctypes releases the
    # interpreter lock during each call) runs in a thread beside the GPU
    # sections that follow, which leave the host cores idle.
    cpu_res, cpu_thread = {}, None
    if not args.no_cpu:
        import threading

        def cpu_job():
            try:
                cpu_res["result"] = cpu_baseline(
                    n, nrhs, headline_nb=args.nb if args.nb in (64, 128, 256) else 128)
            except Exception as e:      # re-raised by the main thread
                cpu_res["error"] = e
        progress("CPU baseline: oracle getrf + getrs, sampled block steps, 1 core (background)")
        cpu_thread = threading.Thread(target=cpu_job, daemon=True)
        cpu_thread.start()
    # the nb sweep of the same workload (the reference's block-size knob,
    # ilaenv's GETRF block size)
    sweep = []
    if not args.quick:
        for nb in (64, 128, 256):
            h.set_lu_blocking(nb, args.ib)
            for _ in range(2):
                step()
            k = max(3, args.steps // 2)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            for _ in range(k):
                step()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - ts) / k
            sweep.append({"nb": nb, "ib": args.ib, "steps": k, "ms_per_step": 1e3 * dt,
                          "solves_per_s": 1.0 / dt, "tflops": lu_flops(n, nrhs) / dt / 1e12})
            progress(f"nb sweep: nb={nb} {1e3 * dt:.2f} ms per step")
        h.set_lu_blocking(args.nb, args.ib)

    peaks = F.probe_peaks(h)
    progress("on-device peaks: " + ", ".join(f"{k} {v:.1f}" for k, v in peaks.items()))

    fl = lu_flops(n, nrhs)
    g, pn = rprof["gemm"], rprof["panel"]
    g_ach = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    p_ach = pn["bytes"] / (pn["ms"] * 1e-3) / 1e9 if pn["ms"] > 0 else 0.0
    headline = (n, nrhs, args.nb, args.ib) == (8192, 16, 128, 32)
    breakdown = {k: {"launches_per_step": v["launches"] / prof_steps,
                     "ms_per_step": v["ms"] / prof_steps,
                     **({"GB_per_s": v["bytes"] / (v["ms"] * 1e-3) / 1e9} if v["ms"] > 0 and v["bytes"] > 0 else {})}
                 for k, v in prof.items() if v["launches"]}
    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "vs_baseline_is": ("mixed+IR solves/s / plain FP64 LU solves/s over the n=16384 "
                           "kappa in {1e2,1e6,1e10} job (mixed_vs_fp64.job.speedup); "
                           "BASELINE.md publishes no number for this metric"),
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) generated on the device)",
        "config": {"workload": f"plain FP64 LU + solve (gesv: getrf carrying B, then the U "
                               f"solve), n={n}, nrhs={nrhs}, "
                               f"A,B ~ U[-1,1), nb={args.nb}, ib={args.ib}",
                   "n": n, "nrhs": nrhs, "nb": args.nb, "ib": args.ib},
        "tflops": fl / elapsed * args.steps / 1e12,
        "frac_fp64_peak": fl / elapsed * args.steps / 1e12 / PEAK_FP64_MFMA,
        "backward_error": eta,
        "backward_error_bound": n * eps64,
        "roofline": {"bound": "hbm", "achieved": p_ach, "peak": PEAK_HBM, "unit": "GB/s",
                     "frac": p_ach / PEAK_HBM,
                     "traffic": LU8192_PANEL_TRAFFIC_BYTES if headline else None,
                     "kernel": "panel_kernel<256,2,G> (cooperative panel factorisation)",
                     "algorithmic_bytes_per_launch": pn["bytes"] / max(1, pn["launches"]),
                     "launches": pn["launches"], "avg_launch_us": 1e3 * pn["ms"] / max(1, pn["launches"]),
                     "share_of_step": pn["ms"] / args.steps / ms,
                     "frac_of_probe": p_ach / peaks["hbm_copy_gbps"]},
        "roofline_gemm": {"bound": "mfma", "achieved": g_ach, "peak": PEAK_FP64_MFMA,
                          "unit": "TFLOP/s", "frac": g_ach / PEAK_FP64_MFMA,
                          "traffic": LU8192_GEMM_TRAFFIC_BYTES if headline else None,
                          "kernel": "gemm_mfma_kernel<double,0,0,P> + gemm_skinny_kernel (trailing updates)",
                          "launches": g["launches"], "avg_launch_us": 1e3 * g["ms"] / max(1, g["launches"]),
                          "frac_of_probe": g_ach / peaks["fp64_mfma_tflops"]},
        "peaks_measured": peaks,
        "breakdown": breakdown,
    }
    if sweep:
        out["nb_sweep"] = sweep
    if not args.quick:
        out["mixed_vs_fp64"] = mixed_section(args.mixed_n)
        out["vs_baseline"] = out["mixed_vs_fp64"]["job"]["speedup"]
        out["batched"] = batched_section()
        out["selector"] = selector_section()
        out["aux"] = aux_section()
        out["gemm"] = gemm_section()
    if cpu_thread is not None:
        progress("waiting for the CPU baseline")
        cpu_thread.join()
        if "error" in cpu_res:
            raise cpu_res["error"]
        cb, per_nb = cpu_res["result"]
        out["cpu_baseline"] = cb
        for rec in out.get("nb_sweep", []):
            c = per_nb.get(rec["nb"])
            if c:
                rec["cpu_projected_s"] = c["trf_s"] + c["trs_s"]
                rec["gpu_vs_cpu"] = rec["solves_per_s"] * (c["trf_s"] + c["trs_s"])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
