"""Benchmark: dense LU solves on one MI355X.

Primary workload (`value`): plain FP64 LU with partial pivoting + solve,
n = 8192, A i.i.d. U[-1,1) (seed 2), nrhs = 16 with B i.i.d. U[-1,1)
(seed 3), blocked right-looking LU through the library's device entry
flx_dgesv (getrf + getrs).  One step = restore A and B from resident
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
workload (the MFMA trailing-update GEMM), from HIP events recorded around
each of its launches on the stream it runs on, over the timed region.
`cpu_baseline` times the oracle's restatement of the reference's generic
getrf on one core over a bounded sample (the first block step of the same
n = 8192 factorisation) and projects the full trf+trs flop count.
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


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(n: int, nrhs: int, nb: int = 64, target_s: float = 15.0):
    """Oracle (restatement of the reference's generic trf, 1 core) on a
    bounded sample of the same factorisation: the first nb-wide block step
    (panel, interchanges, trsm, trailing GEMM) restricted to the first c
    columns of the n x n matrix, leading dimension n.  In a right-looking
    step each trailing column depends only on the panel, so this is that
    step's work for c columns at the workload's memory layout.  c doubles
    until the sample takes about target_s; the rate is projected to the
    full trf+trs flop count."""
    from oracle import oracle as O
    a_full = O.uniform(n, n, 2)
    c, fl, dt = nb * 2, 0.0, 0.0
    while True:
        a = a_full[:, :c]          # column slice of a Fortran array: same lda = n
        fl = O.getrf_steps_flops(n, c, nb, 1)
        t = time.perf_counter()
        O.dgetrf_steps(a, nb, 1)
        dt = time.perf_counter() - t
        progress(f"CPU sample: {c} columns, {fl / 1e9:.2f} GFLOP in {dt:.1f} s")
        if dt * 2.2 > target_s or c * 2 > n:
            break
        c *= 2
        a_full = O.uniform(n, n, 2)
    rate = fl / dt
    total = lu_flops(n, nrhs)
    return {
        "value": rate / total,
        "unit": "solves/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"first nb={nb} block step of the n={n} getrf (panel + laswp + trsm + "
                   f"trailing GEMM) on the first {c} columns, lda={n}: {fl / 1e9:.2f} GFLOP in "
                   f"{dt:.1f} s = {rate / 1e9:.3f} GFLOP/s on 1 core of {cpu_model()} "
                   f"({os.cpu_count()} visible); projected to the {total / 1e9:.0f} GFLOP "
                   f"trf+trs workload"),
    }


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
                         ("fp16_gmres", dict(prec="fp16", refine="gmres", itermax=3, gmres_m=40))):
            bt, st = 1e30, None
            for rep in range(2 if name == "fp32_auto" else 1):
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
                         "meets_n_eps": eta <= n * eps,
                         "speedup_vs_fp64": rec["fp64"]["ms"] / (1e3 * bt)}
        tot64 += rec["fp64"]["ms"]
        totmx += rec["fp32_auto"]["ms"]
        out["per_kappa"].append(rec)
        progress(f"mixed: kappa={kappa:.0e} fp64 {rec['fp64']['ms']:.1f} ms, "
                 f"fp32 auto {rec['fp32_auto']['ms']:.1f} ms (iter {rec['fp32_auto']['iter']}, "
                 f"inner {rec['fp32_auto']['inner_iters']}), "
                 f"fp32+IR {rec['fp32_classic']['ms']:.1f} ms (iter {rec['fp32_classic']['iter']}), "
                 f"fp16+GMRES {rec['fp16_gmres']['ms']:.1f} ms (iter {rec['fp16_gmres']['iter']})")
    out["job"] = {"fp64_solves_per_s": len(kappas) / (tot64 * 1e-3),
                  "mixed_solves_per_s": len(kappas) / (totmx * 1e-3),
                  "speedup": tot64 / totmx}
    return out


def batched_section(count: int = 20000, seed: int = 4):
    """20,000 systems n ~ logU[8,512], A ~ N(0,1), nrhs ~ U{1..8} (FP64 and
    FP32 + FP64 refinement), and 20,000 GEMMs M,N,K ~ logU[16,512], each
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
    for mode, name in ((0, "fp64"), (1, "fp32_ir")):
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
                     "info_nonzero": int((info.cpu().numpy() != 0).sum()),
                     **({"refinement_failures": int((it < 0).sum()),
                         "mean_iters": float(it[it >= 0].mean()) if (it >= 0).any() else 0.0}
                        if mode == 1 else {})}
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
                   "roofline_ms": 1e3 * groof, "roofline_frac": groof / dt, "config": plan.config()}
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


def selector_section(train_count: int = 3000, eval_count: int = 3000):
    """Train the selector on two self-generated batches and compare the
    learned dispatch with FP32 + refinement for every system on a held-out
    batch (summed in-kernel time)."""
    import numpy as np
    import torch
    from flens_amd import selector as S
    t = time.perf_counter()
    Xs, ys, ws = [], [], []
    for seed in (101, 102):
        n, nrhs, kind, A, B = S.generate(train_count, seed, nmax=256)
        f, lab, wt, t0, t1, it1 = S.collect(n, nrhs, A, B)
        Xs.append(f); ys.append(lab); ws.append(wt)
    params, hist = S.train(torch.cat(Xs), np.concatenate(ys), np.concatenate(ws), epochs=120)
    n, nrhs, kind, A, B = S.generate(eval_count, 303, nmax=256)
    _, lab, _, t0, t1, it1 = S.collect(n, nrhs, A, B)
    r2 = S.run(n, nrhs, A, B, 2, params=params)
    t2 = r2["ticks"].astype(np.float64)
    res = {"train_samples": 2 * train_count, "eval_systems": eval_count,
           "loss_first": hist[0], "loss_last": hist[-1],
           "accuracy": float((np.minimum(r2["choice"], 1) == lab).mean()),
           "ticks_fp32_ir": float(t1.sum()), "ticks_learned": float(t2.sum()),
           "ticks_fp64": float(t0.sum()), "gain_vs_fp32_ir": float(t1.sum() / t2.sum()),
           "seconds": time.perf_counter() - t}
    progress(f"selector: accuracy {res['accuracy']:.3f}, gain x{res['gain_vs_fp32_ir']:.2f}")
    return res


# HBM bytes per launch of the trailing-update GEMM in the default workload
# (n = 8192, nrhs = 16, nb = 128, ib = 32): rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over 875 launches, averaged, scaled by what the same counters
# read for a transpose of known size in the same run (fetched = 2.0 x
# FETCH_SIZE, written = 1.0 x WRITE_SIZE): (2 x 88665.4 + 87862.1) KiB.
LU8192_GEMM_TRAFFIC_BYTES = (2.0 * 88665.4 + 87862.1) * 1024.0


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
    # The roofline kernel's launch times: the same K steps again with HIP
    # events around the trailing-update GEMM launches only, on the streams
    # they run on.  The events add about 2 ms per step, so they stay out of
    # the pass that gives `value`.
    h.profile_start(classes=["gemm"])
    t2 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    gprof = h.profile_stop()["gemm"]
    h.sync()
    progress(f"roofline pass (GEMM events): {1e3 * (t3 - t2) / args.steps:.2f} ms per step")
    # per-kernel breakdown from separate steps: the launch profiler's events
    # cost time of their own, so they stay out of the timed region
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

    fl = lu_flops(n, nrhs)
    g = gprof
    achieved = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
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
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1) generated on the device)",
        "config": {"workload": f"plain FP64 LU + solve (getrf+getrs), n={n}, nrhs={nrhs}, "
                               f"A,B ~ U[-1,1), nb={args.nb}, ib={args.ib}",
                   "n": n, "nrhs": nrhs, "nb": args.nb, "ib": args.ib},
        "tflops": fl / elapsed * args.steps / 1e12,
        "backward_error": eta,
        "backward_error_bound": n * eps64,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_MFMA,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA,
                     "traffic": (LU8192_GEMM_TRAFFIC_BYTES
                                 if (n, nrhs, args.nb, args.ib) == (8192, 16, 128, 32) else None),
                     "kernel": "gemm_mfma_kernel<double,0,0> (trailing update)",
                     "launches": g["launches"], "avg_launch_us": 1e3 * g["ms"] / max(1, g["launches"])},
        "breakdown": breakdown,
    }
    if not args.quick:
        out["mixed_vs_fp64"] = mixed_section(args.mixed_n)
        out["batched"] = batched_section()
        out["selector"] = selector_section()
        out["aux"] = aux_section()
        out["gemm"] = gemm_section()
    if not args.no_cpu:
        progress("CPU baseline: oracle getrf, first block step, 1 core")
        out["cpu_baseline"] = cpu_baseline(n, nrhs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
