"""Benchmark: FP64 LU with partial pivoting + solve on one MI355X.

Workload (the plain FP64 LU configuration): n = 8192, A i.i.d. U[-1,1)
(seed 2), nrhs = 16 with B i.i.d. U[-1,1) (seed 3), blocked right-looking
LU through the library's device entry flx_dgesv (getrf + getrs).  One step
= restore A and B from resident pristine copies (device-to-device) + one
solve.  The result of the last step must reach a normwise backward error
<= n*eps64 or the run fails.

Prints ONE JSON line (see the keys below).  `roofline` is for the dominant
kernel (the MFMA trailing-update GEMM), from HIP events recorded around each
of its launches on the stream it runs on, over the timed region.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--nb", type=int, default=256)
    ap.add_argument("--ib", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              f"(the LU path is single-GPU; see DESIGN.md for how it would shard)",
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
    h.profile_start()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = h.profile_stop()
    h.sync()
    progress(f"timed region done: {1e3 * (t1 - t0) / args.steps:.2f} ms per step; checking the solution")
    elapsed = t1 - t0
    ms = 1e3 * elapsed / args.steps

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
    g = prof["gemm"]
    achieved = g["flops"] / (g["ms"] * 1e-3) / 1e12 if g["ms"] > 0 else 0.0
    breakdown = {k: {"launches_per_step": v["launches"] / args.steps,
                     "ms_per_step": v["ms"] / args.steps,
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
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA, "traffic": None,
                     "kernel": "gemm_mfma_kernel<double,0,0> (trailing update)",
                     "launches": g["launches"], "avg_launch_us": 1e3 * g["ms"] / max(1, g["launches"])},
        "breakdown": breakdown,
    }
    if not args.no_cpu:
        progress("CPU baseline: oracle getrf, first block step, 1 core")
        out["cpu_baseline"] = cpu_baseline(n, nrhs)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
