"""The learned precision / tile selector of the batched solver: training data
from the solver's own runs, training with the library's HIP kernels,
dispatch inside the persistent batch kernel.

A class is ladder + 4 x tile: ladders {FP64 LU, FP32 LU + classic
refinement, BF16-operand LU + GMRES-IR, FP16-scaled LU + GMRES-IR}, tiles
{two, one} 16-column tiles per wave strip of the LU's trailing update.

Workflow (all on the GPU):
  1. ``generate``: batches of systems from a mix of generators -- Gaussian,
     uniform, SVD-constructed with kappa log-uniform in [1e1, 1e12] (Haar
     factors per size, random row signs per system), diagonally dominant
     A + n I -- with n log-uniform in [8, 512] and nrhs uniform in 1..8.
  2. ``collect``: every system is solved under each of the 8 classes
     (``FLX_BATCH_TABLE``), recording its in-kernel time, the refinement
     outcome and the features (before any factor for the FP64 runs; with
     the condition estimate of its own factor for the low-precision ones).
  3. ``training_set``: pre-factor samples (condition feature unknown)
     labelled with the class that solved the system fastest, weighted by
     what a wrong class costs (slowest minus fastest time); post-factor
     samples of each low-precision ladder labelled "go on" (that ladder)
     when its refinement converged, else FP64, weighted by the time at
     stake.
  4. ``train``: Adam on the weighted softmax cross-entropy with the
     library's forward / backward kernels (flx_mlp_grad, flx_mlp_adam).
  5. Dispatch: ``FLX_BATCH_LEARNED`` evaluates the network in the persistent
     kernel before the factor (ladder and tile) and again after a
     low-precision factor (FP64 or go on); refinement that misses the
     tolerance within itmax steps ends in the FP64 LU either way.
"""
from __future__ import annotations

import numpy as np

from . import (BATCH_FP64, BATCH_LEARNED, BATCH_RULE, BATCH_TABLE, MLP_IN, MLP_OUT, MLP_PARAMS,
               batched, mlp)

KINDS = ("gauss", "uniform", "svd", "diagdom")
CLASSES = MLP_OUT             # ladder + 4 x tile


def generate(count: int, seed: int, nmin: int = 8, nmax: int = 512, mix=(0.3, 0.2, 0.3, 0.2),
             kappa_range=(1e1, 1e12)):
    """Packed device batch: (n, nrhs, kind, A, B); A, B torch float64,
    column-major per system.  Matrices of one size are generated together."""
    import torch
    rng = np.random.default_rng(seed)
    n = np.clip(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).round(), nmin, nmax)
    n = n.astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    kind = rng.choice(len(KINDS), count, p=np.asarray(mix) / np.sum(mix)).astype(np.int32)
    kap = np.exp(rng.uniform(np.log(kappa_range[0]), np.log(kappa_range[1]), count))
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    offs = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)])
    A = torch.empty(int(offs[-1]), dtype=torch.float64, device="cuda")
    dev = "cuda"
    for k in np.unique(n):
        k = int(k)
        idx = np.nonzero(n == k)[0]
        c = len(idx)
        kk = torch.from_numpy(kind[idx]).to(dev)
        gauss = torch.randn(c, k, k, dtype=torch.float64, device=dev, generator=g)
        unif = torch.rand(c, k, k, dtype=torch.float64, device=dev, generator=g) * 2 - 1
        a = torch.where((kk == 0)[:, None, None], gauss, unif)
        dd = kk == 3
        if bool(dd.any()):
            a[dd] += k * torch.eye(k, dtype=torch.float64, device=dev)
        sv = np.nonzero(kind[idx] == 2)[0]
        if len(sv):
            z = torch.randn(2, k, k, dtype=torch.float64, device=dev, generator=g)
            q, r = torch.linalg.qr(z)
            q = q * torch.sign(torch.diagonal(r, dim1=-2, dim2=-1))[:, None, :]
            expo = np.arange(k) / max(k - 1, 1)
            s = torch.from_numpy(kap[idx[sv]][:, None] ** (-expo[None, :])).to(dev)
            sign = torch.from_numpy(rng.choice([-1.0, 1.0], (len(sv), k))).to(dev)
            vt = q[1].T.contiguous()
            m = torch.matmul(q[0][None] * s[:, None, :], vt[None].expand(len(sv), k, k).contiguous())
            a[torch.from_numpy(sv).to(dev)] = m * sign[:, :, None]
        # column-major: system i's entry (r, c) at offs[i] + c * k + r
        at = a.transpose(1, 2).contiguous().reshape(c, k * k)
        for j, i in enumerate(idx):
            A[int(offs[i]):int(offs[i + 1])] = at[j]
    B = torch.randn(int((n.astype(np.int64) * nrhs).sum()), dtype=torch.float64, device="cuda",
                    generator=g)
    return n, nrhs, kind, A, B


def run(n, nrhs, A, B, mode, params=None, cfg=None, itmax=10, tune=False, features=True):
    """One batched run on copies of A, B.  Returns per-system info, iters,
    ticks, inner steps, the class each system ran, the features, the
    pre-factor choice (learned mode), the results and the batch time in ms
    (best of two replays)."""
    import torch
    count = len(n)
    a, b = A.clone(), B.clone()
    ipiv = torch.zeros(int(n.sum()), dtype=torch.int32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    iters = torch.zeros(count, dtype=torch.int32, device="cuda")
    ticks = torch.zeros(count, dtype=torch.int64, device="cuda")
    inner = torch.zeros(count, dtype=torch.int32, device="cuda")
    used = torch.zeros(count, dtype=torch.int32, device="cuda")
    feat = torch.zeros((count, MLP_IN), dtype=torch.float32, device="cuda") if features else None
    choice = torch.full((count,), -1, dtype=torch.int32, device="cuda") if mode == BATCH_LEARNED else None
    cfgt = None
    if cfg is not None:
        cfgt = torch.from_numpy(np.broadcast_to(np.asarray(cfg, np.int32), (count,)).copy()).to("cuda")
    plan = batched.gesv_plan(n, nrhs, a, b, ipiv, info, iters, mode=mode, itmax=itmax, tune=tune,
                             selector=params, features=feat, ticks=ticks, choice=choice, cfg=cfgt,
                             inner=inner, used=used)
    ms = 1e30
    for rep in range(2):
        if rep:
            a.copy_(A)
            b.copy_(B)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record()
        plan.execute()
        ev1.record()
        torch.cuda.synchronize()
        ms = min(ms, ev0.elapsed_time(ev1))
    bins = plan.bin_tiles() if tune else None
    plan.close()
    return dict(info=info.cpu().numpy(), iters=iters.cpu().numpy(), ticks=ticks.cpu().numpy(),
                inner=inner.cpu().numpy(), used=used.cpu().numpy(), feat=feat,
                choice=choice.cpu().numpy() if choice is not None else None, X=b, LU=a, ms=ms,
                bins=bins)


def collect(n, nrhs, A, B, itmax=10):
    """Every system under each class: ticks (CLASSES x count), iters,
    features per class (device tensors) and the batch times."""
    out = {"ticks": [], "iters": [], "feat": [], "ms": []}
    for c in range(CLASSES):
        r = run(n, nrhs, A, B, BATCH_TABLE, cfg=c, itmax=itmax)
        if not (r["info"] == 0).all():
            raise RuntimeError(f"class {c}: info != 0 on {(r['info'] != 0).sum()} systems")
        out["ticks"].append(r["ticks"].astype(np.float64))
        out["iters"].append(r["iters"])
        out["feat"].append(r["feat"])
        out["ms"].append(r["ms"])
    out["ticks"] = np.stack(out["ticks"])
    out["iters"] = np.stack(out["iters"])
    return out


def training_set(col):
    """(X, y, w) from collect()'s record: X a float32 device tensor."""
    import torch
    T = col["ticks"]
    count = T.shape[1]
    Xs, ys, ws = [], [], []
    # pre-factor samples: the FP64 run's features (condition unknown)
    stake = T.max(0) - T.min(0)
    Xs.append(col["feat"][0])
    ys.append(T.argmin(0).astype(np.int32))
    ws.append(stake)
    # post-factor samples of each low-precision ladder (tile 0 runs)
    for ladder in (1, 2, 3):
        conv = col["iters"][ladder] >= 0
        Xs.append(col["feat"][ladder])
        ys.append(np.where(conv, ladder, 0).astype(np.int32))
        ws.append(np.abs(T[ladder] - T[0]) + 1.0)
    w = np.concatenate(ws)
    w = (w / w.mean()).astype(np.float32)
    X = torch.cat(Xs)
    y = np.concatenate(ys)
    assert X.shape[0] == y.shape[0] == w.shape[0] == 4 * count
    return X, y, w


def train(X, y, w, epochs=60, lr=3e-3, batch=8192, seed=0, params=None, log=None):
    """Adam on the weighted cross-entropy; X (N x MLP_IN) float32 device
    tensor, y int32 / w float32 numpy.  Returns (params, loss history)."""
    import torch
    rng = np.random.default_rng(seed)
    if params is None:
        params = torch.from_numpy(init_params(seed)).to("cuda")
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    yt = torch.from_numpy(np.asarray(y, np.int32)).to("cuda")
    wt = torch.from_numpy(np.asarray(w, np.float32)).to("cuda")
    N = X.shape[0]
    hist, step = [], 0
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(N)).to("cuda")
        for s0 in range(0, N, batch):
            idx = perm[s0:s0 + batch]
            g, loss = mlp.grad(params, X[idx].contiguous(), yt[idx].contiguous(),
                               wt[idx].contiguous())
            step += 1
            mlp.adam(params, g, m, v, step, lr=lr)
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            g, loss = mlp.grad(params, X, yt, wt)
            hist.append(float(loss.item()))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4f}")
    return params, hist


def init_params(seed: int) -> np.ndarray:
    """He-initialised parameters (zero biases) in the library's layout."""
    rng = np.random.default_rng(seed)
    H = 64
    parts = [rng.standard_normal((H, MLP_IN)) * np.sqrt(2.0 / MLP_IN), np.zeros(H),
             rng.standard_normal((H, H)) * np.sqrt(2.0 / H), np.zeros(H),
             rng.standard_normal((MLP_OUT, H)) * np.sqrt(2.0 / H), np.zeros(MLP_OUT)]
    p = np.concatenate([x.ravel() for x in parts]).astype(np.float32)
    assert p.size == MLP_PARAMS
    return p


def predict(params, X):
    return mlp.forward(params, X).argmax(1).cpu().numpy()


def policies(n, nrhs, A, B, params, itmax=10):
    """The learned dispatch against the fixed policies on one batch: every
    class for all systems, FP64 with autotuned tiles per size bin, the
    rule-based ladder (autotuned tiles).  Returns {name: run record}."""
    out = {}
    for c in range(CLASSES):
        out[f"class{c}"] = run(n, nrhs, A, B, BATCH_TABLE, cfg=c, itmax=itmax, features=False)
    out["fp64_tuned"] = run(n, nrhs, A, B, BATCH_FP64, itmax=itmax, tune=True, features=False)
    out["rule_tuned"] = run(n, nrhs, A, B, BATCH_RULE, itmax=itmax, tune=True, features=False)
    out["learned"] = run(n, nrhs, A, B, BATCH_LEARNED, params=params, itmax=itmax)
    return out


def backward_errors(n, nrhs, A, B, X):
    """max over right-hand sides of ||b - A x||_inf / (||A||_inf ||x||_inf +
    ||b||_inf) / (n 2^-53), per system (host arithmetic)."""
    Ah, Bh, Xh = A.cpu().numpy(), B.cpu().numpy(), X.cpu().numpy()
    out = np.zeros(len(n))
    ao = bo = 0
    for i in range(len(n)):
        k, r = int(n[i]), int(nrhs[i])
        a = Ah[ao:ao + k * k].reshape(k, k).T
        if r:
            b = Bh[bo:bo + k * r].reshape(r, k).T
            x = Xh[bo:bo + k * r].reshape(r, k).T
            res = b - a @ x
            eta = np.abs(res).max(0) / (np.abs(a).sum(1).max() * np.abs(x).max(0) + np.abs(b).max(0))
            out[i] = eta.max() / (k * 2.0 ** -53)
        ao += k * k
        bo += k * r
    return out
