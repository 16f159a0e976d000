"""The learned precision selector: training data from the solver's own runs,
training with the library's HIP kernels, dispatch in the batched solver.

Workflow (all on the GPU):
  1. ``generate``: a batch of systems from a mix of generators -- Gaussian,
     uniform, SVD-constructed with kappa log-uniform in [1e1, 1e12] (Haar
     factors per size, random row signs), diagonally dominant A + n I.
  2. ``collect``: run the batch twice, FP64 (mode 0) and FP32 + refinement
     with FP64 fallback (mode 1), recording every system's in-kernel time,
     the mode-1 outcome and the selector features.  The network decides
     after the FP32 factor (its condition feature needs one), when the
     factor is already paid for: going on with the refinement is right
     (class 1) exactly when it converges, and switching to FP64 at once
     (class 0) saves the refinement steps a failing system would waste.
     Samples are weighted by the size of that stake (n^2 nrhs).
  3. ``train``: Adam on the weighted softmax cross-entropy with the
     library's forward / backward kernels (flx_mlp_grad, flx_mlp_adam).
  4. Dispatch: ``flx_batch_gesv_create(..., mode=2, selector=params)``
     evaluates the network inside the persistent kernel after each FP32
     factor and skips the refinement of systems it sends to FP64.
"""
from __future__ import annotations

import numpy as np

from . import MLP_IN, MLP_OUT, MLP_PARAMS, batched, default_handle, mlp

KINDS = ("gauss", "uniform", "svd", "diagdom")


def generate(count: int, seed: int, nmin: int = 8, nmax: int = 512, mix=(0.35, 0.15, 0.35, 0.15),
             kappa_range=(1e1, 1e12)):
    """Packed device batch: (n, nrhs, kind, A, B) with A, B torch float64."""
    import torch
    rng = np.random.default_rng(seed)
    n = np.clip(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count)).round(), nmin, nmax)
    n = n.astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    kind = rng.choice(len(KINDS), count, p=np.asarray(mix) / np.sum(mix))
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    offs = np.concatenate([[0], np.cumsum(n.astype(np.int64) ** 2)])
    A = torch.empty(int(offs[-1]), dtype=torch.float64, device="cuda")
    haar = {}

    def haar_of(k):
        if k not in haar:
            z = torch.randn(2, k, k, dtype=torch.float64, device="cuda", generator=g)
            q, r = torch.linalg.qr(z)
            haar[k] = q * torch.sign(torch.diagonal(r, dim1=-2, dim2=-1))[:, None, :]
        return haar[k]

    for i in range(count):
        k = int(n[i])
        if kind[i] == 0:
            a = torch.randn(k, k, dtype=torch.float64, device="cuda", generator=g)
        elif kind[i] == 1:
            a = torch.rand(k, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        elif kind[i] == 2:
            kap = float(np.exp(rng.uniform(np.log(kappa_range[0]), np.log(kappa_range[1]))))
            uv = haar_of(k)
            s = torch.from_numpy(kap ** (-np.arange(k) / max(k - 1, 1))).to("cuda")
            sign = torch.from_numpy(rng.choice([-1.0, 1.0], k)).to("cuda")
            a = ((uv[0] * s) @ uv[1].T.contiguous()) * sign[:, None]
        else:
            a = torch.rand(k, k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
            a = a + k * torch.eye(k, dtype=torch.float64, device="cuda")
        A[int(offs[i]):int(offs[i + 1])] = a.T.contiguous().reshape(-1)   # column-major
    B = torch.randn(int((n.astype(np.int64) * nrhs).sum()), dtype=torch.float64, device="cuda",
                    generator=g)
    return n, nrhs, kind, A, B


def run(n, nrhs, A, B, mode, params=None, itmax=10, tune=False):
    """One batched run on copies of A, B; returns per-system info, iters,
    ticks, features, choice and the copies' results."""
    import torch
    count = len(n)
    a, b = A.clone(), B.clone()
    ipiv = torch.zeros(int(n.sum()), dtype=torch.int32, device="cuda")
    info = torch.zeros(count, dtype=torch.int32, device="cuda")
    iters = torch.zeros(count, dtype=torch.int32, device="cuda")
    ticks = torch.zeros(count, dtype=torch.int64, device="cuda")
    feat = torch.zeros((count, MLP_IN), dtype=torch.float32, device="cuda") if mode >= 1 else None
    choice = torch.full((count,), -1, dtype=torch.int32, device="cuda") if mode == 2 else None
    plan = batched.gesv_plan(n, nrhs, a, b, ipiv, info, iters, mode=mode, itmax=itmax, tune=tune,
                             selector=params, features=feat, ticks=ticks, choice=choice)
    plan.execute()
    torch.cuda.synchronize()
    plan.close()
    return dict(info=info.cpu().numpy(), iters=iters.cpu().numpy(), ticks=ticks.cpu().numpy(),
                feat=feat, choice=choice.cpu().numpy() if choice is not None else None,
                X=b, LU=a)


def collect(n, nrhs, A, B, itmax=10):
    """Features, labels and weights from the FP64 and FP32+IR runs."""
    r0 = run(n, nrhs, A, B, 0, itmax=itmax)
    r1 = run(n, nrhs, A, B, 1, itmax=itmax)
    t0 = r0["ticks"].astype(np.float64)
    t1 = r1["ticks"].astype(np.float64)
    label = (r1["iters"] >= 0).astype(np.int32)
    stake = n.astype(np.float64) ** 2 * nrhs
    weight = (stake / stake.mean()).astype(np.float32)
    return r1["feat"], label, weight, t0, t1, r1["iters"]


def train(X, y, w, epochs=200, lr=3e-3, batch=4096, seed=0, params=None, log=None):
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
