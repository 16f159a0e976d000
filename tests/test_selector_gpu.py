"""GPU tests of the learned precision selector: the network's HIP forward,
backward and Adam kernels against the CPU restatement (oracle/mlp.py, itself
pinned by finite differences in test_oracle.py), training on solves the
batched solver runs itself, and the gain of the trained network's dispatch
over the same batch solved without it (FP32 + refinement with FP64 fallback
for every system)."""
import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def test_network_kernels_match_cpu_oracle(gpu):
    import torch
    rng = np.random.default_rng(1)
    p = M.init(7)
    p[M.PARAMS - M.OUT:] = rng.standard_normal(M.OUT).astype(np.float32) * 0.1
    X = rng.standard_normal((333, M.IN)).astype(np.float32)
    y = rng.integers(0, M.OUT, 333).astype(np.int32)
    w = rng.uniform(0.5, 2.0, 333).astype(np.float32)
    pt, Xt = torch.from_numpy(p).cuda(), torch.from_numpy(X).cuda()
    z = gpu.mlp.forward(pt, Xt).cpu().numpy()
    zr, _ = M.forward(p, X)
    assert np.abs(z - zr).max() <= 1e-4 * max(1.0, np.abs(zr).max())
    g, loss = gpu.mlp.grad(pt, Xt, torch.from_numpy(y).cuda(), torch.from_numpy(w).cuda())
    lr_, gr = M.loss_and_grad(p, X, y, w)
    g = g.cpu().numpy()
    assert abs(loss.item() - lr_) <= 1e-4 * abs(lr_)
    assert np.abs(g - gr).max() <= 1e-4 * np.abs(gr).max(), np.abs(g - gr).max()
    m = torch.zeros_like(pt)
    v = torch.zeros_like(pt)
    gt = torch.from_numpy(gr.astype(np.float32)).cuda()
    for step in (1, 2):
        gpu.mlp.adam(pt, gt, m, v, step, lr=1e-2, weight_decay=1e-4)
    pr, mr, vr = p.astype(np.float64), np.zeros(M.PARAMS), np.zeros(M.PARAMS)
    for step in (1, 2):
        pr, mr, vr = M.adam(pr, gr.astype(np.float32).astype(np.float64), mr, vr, step, lr=1e-2,
                            wd=1e-4)
    assert np.abs(pt.cpu().numpy() - pr).max() <= 1e-5


def test_selector_trained_on_own_solves_dispatches_within_best_fixed_policy(gpu):
    """Train on batches solved under every class (4 ladders x 2 tiles), then
    dispatch a held-out batch with the network fused into the persistent
    kernel.  Every system must reach the FP64 backward-error level; the
    learned dispatch's summed in-kernel time must stay within 1 % (timing
    noise) of the best fixed policy (every class for all systems, FP64 and
    the rule-based ladder with autotuned tiles) and at most 3/4 of every
    fixed policy that uses a low-precision ladder: on this kernel the
    low-precision ladders are slower than FP64 for every system, so the
    network has to learn to keep the systems away from them and to pick
    the tile."""
    import torch
    from flens_amd import selector as S
    Xs, ys, ws = [], [], []
    for seed in (101, 102, 103):
        n, nrhs, kind, A, B = S.generate(6000, seed)
        col = S.collect(n, nrhs, A, B)
        X, y, w = S.training_set(col)
        Xs.append(X)
        ys.append(y)
        ws.append(w)
    X = torch.cat(Xs)
    y = np.concatenate(ys)
    w = np.concatenate(ws)
    params, hist = S.train(X, y, w, epochs=40)
    assert hist[-1] < hist[0]
    n, nrhs, kind, A, B = S.generate(4000, 303)
    pol = S.policies(n, nrhs, A, B, params)
    ln = pol["learned"]
    assert (ln["info"] == 0).all()
    eta = S.backward_errors(n, nrhs, A, B, ln["X"])
    assert eta.max() <= 1.0, (int(eta.argmax()), eta.max())
    ticks = {k: float(v["ticks"].astype(np.float64).sum()) for k, v in pol.items()}
    fixed = {k: v for k, v in ticks.items() if k != "learned"}
    lowp = {k: v for k, v in fixed.items()
            if k in ("rule_tuned",) or (k.startswith("class") and int(k[5:]) % 4 != 0)}
    best = min(fixed, key=fixed.get)
    print("summed ticks:", {k: f"{v:.3e}" for k, v in ticks.items()}, "best fixed:", best,
          "chosen ladders:", np.bincount(ln["used"] & 3, minlength=4))
    assert ticks["learned"] <= 1.01 * fixed[best]
    assert ticks["learned"] <= 0.75 * min(lowp.values())
