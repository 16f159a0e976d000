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


def test_trained_selector_beats_dispatch_without_it(gpu):
    import torch
    from flens_amd import selector as S
    Xs, ys, ws = [], [], []
    for seed in (101, 102):
        n, nrhs, kind, A, B = S.generate(3000, seed, nmax=256)
        f, lab, wt, t0, t1, it1 = S.collect(n, nrhs, A, B)
        Xs.append(f)
        ys.append(lab)
        ws.append(wt)
    X = torch.cat(Xs)
    y = np.concatenate(ys)
    w = np.concatenate(ws)
    params, hist = S.train(X, y, w, epochs=120, lr=3e-3, batch=2048)
    assert hist[-1] < hist[0]
    # held-out batch: learned dispatch (mode 2) against FP32+IR for all (mode 1)
    n, nrhs, kind, A, B = S.generate(3000, 303, nmax=256)
    _, lab, _, t0, t1, it1 = S.collect(n, nrhs, A, B)
    r2 = S.run(n, nrhs, A, B, 2, params=params)
    t2 = r2["ticks"].astype(np.float64)
    assert (r2["info"] == 0).all()
    # every system still reaches the FP64 backward-error level
    Xh = r2["X"].cpu().numpy()
    Ah = A.cpu().numpy()
    Bh = B.cpu().numpy()
    ao = bo = 0
    for i in range(len(n)):
        k, r = int(n[i]), int(nrhs[i])
        a = Ah[ao:ao + k * k].reshape(k, k).T
        b = Bh[bo:bo + k * r].reshape(r, k).T
        x = Xh[bo:bo + k * r].reshape(r, k).T
        res = b - a @ x
        eta = (np.abs(res).max(0) / (np.abs(a).sum(1).max() * np.abs(x).max(0) + np.abs(b).max(0))).max()
        assert eta <= k * 2.0 ** -53, (i, k, eta)
        ao += k * k
        bo += k * r
    acc = float((np.minimum(r2["choice"], 1) == lab).mean())
    print(f"selector accuracy {acc:.3f}; summed ticks: fp64 {t0.sum():.3e} fp32+ir {t1.sum():.3e} "
          f"learned {t2.sum():.3e}")
    assert acc >= 0.9
    assert 0.0 < lab.mean() < 1.0          # both outcomes occur in the held-out batch
    assert t2.sum() < t1.sum()
