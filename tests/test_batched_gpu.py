"""GPU tests of the batched plans (flx_batch_*): thousands of uneven LU
solves and GEMMs in one persistent, graph-captured launch.

Each problem must meet the single-problem bar of the FP64 path: INFO as
LAPACK (flens/lapack/ge/tf2.tcc:101-126), pivots equal to the oracle's
partial pivoting except at near-ties, dget01 / dget02 ratios < 30; GEMMs
within 4 K eps64 (|alpha||A||B| + |beta||C|) elementwise and exact on
integer data, beta = 0 computing 0 * C as the reference's generic gemm.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
TAU = 1e-9


def make_solves(count, seed, nmax=512, singular=(), illcond=()):
    from flens_amd import testmat as T
    rng = np.random.default_rng(seed)
    n = np.clip(np.exp(rng.uniform(0.0, np.log(nmax), count)).round(), 1, nmax).astype(np.int32)
    nrhs = rng.integers(1, 9, count).astype(np.int32)
    A, B = [], []
    for i in range(count):
        a = np.asfortranarray(rng.standard_normal((n[i], n[i])))
        if i in singular:
            a[:, min(2, n[i] - 1)] = 0.0
        if i in illcond and n[i] > 4:
            a = T.svd_matrix(int(n[i]), 1e13, seed + i)
        A.append(a)
        B.append(np.asfortranarray(rng.standard_normal((n[i], nrhs[i]))))
    return n, nrhs, A, B


def pack(mats):
    return np.concatenate([m.ravel(order="F") for m in mats]) if mats else np.zeros(0)


def unpack(vec, shapes):
    out, o = [], 0
    for (r, c) in shapes:
        out.append(np.asfortranarray(vec[o:o + r * c].reshape((c, r)).T))
        o += r * c
    return out


def run_solves(gpu, n, nrhs, A, B, mode, tune=False, reps=1, cfg=None, itmax=5):
    import torch
    dev = "cuda"
    cfgt = None
    if cfg is not None:
        cfgt = torch.from_numpy(np.broadcast_to(np.asarray(cfg, np.int32), (len(n),)).copy()).to(dev)
    At = torch.from_numpy(pack(A)).to(dev)
    Bt = torch.from_numpy(pack(B)).to(dev)
    A0, B0 = At.clone(), Bt.clone()
    ipiv = torch.zeros(int(n.sum()), dtype=torch.int32, device=dev)
    info = torch.zeros(len(n), dtype=torch.int32, device=dev)
    iters = torch.zeros(len(n), dtype=torch.int32, device=dev)
    plan = gpu.batched.gesv_plan(n, nrhs, At, Bt, ipiv, info, iters, mode=mode, tune=tune,
                                 cfg=cfgt, itmax=itmax)
    outs = []
    for _ in range(reps):
        At.copy_(A0)
        Bt.copy_(B0)
        plan.execute()
        torch.cuda.synchronize()
        outs.append((At.cpu().numpy(), Bt.cpu().numpy(), ipiv.cpu().numpy(), info.cpu().numpy(),
                     iters.cpu().numpy()))
    conf = plan.config()
    conf["bin_tiles"] = plan.bin_tiles()
    plan.close()
    return outs, conf


def test_batched_fp64_solves_vs_oracle(gpu):
    count = 300
    sing = {5, 77, 150}
    n, nrhs, A, B = make_solves(count, 11, singular=sing)
    (outs, cfg) = run_solves(gpu, n, nrhs, A, B, mode=0)
    lu_all, x_all, piv_all, info, _ = outs[0]
    lus = unpack(lu_all, [(k, k) for k in n])
    xs = unpack(x_all, [(k, r) for k, r in zip(n, nrhs)])
    po = np.concatenate([[0], np.cumsum(n)])
    for i in range(count):
        a, b, lu, x = A[i], B[i], lus[i], xs[i]
        piv = piv_all[po[i]:po[i + 1]]
        ref = a.copy(order="F")
        rpiv, rinfo, gap = O.dgetrf_gaps(ref)
        assert info[i] == rinfo, (i, n[i], info[i], rinfo)
        d = np.nonzero(piv != rpiv)[0]
        assert d.size == 0 or gap[d[0]] < TAU, (i, n[i], d[:3])
        if n[i] <= 160:
            assert O.dget01(a, lu, piv) < 30, i
        if rinfo == 0:
            assert O.dget02(a, x, b) < 30, (i, n[i])
        else:
            assert np.array_equal(x, b)      # the solve is skipped on INFO > 0


EDGE_N = (1, 2, 15, 16, 17, 31, 32, 33, 47, 48, 63, 64, 65, 95, 96, 97, 127, 160, 255, 256,
          257, 300, 383, 449, 480, 481, 500, 511, 512)


def test_batched_fp64_edge_sizes(gpu):
    """Panel, strip and tile boundaries (multiples of 16 and 32, one row
    per thread up to 256 rows, two above) up to the largest n: the whole
    factor (dget01, so the final row order of every L column), pivots and
    solutions."""
    rng = np.random.default_rng(21)
    n = np.array(EDGE_N, dtype=np.int32)
    nrhs = rng.integers(1, 9, len(n)).astype(np.int32)
    A = [np.asfortranarray(rng.standard_normal((k, k))) for k in n]
    B = [np.asfortranarray(rng.standard_normal((k, r))) for k, r in zip(n, nrhs)]
    (outs, cfg) = run_solves(gpu, n, nrhs, A, B, mode=0)
    lu_all, x_all, piv_all, info, _ = outs[0]
    lus = unpack(lu_all, [(k, k) for k in n])
    xs = unpack(x_all, [(k, r) for k, r in zip(n, nrhs)])
    po = np.concatenate([[0], np.cumsum(n)])
    for i in range(len(n)):
        piv = piv_all[po[i]:po[i + 1]]
        rpiv, rinfo, gap = O.dgetrf_gaps(A[i].copy(order="F"))
        assert info[i] == rinfo == 0, (n[i], info[i], rinfo)
        d = np.nonzero(piv != rpiv)[0]
        assert d.size == 0 or gap[d[0]] < TAU, (n[i], d[:3])
        assert O.dget01(A[i], lus[i], piv) < 30, n[i]
        assert O.dget02(A[i], xs[i], B[i]) < 30, n[i]
    # the FP32 factor + FP64 refinement on the same sizes
    (outs, cfg) = run_solves(gpu, n, nrhs, A, B, mode=1)
    xs = unpack(outs[0][1], [(k, r) for k, r in zip(n, nrhs)])
    for i in range(len(n)):
        a, b, x = A[i], B[i], xs[i]
        r = b - a @ x
        an = np.abs(a).sum(1).max()
        eta = max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
                  for j in range(b.shape[1]))
        assert eta <= n[i] * EPS, (n[i], eta, outs[0][4][i])


def test_batched_fp64_edge_sizes_one_tile_strips(gpu):
    """The FP64 LU with the other tile configuration (one 16-column tile per
    wave strip of the trailing update): the same factors, pivots and
    solutions as the oracle at every boundary size."""
    rng = np.random.default_rng(22)
    n = np.array(EDGE_N, dtype=np.int32)
    nrhs = rng.integers(1, 9, len(n)).astype(np.int32)
    A = [np.asfortranarray(rng.standard_normal((k, k))) for k in n]
    B = [np.asfortranarray(rng.standard_normal((k, r))) for k, r in zip(n, nrhs)]
    (outs, _) = run_solves(gpu, n, nrhs, A, B, mode=gpu.BATCH_TABLE, cfg=4)
    lu_all, x_all, piv_all, info, _ = outs[0]
    lus = unpack(lu_all, [(k, k) for k in n])
    xs = unpack(x_all, [(k, r) for k, r in zip(n, nrhs)])
    po = np.concatenate([[0], np.cumsum(n)])
    for i in range(len(n)):
        piv = piv_all[po[i]:po[i + 1]]
        rpiv, rinfo, gap = O.dgetrf_gaps(A[i].copy(order="F"))
        assert info[i] == rinfo == 0, (n[i], info[i], rinfo)
        d = np.nonzero(piv != rpiv)[0]
        assert d.size == 0 or gap[d[0]] < TAU, (n[i], d[:3])
        assert O.dget01(A[i], lus[i], piv) < 30, n[i]
        assert O.dget02(A[i], xs[i], B[i]) < 30, n[i]


def backward_error(a, b, x):
    r = b - a @ x
    an = np.abs(a).sum(1).max()
    return max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
               for j in range(b.shape[1]))


@pytest.mark.parametrize("mode", ["BATCH_BF16_GMRES", "BATCH_FP16_GMRES", "BATCH_RULE"])
def test_batched_low_precision_ladders(gpu, mode):
    """BF16-operand and FP16-scaled LU + GMRES-IR, and the rule-based ladder,
    both tiles: every system reaches the FP64 backward-error level (through
    the FP64 LU where refinement fails), most by refinement; the
    kappa = 1e13 systems end in the FP64 LU."""
    count = 160
    ill = {3, 40, 41, 90}
    n, nrhs, A, B = make_solves(count, 14, illcond=ill)
    m = getattr(gpu, mode)
    for tile_cfg in (None, 1):
        if tile_cfg is None:
            outs, _ = run_solves(gpu, n, nrhs, A, B, mode=m, itmax=10)
        else:
            ladder = {"BATCH_BF16_GMRES": 2, "BATCH_FP16_GMRES": 3, "BATCH_RULE": 1}[mode]
            if mode == "BATCH_RULE":
                outs, _ = run_solves(gpu, n, nrhs, A, B, mode=m, itmax=10, tune=True)
            else:
                outs, _ = run_solves(gpu, n, nrhs, A, B, mode=gpu.BATCH_TABLE, itmax=10,
                                     cfg=ladder + 4)
        lu_all, x_all, piv_all, info, iters = outs[0]
        xs = unpack(x_all, [(k, r) for k, r in zip(n, nrhs)])
        conv = 0
        for i in range(count):
            assert info[i] == 0, (i, info[i])
            eta = backward_error(A[i], B[i], xs[i])
            assert eta <= n[i] * EPS, (mode, tile_cfg, i, n[i], eta, iters[i])
            if i in ill and n[i] > 4:
                assert iters[i] < 0, (i, iters[i])
            conv += iters[i] >= 0
        # FP16 (unit roundoff 2^-11) leaves Gaussian systems with kappa u
        # near 1 to the FP64 LU more often
        need = 0.8 * count if mode == "BATCH_FP16_GMRES" else count - len(ill) - 10
        assert conv >= need, (mode, tile_cfg, conv)


def test_batched_mixed_solves_and_fallback(gpu):
    count = 200
    ill = {3, 40, 41, 90}
    n, nrhs, A, B = make_solves(count, 12, illcond=ill)
    outs, cfg = run_solves(gpu, n, nrhs, A, B, mode=1)
    lu_all, x_all, piv_all, info, iters = outs[0]
    xs = unpack(x_all, [(k, r) for k, r in zip(n, nrhs)])
    conv = 0
    for i in range(count):
        a, b, x = A[i], B[i], xs[i]
        assert info[i] == 0
        r = b - a @ x
        an = np.abs(a).sum(1).max()
        eta = max(np.abs(r[:, j]).max() / (an * np.abs(x[:, j]).max() + np.abs(b[:, j]).max())
                  for j in range(b.shape[1]))
        assert eta <= n[i] * EPS, (i, n[i], eta, iters[i])
        if i in ill and n[i] > 4:
            assert iters[i] < 0, (i, iters[i])       # refinement cannot reach FP64 accuracy
        conv += iters[i] >= 0
    assert conv >= count - len(ill) - 5


def test_batched_graph_replays_and_tuning(gpu):
    n, nrhs, A, B = make_solves(120, 13)
    outs, cfg = run_solves(gpu, n, nrhs, A, B, mode=0, reps=3)
    for o in outs[1:]:
        for u, v in zip(o, outs[0]):
            assert np.array_equal(u, v)
    outs_t, cfg_t = run_solves(gpu, n, nrhs, A, B, mode=1, tune=True, reps=2)
    assert cfg_t["grid"] > 0 and cfg_t["tuned_ms"] > 0
    assert len(cfg_t["bin_tiles"]) == 6 and set(cfg_t["bin_tiles"]) <= {0, 1}
    assert np.array_equal(outs_t[0][1], outs_t[1][1])


def test_batched_gemm(gpu):
    import torch
    rng = np.random.default_rng(5)
    count = 250
    # sizes over the whole range of the batched GEMM workload (up to 512),
    # with full 512^3 problems and tile-ragged ones
    m, n, k = (np.clip(np.exp(rng.uniform(0, np.log(512), count)).round(), 1, 512).astype(np.int32)
               for _ in range(3))
    m[7] = 0
    k[9] = 0
    m[40:43] = n[40:43] = k[40:43] = 512
    m[43], n[43], k[43] = 511, 509, 512
    m[44], n[44], k[44] = 512, 17, 500
    alpha = rng.uniform(-2, 2, count)
    beta = rng.uniform(-2, 2, count)
    beta[11] = 0.0
    integer = set(range(20, 30))
    As, Bs, Cs = [], [], []
    for i in range(count):
        if i in integer:
            As.append(np.asfortranarray(rng.integers(-8, 9, (m[i], k[i])).astype(float)))
            Bs.append(np.asfortranarray(rng.integers(-8, 9, (k[i], n[i])).astype(float)))
            Cs.append(np.asfortranarray(rng.integers(-8, 9, (m[i], n[i])).astype(float)))
            alpha[i], beta[i] = 2.0, -3.0
        else:
            As.append(np.asfortranarray(rng.uniform(-1, 1, (m[i], k[i]))))
            Bs.append(np.asfortranarray(rng.uniform(-1, 1, (k[i], n[i]))))
            Cs.append(np.asfortranarray(rng.uniform(-1, 1, (m[i], n[i]))))
    Cs[11][:] = np.nan
    At, Bt, Ct = (torch.from_numpy(pack(x)).cuda() for x in (As, Bs, Cs))
    C0 = Ct.clone()
    for tune in (False, True):
        Ct.copy_(C0)
        plan = gpu.batched.gemm_plan(m, n, k, alpha, beta, At, Bt, Ct, tune=tune)
        plan.execute()
        torch.cuda.synchronize()
        out = unpack(Ct.cpu().numpy(), [(a, b) for a, b in zip(m, n)])
        plan.close()
        for i in range(count):
            want = beta[i] * Cs[i] + alpha[i] * (As[i] @ Bs[i])
            if i == 11:
                assert np.isnan(out[i]).all()
                continue
            if i in integer:
                assert np.array_equal(out[i], want), i
                continue
            bound = 4 * max(k[i], 1) * EPS * (abs(alpha[i]) * np.abs(As[i]) @ np.abs(Bs[i])
                                              + abs(beta[i]) * np.abs(Cs[i]))
            assert np.all(np.abs(out[i] - want) <= bound + 1e-300), (i, tune)
