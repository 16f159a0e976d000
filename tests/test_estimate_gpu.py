"""GPU parity for the norm, equilibration, condition-estimate, refinement and
expert-driver entry points (flx_dlange, flx_dgeequ, flx_dlaqge, flx_dgecon,
flx_dgerfs, flx_dgesvx and their Fortran seams) against the oracle.

Bar: bit-exact for the max norm, the 1- and infinity-norms (summed in the
reference's order), the equilibration scalings and laq's scaling of A;
floating point within stated tolerances where the device sums in parallel
(Frobenius norm, residuals) or solves with a differently rounded factor:
  * Frobenius norm: relative 1e-14;
  * rcond: relative 1e-9 of the oracle's on the same factor (lacn2 takes the
    same path unless two candidates tie to rounding);
  * refinement: the refined X within (n+1) eps ||x|| kappa of the oracle's,
    berr <= 4 n eps, ferr >= the true forward error and within a factor 2 of
    the oracle's bound.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53


def rnd(m, n, seed, scale_rows=False):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((m, n))
    if scale_rows:
        a *= np.logspace(-6, 6, m)[:, None]
    return np.asfortranarray(a)


def same(x, y):
    x, y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
    return x.shape == y.shape and np.array_equal(x.view(np.int64), y.view(np.int64))


def dev(gpu, a):
    return gpu.DeviceMatrix.from_numpy(a)


@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (64, 64), (200, 130), (130, 200), (1000, 1000)])
def test_lange_matches_oracle(gpu, shape):
    a = rnd(*shape, 3)
    for variant in range(3):
        b = a.copy(order="F")
        if variant == 1:
            b[0, -1] = np.nan
        if variant == 2:
            b[-1, 0] = np.inf
        d = dev(gpu, b)
        for nm in "MOI":
            assert same(gpu.lapack.lan(nm, d), O.dlange(nm, b)), (shape, variant, nm)
        f, fo = gpu.lapack.lan("F", d), O.dlange("F", b)
        assert (np.isnan(f) and np.isnan(fo)) or f == fo or abs(f - fo) <= 1e-14 * fo
    c = a.copy(order="F")
    c[0, 0], c[-1, -1] = np.inf, -np.inf
    if c.size > 1:
        assert np.isnan(gpu.lapack.lan("F", dev(gpu, c)))   # lassq's inf / inf
    assert gpu.lapack.lan("M", gpu.DeviceMatrix.empty(0, 5)) == 0.0


def test_lange_seam(gpu):
    a = rnd(33, 21, 4)
    for nm in "MOIF":
        assert gpu.lapack.lan(nm, a) == gpu.lapack.lan(nm, dev(gpu, a))


@pytest.mark.parametrize("shape", [(50, 50), (31, 77), (300, 90), (1024, 1024)])
def test_equ_laq_bit_exact(gpu, shape):
    import torch
    a = rnd(*shape, 10 + shape[0], scale_rows=True)
    r, c, rc, cc, am, info = gpu.lapack.equ(dev(gpu, a))
    ro, co, rco, cco, amo, infoo = O.dgeequ(a)
    assert info == infoo == 0
    assert same(r.cpu().numpy(), ro) and same(c.cpu().numpy(), co)
    assert same(rc, rco) and same(cc, cco) and same(am, amo)
    d = dev(gpu, a)
    eq = gpu.lapack.laq(d, r, c, rc, cc, am)
    b = a.copy(order="F")
    assert eq == O.dlaqge(b, ro, co, rco, cco, amo)
    assert same(d.numpy(), b)
    # seam
    r2, c2, rc2, cc2, am2, info2 = gpu.lapack.equ(a)
    assert same(r2, ro) and same(c2, co) and info2 == 0
    torch.cuda.synchronize()


def test_equ_zero_row_column(gpu):
    a = rnd(6, 5, 5)
    a[3, :] = 0.0
    assert gpu.lapack.equ(dev(gpu, a))[5] == O.dgeequ(a)[5] == 4
    b = rnd(6, 5, 6)
    b[:, 2] = 0.0
    assert gpu.lapack.equ(dev(gpu, b))[5] == O.dgeequ(b)[5] == 9


@pytest.mark.parametrize("n", [1, 2, 37, 300, 1500])
def test_con_matches_oracle(gpu, n):
    a = rnd(n, n, 40 + n)
    lu = a.copy(order="F")
    piv, info = O.dgetrf(lu)
    d = dev(gpu, lu)
    for nm in "OI":
        an = O.dlange(nm, a)
        rc, rco = gpu.lapack.con(nm, d, an), O.dgecon(nm, lu, an)
        assert abs(rc - rco) <= 1e-9 * rco, (n, nm, rc, rco)
        assert gpu.lapack.con(nm, lu, an) == rc          # the dgecon_ seam


def test_con_diagonal_and_singular(gpu):
    dvals = np.array([3.0, -0.5, 8.0, 0.25, -2.0])
    a = np.asfortranarray(np.diag(dvals))
    for nm in "OI":
        rc = gpu.lapack.con(nm, dev(gpu, a), O.dlange(nm, a))
        assert abs(rc - 0.25 / 8.0) <= 4 * EPS
    u = np.asfortranarray(np.triu(rnd(6, 6, 31)))
    u[4, 4] = 0.0
    assert gpu.lapack.con("O", dev(gpu, u), 1.0) == 0.0 == O.dgecon("O", u, 1.0)


@pytest.mark.parametrize("n,tr", [(1, 0), (5, 1), (120, 0), (500, 1), (2000, 0)])
def test_rfs_matches_oracle(gpu, n, tr):
    rng = np.random.default_rng(60 + n)
    a = rnd(n, n, 61 + n)
    xt = np.asfortranarray(rng.standard_normal((n, 3)))
    b = np.asfortranarray((a.T if tr else a) @ xt)
    lu = a.copy(order="F")
    piv, _ = O.dgetrf(lu)
    x0 = b.copy(order="F")
    O.dgetrs(tr, lu, piv, x0)
    xo = x0.copy(order="F")
    fo, bo = O.dgerfs(tr, a, lu, piv, b, xo)
    import torch
    pd = torch.from_numpy(piv.astype(np.int32)).cuda()
    X = dev(gpu, x0)
    f, be = gpu.lapack.rfs(tr, dev(gpu, a), dev(gpu, lu), pd, dev(gpu, b), X)
    x = X.numpy()
    kappa = np.linalg.cond(a, 1)
    scale = np.abs(xo).max(axis=0)
    assert np.all(np.abs(x - xo).max(axis=0) <= 4 * (n + 1) * EPS * kappa * scale)
    assert np.all(be <= 4 * n * EPS + 1e-300)
    fwd = np.abs(x - xt).max(axis=0) / np.abs(x).max(axis=0)
    assert np.all(fwd <= f), (fwd, f)
    assert np.all(f <= 2 * fo) and np.all(fo <= 2 * f), (f, fo)
    # the dgerfs_ seam
    x2 = x0.copy(order="F")
    f2, b2 = gpu.lapack.rfs(tr, a, lu, piv, b, x2)
    assert np.all(fwd <= 2 * f2)


@pytest.mark.parametrize("n", [1, 4, 60, 400])
def test_svx_matches_oracle(gpu, n):
    import torch
    a0 = rnd(n, n, 100 + n, scale_rows=True)
    rng = np.random.default_rng(101 + n)
    xt = rng.standard_normal((n, 2))
    for fact in "NE":
        for tr in (0, 1):
            b0 = np.asfortranarray((a0.T if tr else a0) @ xt)
            # oracle
            ao, bo = a0.copy(order="F"), b0.copy(order="F")
            afo, pivo = np.zeros((n, n), order="F"), np.zeros(n, dtype=np.int32)
            ro, co = np.zeros(n), np.zeros(n)
            do = O.dgesvx(fact, tr, ao, afo, pivo, "N", ro, co, bo)
            # device
            A, B = dev(gpu, a0), dev(gpu, b0)
            AF, X = gpu.DeviceMatrix.empty(n, n), gpu.DeviceMatrix.empty(n, 2)
            piv = torch.zeros(max(1, n), dtype=torch.int32, device="cuda")
            r = torch.zeros(max(1, n), dtype=torch.float64, device="cuda")
            c = torch.zeros(max(1, n), dtype=torch.float64, device="cuda")
            d = gpu.lapack.svx(fact, tr, A, AF, piv, "N", r, c, B, X)
            if d["info"] != do["info"] or not np.isfinite(X.numpy()).all():
                # say which stage went wrong: the factor or the solves
                A3 = dev(gpu, a0)
                gpu.lapack.trf(A3)
                Bt = dev(gpu, b0)
                gpu.lapack.trs(tr, AF, piv, Bt)
                print("svx mismatch:", n, fact, tr, d, "AF == trf(A):",
                      np.array_equal(AF.numpy(), A3.numpy()), "AF NaN:", int(np.isnan(AF.numpy()).sum()),
                      "trs(AF) NaN:", np.isnan(Bt.numpy()).sum(0).tolist())
            assert d["info"] == do["info"] and d["equed"] == do["equed"], (fact, tr, d, do)
            assert same(r.cpu().numpy()[:n], ro) and same(c.cpu().numpy()[:n], co)
            assert same(A.numpy(), ao) and same(B.numpy(), bo)   # equilibration bit-exact
            assert abs(d["rcond"] - do["rcond"]) <= 1e-6 * do["rcond"], (d["rcond"], do["rcond"])
            assert abs(d["rpvgrw"] - do["rpvgrw"]) <= 1e-12 * do["rpvgrw"]
            x = X.numpy()
            fwd = np.abs(x - xt).max(axis=0) / np.abs(x).max(axis=0)
            assert np.all(fwd <= d["ferr"]), (fwd, d["ferr"])
            assert np.all(d["berr"] <= 4 * max(n, 1) * EPS + 1e-300)
            # the dgesvx_ seam on host arrays
            ah, bh = a0.copy(order="F"), b0.copy(order="F")
            afh, pivh = np.zeros((n, n), order="F"), np.zeros(max(1, n), dtype=np.int32)
            rh, ch, xh = np.zeros(max(1, n)), np.zeros(max(1, n)), np.zeros((n, 2), order="F")
            dh = gpu.lapack.svx(fact, tr, ah, afh, pivh, "N", rh, ch, bh, xh)
            assert dh["info"] == do["info"] and dh["equed"] == do["equed"]
            assert np.array_equal(xh, x) and dh["rcond"] == d["rcond"]


def test_svx_factored_and_singular(gpu):
    import torch
    n = 50
    a0 = rnd(n, n, 7, scale_rows=True)
    b0 = rnd(n, 1, 8)
    A, B = dev(gpu, a0), dev(gpu, b0)
    AF, X = gpu.DeviceMatrix.empty(n, n), gpu.DeviceMatrix.empty(n, 1)
    piv = torch.zeros(n, dtype=torch.int32, device="cuda")
    r = torch.zeros(n, dtype=torch.float64, device="cuda")
    c = torch.zeros(n, dtype=torch.float64, device="cuda")
    d = gpu.lapack.svx("E", 0, A, AF, piv, "N", r, c, B, X)
    x1 = X.numpy()
    # fact 'F' reuses AF, ipiv and the scalings found above: A as equilibrated
    # by the first call, B the original right-hand side (scaled inside again)
    X2 = gpu.DeviceMatrix.empty(n, 1)
    d2 = gpu.lapack.svx("F", 0, A, AF, piv, d["equed"], r, c, dev(gpu, b0), X2)
    assert d2["info"] == d["info"] == 0 and d2["equed"] == d["equed"]
    assert np.allclose(X2.numpy(), x1, rtol=1e-12, atol=0)
    s = rnd(8, 8, 90)
    s[:, 5] = 0.0
    As, Bs = dev(gpu, s), dev(gpu, rnd(8, 1, 91))
    AFs, Xs = gpu.DeviceMatrix.empty(8, 8), gpu.DeviceMatrix.empty(8, 1)
    ps = torch.zeros(8, dtype=torch.int32, device="cuda")
    rs = torch.zeros(8, dtype=torch.float64, device="cuda")
    cs = torch.zeros(8, dtype=torch.float64, device="cuda")
    ds = gpu.lapack.svx("N", 0, As, AFs, ps, "N", rs, cs, Bs, Xs)
    assert ds["info"] == 6 and ds["rcond"] == 0.0


# ----------------------------------------------------------------- getri ---
@pytest.mark.parametrize("n", [1, 7, 64, 65, 200, 1000, 2500])
def test_getri_matches_oracle(gpu, n):
    import torch
    a = rnd(n, n, 120 + n)
    lu = a.copy(order="F")
    piv, _ = O.dgetrf(lu)
    xo = lu.copy(order="F")
    assert O.dgetri(xo, piv) == 0
    A = dev(gpu, lu)
    pd = torch.from_numpy(piv.astype(np.int32)).cuda()
    assert gpu.lapack.tri(A, pd) == 0
    x = A.numpy()
    # dget03 ratio of the device inverse, and closeness to the oracle's
    r = np.linalg.norm(np.eye(n) - a @ x, 1)
    ratio = r / (n * np.linalg.norm(a, 1) * np.linalg.norm(x, 1) * EPS)
    assert ratio < 30, (n, ratio)
    kappa = np.linalg.norm(a, 1) * np.linalg.norm(xo, 1)
    assert np.abs(x - xo).max() <= 8 * n * EPS * kappa * np.abs(xo).max(), n
    # the dgetri_ seam
    xh = lu.copy(order="F")
    assert gpu.lapack.tri(xh, piv) == 0
    assert np.array_equal(xh, x)


def test_getri_closed_forms_and_singular(gpu):
    import torch
    rng = np.random.default_rng(110)
    n = 300
    p = np.asfortranarray(np.eye(n)[rng.permutation(n)])
    lu = p.copy(order="F")
    piv, _ = O.dgetrf(lu)
    A = dev(gpu, lu)
    assert gpu.lapack.tri(A, torch.from_numpy(piv.astype(np.int32)).cuda()) == 0
    assert np.array_equal(A.numpy(), p.T)                  # inv(P) = P^T exactly
    s = rnd(70, 70, 111)
    s[:, 40] = 0.0
    lu = s.copy(order="F")
    piv, info = O.dgetrf(lu)
    A = dev(gpu, lu)
    assert gpu.lapack.tri(A, torch.from_numpy(piv.astype(np.int32)).cuda()) == info == 41
    assert np.array_equal(A.numpy(), lu)                   # untouched


def test_empty_and_zero_rhs_cases(gpu):
    """Quick returns as in the reference: n = 0 and nrhs = 0."""
    import torch
    e = gpu.DeviceMatrix.empty(0, 0)
    assert gpu.lapack.lan("O", e) == 0.0
    r, c, rc, cc, am, info = gpu.lapack.equ(e)
    assert (rc, cc, am, info) == (1.0, 1.0, 0.0, 0)
    assert gpu.lapack.con("O", e, 1.0) == 1.0
    assert gpu.lapack.tri(e, torch.zeros(1, dtype=torch.int32, device="cuda")) == 0
    n = 5
    a = rnd(n, n, 3)
    lu = a.copy(order="F")
    piv, _ = O.dgetrf(lu)
    pd = torch.from_numpy(piv.astype(np.int32)).cuda()
    B0 = gpu.DeviceMatrix.empty(n, 0)
    X0 = gpu.DeviceMatrix.empty(n, 0)
    f, b = gpu.lapack.rfs(0, dev(gpu, a), dev(gpu, lu), pd, B0, X0)
    assert len(f) == 0 and len(b) == 0
    assert gpu.lapack.con("O", dev(gpu, lu), 0.0) == 0.0      # anorm = 0 -> 0
