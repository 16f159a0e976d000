"""CPU tests: pin the oracle's restatement of lan, equ, laq, con (lacn2,
latrs, rscl), rfs and svx (oracle/aux_body.h).

Pins, each chosen so a plausible mistake fails one of them:
  * bit-exact agreement with the compiled reference (oracle/_ref, built from
    the reference's own headers) over random, scaled, singular and
    NaN/Inf-carrying inputs, when that build is present;
  * closed forms: norms against Python loops in the reference's summation
    order (bit-exact) and numpy; equilibration of power-of-two matrices;
    the condition estimate of diagonal matrices (exact) and its lower-bound
    property on random ones; refinement of an integer system (berr = 0) and
    the forward-error bound property on random ones;
  * lacn2 / lassq / std::max quirks the reference has (two infinities give
    a NaN Frobenius norm, NaN entries never win a max).
"""
import numpy as np
import pytest

from oracle import oracle as O

HAVE_REF = O.have_ref()
needs_ref = pytest.mark.skipif(not HAVE_REF, reason="compiled reference (oracle/_ref) absent")
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


# ----------------------------------------------------------------- lange ---
def seq_sums(a):
    """Column and row sums of |a| in the reference's sequential order."""
    m, n = a.shape
    col = [0.0] * n
    for j in range(n):
        s = 0.0
        for i in range(m):
            s += abs(float(a[i, j]))
        col[j] = s
    row = [0.0] * m
    for j in range(n):
        for i in range(m):
            row[i] += abs(float(a[i, j]))
    return col, row


def test_lange_closed_forms():
    a = rnd(23, 17, 1)
    assert O.dlange("M", a) == np.abs(a).max()
    col, row = seq_sums(a)
    assert O.dlange("O", a) == max(col)
    assert O.dlange("I", a) == max(row)
    assert abs(O.dlange("F", a) - np.linalg.norm(a)) <= 4 * EPS * np.linalg.norm(a)
    z = np.zeros((5, 4), order="F")
    for nm in "MOIF":
        assert O.dlange(nm, z) == 0.0
        assert O.dlange(nm, np.zeros((0, 3), order="F")) == 0.0


def test_lange_nan_inf_semantics():
    a = rnd(9, 7, 2)
    b = a.copy(order="F")
    b[3, 2] = np.nan
    # std::max never picks a NaN: M skips it, O / I skip the NaN column / row sum
    assert O.dlange("M", b) == np.abs(np.where(np.isnan(b), 0, b)).max()
    col, row = seq_sums(b)
    assert O.dlange("O", b) == max(v for v in col if v == v)
    assert O.dlange("I", b) == max(v for v in row if v == v)
    assert np.isnan(O.dlange("F", b))
    c = a.copy(order="F")
    c[0, 0] = np.inf
    assert O.dlange("F", c) == np.inf          # one infinity: lassq gives inf
    c[5, 5] = -np.inf
    assert np.isnan(O.dlange("F", c))          # two: inf / inf inside lassq


@needs_ref
@pytest.mark.parametrize("shape", [(1, 1), (7, 5), (40, 40), (100, 60), (33, 129)])
def test_lange_vs_reference(shape):
    a = rnd(*shape, 3)
    for variant in range(3):
        b = a.copy(order="F")
        if variant == 1:
            b[0, -1] = np.nan
        if variant == 2:
            b[-1, 0] = np.inf
        for nm in "MOIF":
            assert same(O.dlange(nm, b), O.dlange(nm, b, use_ref=True)), (shape, variant, nm)


# ----------------------------------------------------------- equ / laq ---
def test_equ_power_of_two_closed_form():
    rng = np.random.default_rng(4)
    e = rng.integers(-20, 20, size=(12, 9))
    a = np.asfortranarray(np.ldexp(np.where(rng.random((12, 9)) < 0.5, -1.0, 1.0), e))
    r, c, rc, cc, am, info = O.dgeequ(a)
    assert info == 0
    rmax = np.abs(a).max(axis=1)
    assert np.array_equal(r, 1.0 / rmax)
    assert np.array_equal(c, 1.0 / (np.abs(a) * r[:, None]).max(axis=0))
    assert am == rmax.max() and rc == rmax.min() / rmax.max()


def test_equ_zero_row_and_column_info():
    a = rnd(6, 5, 5)
    a[3, :] = 0.0
    assert O.dgeequ(a)[5] == 4                   # first zero row, 1-based
    b = rnd(6, 5, 6)
    b[:, 2] = 0.0
    assert O.dgeequ(b)[5] == 6 + 3               # m + first zero column


def test_laqge_modes():
    a = rnd(8, 8, 7)
    r, c = np.full(8, 2.0), np.full(8, 4.0)
    for rowcnd, colcnd, amax, mode in [(1, 1, 1, "N"), (1, 0.01, 1, "C"), (0.01, 1, 1, "R"),
                                       (0.01, 0.01, 1, "B"), (1, 1, 1e-300, "R")]:
        b = a.copy(order="F")
        assert O.dlaqge(b, r, c, rowcnd, colcnd, amax) == mode
        f = {"N": 1, "C": 4, "R": 2, "B": 8}[mode]
        assert np.array_equal(b, a * f)


@needs_ref
def test_equ_laq_vs_reference():
    for seed, (m, n) in enumerate([(50, 50), (31, 77), (90, 12)]):
        a = rnd(m, n, 10 + seed, scale_rows=True)
        r1, r2 = O.dgeequ(a), O.dgeequ(a, use_ref=True)
        for u, v in zip(r1, r2):
            assert same(u, v) or (np.isnan(u) and np.isnan(v))
        b1, b2 = a.copy(order="F"), a.copy(order="F")
        assert O.dlaqge(b1, *r1[:5]) == O.dlaqge(b2, *r2[:5], use_ref=True)
        assert same(b1, b2)


# ------------------------------------------------------------------ con ---
def test_con_diagonal_exact():
    d = np.array([3.0, -0.5, 8.0, 0.25, -2.0])
    a = np.asfortranarray(np.diag(d))
    lu = a.copy(order="F")
    O.dgetrf(lu)
    for nm in "OI":
        rc = O.dgecon(nm, lu, O.dlange(nm, a))
        assert abs(rc - np.abs(d).min() / np.abs(d).max()) <= 4 * EPS


def test_con_lower_bound_and_quality():
    for n in (2, 9, 60, 200):
        a = rnd(n, n, 20 + n)
        lu = a.copy(order="F")
        O.dgetrf(lu)
        inv = np.linalg.inv(a)
        for nm, p in (("O", 1), ("I", np.inf)):
            rc = O.dgecon(nm, lu, O.dlange(nm, a))
            true = 1.0 / (np.linalg.norm(a, p) * np.linalg.norm(inv, p))
            # lacn2 estimates ||inv(A)|| from below, so rcond from above
            assert true * (1 - 1e-10) <= rc <= 10 * true, (n, nm, rc, true)


def test_con_singular_and_trivial():
    lu = np.asfortranarray(np.triu(rnd(6, 6, 31)))
    lu[4, 4] = 0.0                          # exact zero on U's diagonal
    assert O.dgecon("O", lu, 1.0) == 0.0
    assert O.dgecon("O", np.ones((1, 1), order="F"), 0.0) == 0.0
    assert O.dgecon("O", np.zeros((0, 0), order="F"), 1.0) == 1.0


@needs_ref
@pytest.mark.parametrize("n", [1, 2, 5, 37, 120, 300])
def test_con_vs_reference(n):
    a = rnd(n, n, 40 + n)
    lu = a.copy(order="F")
    O.dgetrf(lu)
    for nm in "OI":
        an = O.dlange(nm, a)
        assert same(O.dgecon(nm, lu, an), O.dgecon(nm, lu, an, use_ref=True))


# ------------------------------------------------------------------ rfs ---
def test_rfs_integer_system_exact():
    # A = P D U (U unit upper integer, D powers of two): partial pivoting
    # recovers L = I, U' = D U exactly, so the solve of an integer system is
    # exact and refinement must report berr = 0 without changing x
    rng = np.random.default_rng(50)
    n = 12
    u = np.triu(rng.integers(-3, 4, size=(n, n)).astype(float), 1) + np.eye(n)
    d = np.diag(np.ldexp(1.0, rng.integers(-3, 4, size=n)))
    p = np.eye(n)[rng.permutation(n)]
    a = np.asfortranarray(p @ d @ u)
    xt = np.asfortranarray(rng.integers(-9, 10, size=(n, 2)).astype(float))
    b = np.asfortranarray(a @ xt)
    lu = a.copy(order="F")
    piv, _ = O.dgetrf(lu)
    x = b.copy(order="F")
    O.dgetrs(0, lu, piv, x)
    ferr, berr = O.dgerfs(0, a, lu, piv, b, x)
    assert np.array_equal(x, xt) and np.all(berr == 0.0)
    assert np.all(ferr >= 0.0) and np.all(ferr < 1e-8)   # the (n+1) eps |inv(A)| (|A||x|+|b|) term


def test_rfs_bounds_hold():
    for n, tr in ((30, 0), (80, 1), (150, 0)):
        rng = np.random.default_rng(60 + n)
        a = rnd(n, n, 61 + n)
        xt = np.asfortranarray(rng.standard_normal((n, 3)))
        b = np.asfortranarray((a.T if tr else a) @ xt)
        lu = a.copy(order="F")
        piv, _ = O.dgetrf(lu)
        x = b.copy(order="F")
        O.dgetrs(tr, lu, piv, x)
        ferr, berr = O.dgerfs(tr, a, lu, piv, b, x)
        fwd = np.abs(x - xt).max(axis=0) / np.abs(x).max(axis=0)
        assert np.all(fwd <= ferr), (n, fwd, ferr)
        assert np.all(berr <= 4 * EPS * n)


@needs_ref
def test_rfs_vs_reference():
    for n in (1, 2, 5, 37, 120):
        a = rnd(n, n, 70 + n)
        b = rnd(n, 3, 71 + n)
        lu = a.copy(order="F")
        piv, _ = O.dgetrf(lu)
        for tr in (0, 1):
            x0 = b.copy(order="F")
            O.dgetrs(tr, lu, piv, x0)
            x1, x2 = x0.copy(order="F"), x0.copy(order="F")
            f1, b1 = O.dgerfs(tr, a, lu, piv, b, x1)
            f2, b2 = O.dgerfs(tr, a, lu, piv, b, x2, use_ref=True)
            assert same(x1, x2) and same(f1, f2) and same(b1, b2), (n, tr)


# ------------------------------------------------------------------ svx ---
def _svx(fact, tr, a0, b0, use_ref, equed="N", r=None, c=None, af=None, piv=None):
    n = a0.shape[0]
    a, b = a0.copy(order="F"), b0.copy(order="F")
    af = np.zeros((n, n), order="F") if af is None else af.copy(order="F")
    piv = np.zeros(n, dtype=np.int32) if piv is None else piv.copy()
    r = np.zeros(n) if r is None else r.copy()
    c = np.zeros(n) if c is None else c.copy()
    d = O.dgesvx(fact, tr, a, af, piv, equed, r, c, b, use_ref=use_ref)
    return a, af, piv, r, c, b, d


def test_svx_equilibrates_and_solves():
    n = 60
    a0 = rnd(n, n, 80, scale_rows=True)
    rng = np.random.default_rng(81)
    xt = rng.standard_normal((n, 2))
    b0 = np.asfortranarray(a0 @ xt)
    a, af, piv, r, c, b, d = _svx("E", 0, a0, b0, False)
    assert d["info"] == 0 and d["equed"] in ("R", "B")
    assert np.array_equal(b, b0 * r[:, None])   # B scaled by diag(r)
    fwd = np.abs(d["x"] - xt).max(axis=0) / np.abs(d["x"]).max(axis=0)
    assert np.all(fwd <= d["ferr"])
    # rcond equals con on the equilibrated matrix
    lu = a.copy(order="F")
    O.dgetrf(lu)
    assert d["rcond"] == O.dgecon("O", lu, O.dlange("O", a))


def test_svx_singular_info():
    a0 = rnd(8, 8, 90)
    a0[:, 5] = 0.0
    _, _, _, _, _, _, d = _svx("N", 0, a0, rnd(8, 1, 91), False)
    assert d["info"] == 6 and d["rcond"] == 0.0


@needs_ref
@pytest.mark.parametrize("n", [1, 4, 50, 200])
def test_svx_vs_reference(n):
    a0 = rnd(n, n, 100 + n, scale_rows=True)
    b0 = rnd(n, 2, 101 + n)
    for fact in "NE":
        for tr in (0, 1):
            o = _svx(fact, tr, a0, b0, False)
            f = _svx(fact, tr, a0, b0, True)
            for u, v in zip(o[:6], f[:6]):
                assert same(u, v), (n, fact, tr)
            for k in ("x", "ferr", "berr", "rcond", "rpvgrw", "info"):
                assert same(o[6][k], f[6][k]), (n, fact, tr, k)
    # fact 'F' with the equilibration found above (the reference takes EQUED in)
    a, af, piv, r, c, b, d = _svx("E", 0, a0, b0, False)
    o = _svx("F", 0, a, b, False, equed=d["equed"], r=r, c=c, af=af, piv=piv)
    f = _svx("F", 0, a, b, True, equed=d["equed"], r=r, c=c, af=af, piv=piv)
    for k in ("x", "ferr", "berr", "rcond", "rpvgrw", "info"):
        assert same(o[6][k], f[6][k]), (n, "F", k)


# ----------------------------------------------------------------- getri ---
def test_getri_closed_forms():
    rng = np.random.default_rng(110)
    n = 9
    p = np.asfortranarray(np.eye(n)[rng.permutation(n)])
    lu = p.copy(order="F")
    piv, _ = O.dgetrf(lu)
    assert O.dgetri(lu, piv) == 0 and np.array_equal(lu, p.T)   # inv(P) = P^T exactly
    d = np.ldexp(1.0, rng.integers(-5, 6, size=n)) * np.where(rng.random(n) < 0.5, -1, 1)
    lu = np.asfortranarray(np.diag(d))
    piv, _ = O.dgetrf(lu)
    assert O.dgetri(lu, piv) == 0 and np.array_equal(lu, np.diag(1.0 / d))
    s = rnd(6, 6, 111)
    s[:, 2] = 0.0
    lu = s.copy(order="F")
    piv, info = O.dgetrf(lu)
    before = lu.copy(order="F")
    assert O.dgetri(lu, piv) == info == 3 and np.array_equal(lu, before)


def test_getri_inverse_residual():
    for n in (1, 7, 64, 130):
        a = rnd(n, n, 120 + n)
        lu = a.copy(order="F")
        piv, _ = O.dgetrf(lu)
        assert O.dgetri(lu, piv) == 0
        # LAPACK dget03: ||I - A inv(A)||_1 / (n ||A||_1 ||inv(A)||_1 eps) < 30
        r = np.linalg.norm(np.eye(n) - a @ lu, 1)
        ratio = r / (n * np.linalg.norm(a, 1) * np.linalg.norm(lu, 1) * EPS)
        assert ratio < 30, (n, ratio)


@needs_ref
def test_getri_vs_reference():
    # bit-exact on the reference's unblocked path (n <= 64), close beyond
    for n in (1, 2, 5, 40, 64, 65, 150):
        a = rnd(n, n, 130 + n)
        lu = a.copy(order="F")
        piv, _ = O.dgetrf(lu)
        x1, x2 = lu.copy(order="F"), lu.copy(order="F")
        assert O.dgetri(x1, piv) == O.dgetri(x2, piv, use_ref=True) == 0
        if n <= 64:
            assert same(x1, x2), n
        else:
            assert np.abs(x1 - x2).max() <= 64 * n * EPS * np.abs(x2).max(), n


def test_empty_quick_returns():
    e = np.zeros((0, 0), order="F")
    assert O.dlange("F", e) == 0.0
    r, c, rc, cc, am, info = O.dgeequ(e)
    assert (rc, cc, am, info) == (1.0, 1.0, 0.0, 0)
    assert O.dgecon("I", e, 1.0) == 1.0
    assert O.dgetri(e, np.zeros(1, dtype=np.int32)) == 0
    a = rnd(4, 4, 5)
    lu = a.copy(order="F")
    piv, _ = O.dgetrf(lu)
    f, b = O.dgerfs(0, a, lu, piv, np.zeros((4, 0), order="F"), np.zeros((4, 0), order="F"))
    assert len(f) == 0 and len(b) == 0
