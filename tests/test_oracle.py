"""CPU tests: pin the oracle (the C restatement of the FLENS generic path).

Each check is chosen so a plausible restatement mistake (dropped term, wrong
sign/index, transposed operand, wrong loop order) fails one of them:
  * bit-exact SHA-256 agreement with fixtures produced by the compiled
    reference (tests/golden/small.json, lu8192_nb*.json),
  * the 4x4 known answer of flens/examples/lapack-getrf.cc:43-46,
  * exact-zero / rank-deficient / tie cases (INFO and pivot conventions),
  * closed forms: permutation matrices (pivots = inverse permutation),
    diagonal matrices (identity pivots), integer GEMM (exact),
  * brute force against a textbook partial-pivot LU in long double (n<=8),
  * the checkers themselves reject a planted error.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def small():
    with open(os.path.join(GOLD, "small.json")) as f:
        return json.load(f)


def test_rng_numpy_matches_c():
    import ctypes as C
    a = O.uniform(37, 11, 1234, offset=5)
    b = np.zeros((37, 11), order="F")
    O.lib().orc_fill_uniform(b.ctypes.data_as(C.POINTER(C.c_double)), 37, 11, 37, 1234, 5)
    assert np.array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0
    # values are on the 2^-52 grid
    assert np.array_equal(a * 2.0 ** 52, np.round(a * 2.0 ** 52))


def test_gemm_grid_bit_exact(small):
    for c in small["gemm"]:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        s = c["seed"]
        a = O.uniform(k if ta else m, m if ta else k, s)
        b = O.uniform(n if tb else k, k if tb else n, s + 1)
        cc = O.uniform(m, n, s + 2)
        O.dgemm(ta, tb, c["alpha"], a, b, c["beta"], cc)
        assert sha(cc) == c["sha"], c


def test_lu_grid_bit_exact(small):
    for c in small["lu"]:
        a = O.uniform(c["m"], c["n"], c["seed"])
        piv, info = O.dgetrf(a, nb=c["nb"])
        assert info == c["info"] and piv.tolist() == c["piv"], c
        assert sha(a) == c["sha_lu"], c
        if c["m"] == c["n"]:
            for nrhs in (1, 2, 15, 16):
                for trans in (0, 1):
                    b = O.uniform(c["n"], nrhs, c["seed"] + 1)
                    O.dgetrs(trans, a, piv, b)
                    assert sha(b) == c[f"sha_x_{nrhs}_{trans}"], (c["m"], nrhs, trans)


def test_known_answer_4x4(small):
    # flens/examples/lapack-getrf.cc:43-46: trf on [A|b], then U solve
    ab = np.asfortranarray(np.array([[2, 3, -1, 0, 20], [-6, -5, 0, 2, -33],
                                     [2, -5, 6, -6, -43], [4, 6, 2, -3, 49]], float))
    piv, info = O.dgetrf(ab)
    assert piv.tolist() == [2, 3, 4, 4] and info == 0
    assert piv.tolist() == small["special"]["kat_getrf"]["piv"]
    assert ab.T.ravel().tolist() == small["special"]["kat_getrf"]["lu"]
    u = np.triu(ab[:, :4])
    x = np.linalg.solve(u, ab[:, 4])
    assert np.allclose(x, [1, 9, 9, 9])


def test_exact_zero_and_tie_cases(small):
    sp = small["special"]
    for key in ("zero_col", "rank_def", "zero_row", "tie"):
        a = np.array(np.array(sp[key]["a"], float).reshape(-1, int(len(sp[key]["a"]) ** 0.5)).T, order="F")
        piv, info = O.dgetrf(a)
        assert piv.tolist() == sp[key]["piv"] and info == sp[key]["info"], key
    # tie: |2| vs |-2| in column 1 -> the first (row 2 here since a11=1) wins
    assert sp["tie"]["piv"][0] == 2
    # the generic GEMM computes 0*C for beta = 0 (NaN propagates)
    assert sp["beta0_nan"]["all_nan"]
    a, b = O.uniform(3, 3, 1), O.uniform(3, 3, 2)
    c = np.asfortranarray(np.full((3, 3), np.nan))
    O.dgemm(0, 0, 1.0, a, b, 0.0, c)
    assert np.isnan(c).all()


def test_permutation_and_diagonal_closed_forms():
    rng = np.random.default_rng(3)
    for n in (1, 5, 64, 130):
        p = rng.permutation(n)
        a = np.asfortranarray(np.eye(n)[p])           # row i of A = e_{p[i]}
        piv, info = O.dgetrf(a, nb=32)
        # apply the interchanges to the identity ordering -> P A = I
        perm = np.arange(n)
        for i, q in enumerate(piv - 1):
            perm[[i, q]] = perm[[q, i]]
        assert info == 0
        assert np.array_equal(np.eye(n)[p][perm], np.eye(n))
        assert np.array_equal(a, np.eye(n))           # L = U = I
        d = np.asfortranarray(np.diag(rng.uniform(1, 2, n)))
        piv, info = O.dgetrf(d.copy(order="F"), nb=32)
        assert np.array_equal(piv, np.arange(1, n + 1))


def test_integer_gemm_exact():
    rng = np.random.default_rng(1)
    for (ta, tb) in [(0, 0), (0, 1), (1, 0), (1, 1)]:
        m, n, k = 17, 23, 31
        a = np.asfortranarray(rng.integers(-8, 9, (k, m) if ta else (m, k)).astype(float))
        b = np.asfortranarray(rng.integers(-8, 9, (n, k) if tb else (k, n)).astype(float))
        c = np.asfortranarray(rng.integers(-8, 9, (m, n)).astype(float))
        want = 2.0 * (a.T if ta else a) @ (b.T if tb else b) - 3.0 * c
        O.dgemm(ta, tb, 2.0, a, b, -3.0, c)
        assert np.array_equal(c, want)


def _textbook_lu(a):
    """Partial-pivot LU (Golub & Van Loan alg. 3.4.1) in long double."""
    a = a.astype(np.longdouble).copy()
    n = a.shape[0]
    piv = np.zeros(n, int)
    for j in range(n):
        p = j + int(np.argmax(np.abs(a[j:, j])))  # argmax keeps the first max
        piv[j] = p + 1
        a[[j, p]] = a[[p, j]]
        if a[j, j] != 0:
            a[j + 1:, j] /= a[j, j]
        a[j + 1:, j + 1:] -= np.outer(a[j + 1:, j], a[j, j + 1:])
    return a, piv


def test_brute_force_small():
    for n in range(1, 9):
        for seed in range(20):
            a0 = O.uniform(n, n, 900 + 31 * n + seed)
            lu_ref, piv_ref = _textbook_lu(a0)
            a = a0.copy(order="F")
            piv, info = O.dgetrf(a, nb=3)
            assert piv.tolist() == piv_ref.tolist()
            assert np.allclose(a, lu_ref.astype(float), rtol=1e-12, atol=1e-12)
            assert O.dget01(a0, a, piv) < 30


def test_checkers_reject_planted_error():
    a0 = O.uniform(40, 40, 77)
    a = a0.copy(order="F")
    piv, _ = O.dgetrf(a)
    assert O.dget01(a0, a, piv) < 30
    bad = a.copy(order="F")
    bad[17, 5] += 1e-9                                 # a dropped-term sized error
    assert O.dget01(a0, bad, piv) > 30
    wrong_piv = piv.copy()
    wrong_piv[3] = 4 if wrong_piv[3] != 4 else 5
    assert O.dget01(a0, a, wrong_piv) > 30
    b0 = O.uniform(40, 3, 78)
    x = b0.copy(order="F")
    O.dgetrs(0, a, piv, x)
    assert O.dget02(a0, x, b0) < 30
    x[0, 0] += 1e-9
    assert O.dget02(a0, x, b0) > 30
    # DMMCH
    m = 20
    aa, bb, c0 = O.uniform(m, m, 1), O.uniform(m, m, 2), O.uniform(m, m, 3)
    c = c0.copy(order="F")
    O.dgemm(0, 0, 0.7, aa, bb, 1.3, c)
    assert O.dmmch(0, 0, 0.7, aa, bb, 1.3, c0, c) < 16
    c[3, 4] += 1e-12
    assert O.dmmch(0, 0, 0.7, aa, bb, 1.3, c0, c) > 16


def test_laswp_forward_backward_inverse():
    a0 = O.uniform(30, 7, 5)
    piv = np.array([3, 5, 3, 30, 5, 6, 7, 29], np.int32)
    a = a0.copy(order="F")
    O.dlaswp(a, 1, 8, piv, 1)
    O.dlaswp(a, 1, 8, piv, -1)
    assert np.array_equal(a, a0)


def test_dsgesv_semantics():
    n = 60
    rng = np.random.default_rng(0)
    a = np.asfortranarray(O.uniform(n, n, 3) + n * np.eye(n))
    b = O.uniform(n, 2, 4)
    x, piv, info, it = O.dsgesv(a.copy(order="F"), b)
    assert info == 0 and 0 <= it <= 30
    assert np.linalg.norm(a @ x - b, np.inf) <= n * 2.0 ** -53 * np.linalg.norm(a, np.inf) * np.linalg.norm(x, np.inf) * 10
    # no refinement allowed and a badly scaled system -> fallback code -1
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ill = np.asfortranarray(q @ np.diag(np.logspace(0, -12, n)) @ q.T)
    x, piv, info, it = O.dsgesv(ill.copy(order="F"), b, itermax=0)
    assert it == -1
    # overflow in the single-precision copy -> -2
    big = a.copy(order="F")
    big[0, 0] = 1e300
    x, piv, info, it = O.dsgesv(big, b)
    assert it == -2


@pytest.mark.skipif(not O.have_ref(), reason="reference sources not built here")
def test_oracle_vs_compiled_reference_random():
    rng = np.random.default_rng(11)
    for _ in range(6):
        m, n = (int(x) for x in rng.integers(1, 200, 2))
        nb = int(rng.choice([1, 2, 8, 32, 64]))
        a0 = O.uniform(m, n, int(rng.integers(1 << 30)))
        a1, a2 = a0.copy(order="F"), a0.copy(order="F")
        p1, i1 = O.dgetrf(a1, nb=nb)
        p2, i2 = O.dgetrf(a2, nb=nb, use_ref=True)
        assert i1 == i2 and np.array_equal(p1, p2) and sha(a1) == sha(a2)


def test_big_lu_fixture_consistency():
    """The n=8192 fixtures were produced by the reference; the oracle's own
    run of the same input was compared bit for bit when they were made."""
    found = 0
    for nb in (64, 128, 256):
        path = os.path.join(GOLD, f"lu8192_nb{nb}.json")
        if not os.path.exists(path):
            continue
        found += 1
        meta = json.load(open(path))
        z = np.load(os.path.join(GOLD, f"lu8192_nb{nb}.npz"))
        assert meta["oracle_bit_equal"]
        piv = z["piv"]
        assert piv.shape == (8192,) and piv.min() >= 1 and piv.max() <= 8192
        assert np.all(piv >= np.arange(1, 8193))
    assert found >= 1


def test_selector_network_oracle_against_finite_differences():
    """Pins the CPU restatement of the selector network: its analytic
    gradient equals central finite differences of its loss."""
    from oracle import mlp as M
    rng = np.random.default_rng(0)
    p = M.init(3).astype(np.float64)
    p[M.PARAMS - M.OUT:] = rng.standard_normal(M.OUT) * 0.1     # nonzero b3
    X = rng.standard_normal((9, M.IN))
    y = rng.integers(0, M.OUT, 9)
    w = rng.uniform(0.5, 2.0, 9)
    loss, g = M.loss_and_grad(p, X, y, w)
    idx = rng.choice(M.PARAMS, 60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(M.PARAMS)
        e[i] = h
        fd = (M.loss_and_grad(p + e, X, y, w)[0] - M.loss_and_grad(p - e, X, y, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # a planted mistake -- the backward pass without the second ReLU mask --
    # disagrees with the finite differences
    d = M.unpack(p)
    z, (X_, a1, h1, a2, h2) = M.forward(p, X)
    nb = X.shape[0]
    e = np.exp(z - z.max(1, keepdims=True))
    d3 = e / e.sum(1, keepdims=True)
    d3[np.arange(nb), y] -= 1.0
    d3 *= w[:, None] / nb
    bad_d2 = d3 @ d["W3"]                       # mask dropped
    bad_gb2 = bad_d2.sum(0)
    o = M.IN * M.H + M.H + M.H * M.H             # offset of b2
    assert np.abs(bad_gb2 - g[o:o + M.H]).max() > 1e-3 * np.abs(g[o:o + M.H]).max()


# ---------------------------------------------------------------- dsgesv ---
def _dsgesv_ratios(a, b, itmax=30):
    """The refinement's exit ratios max_j ||r_j||_inf / (||x_j||_inf cte),
    cte = ||A||_inf 2^-53 sqrt(n), for the iterates x_0, x_1, ... built from
    the oracle's FP32 LU (pinned below on exact cases) with FP64 residuals
    formed by numpy (another summation order than the oracle's DGEMM)."""
    n = a.shape[0]
    sa = np.asfortranarray(a.astype(np.float32))
    piv, info = O.sgetrf(sa)
    assert info == 0
    sx = np.asfortranarray(b.astype(np.float32))
    O.sgetrs(0, sa, piv, sx)
    x = sx.astype(np.float64)
    cte = np.abs(a).sum(1).max() * 2.0 ** -53 * np.sqrt(n)
    out = []
    for _ in range(itmax + 1):
        r = b - a @ x
        out.append(max(np.abs(r[:, j]).max() / (np.abs(x[:, j]).max() * cte)
                       for j in range(b.shape[1])))
        if out[-1] <= 1.0:
            break
        sr = np.asfortranarray(r.astype(np.float32))
        O.sgetrs(0, sa, piv, sr)
        x = x + sr.astype(np.float64)
    return out


def test_dsgesv_exit_rule_exact():
    """orc_dsgesv stops at the first iterate with ||r||_inf <= ||x||_inf
    ||A||_inf 2^-53 sqrt(n) (netlib dsgesv.f) and not one iterate earlier or
    later.  The two cases are chosen so that a threshold off by a factor of
    1.6 either way moves ITER: in the first the last rejected iterate misses
    by 1.56x, in the second the accepted one passes with 0.71x."""
    from flens_amd import testmat as T
    for seed, kappa, want in ((0, 1e6, 4), (5, 1e7, 5)):
        a = T.svd_matrix(48, kappa, seed)
        b, _ = T.rhs_from_solution(a, 1, seed)
        ratios = _dsgesv_ratios(a, b)
        assert len(ratios) == want + 1 and ratios[-1] <= 1.0
        assert ratios[-2] > 1.4 and (seed != 0 or ratios[-2] < 1.7)
        assert seed != 5 or 0.6 < ratios[-1] < 0.8
        x, piv, info, it = O.dsgesv(a.copy(order="F"), b)
        assert info == 0 and it == want, (seed, it, ratios)


def test_dsgesv_fallback_codes_on_the_oracle():
    """ITER -3 (an exactly singular FP32 factor; the FP64 fallback's INFO is
    the zero column) and ITER -31 (no convergence within 30 steps at
    kappa = 1e12; the FP64 fallback's X solves the system and A holds its
    factors), checked on the oracle alone."""
    from flens_amd import testmat as T
    n = 50
    a = np.asfortranarray(O.uniform(n, n, 5))
    a[:, 17] = 0.0
    b = O.uniform(n, 1, 2)
    x, piv, info, it = O.dsgesv(a.copy(order="F"), b)
    assert it == -3 and info == 18
    a = T.svd_matrix(n, 1e12, 2)
    b, _ = T.rhs_from_solution(a, 1, 3)
    ac = a.copy(order="F")
    x, piv, info, it = O.dsgesv(ac, b)
    assert it == -31 and info == 0
    assert O.dget02(a, x, b) < 30
    assert O.dget01(a, ac, piv) < 30
    lu = a.copy(order="F")
    p64, _ = O.dgetrf(lu)
    assert np.array_equal(piv, p64) and np.array_equal(ac, lu)


def test_sgetrf_exact_cases():
    """orc_sgetrf on matrices whose partial-pivot LU is exact in FP32:
    A = P^T L U with |l_ij| <= 1/2 dyadic, U integer with a dominant
    diagonal -- every product and sum is representable, so the factors are
    L and U exactly and the pivots are those of P (a closed form).  The FP64
    oracle gives the same bits."""
    rng = np.random.default_rng(3)
    for n in (1, 5, 33, 70, 130):
        l = np.tril(rng.integers(-2, 3, (n, n)) / 4.0, -1) + np.eye(n)
        u = np.triu(rng.integers(-3, 4, (n, n)).astype(float), 1) + np.diag(rng.integers(8, 16, n) * 1.0)
        a = l @ u
        # a row permutation applied as a sequence of interchanges with
        # ipiv[i] >= i that partial pivoting must reproduce: the rows of
        # L U are moved so each column's pivot sits below the diagonal
        perm_piv = np.array([rng.integers(i, n) for i in range(n)], np.int32) + 1
        pa = a.copy()
        for i in range(n - 1, -1, -1):
            q = perm_piv[i] - 1
            pa[[i, q]] = pa[[q, i]]
        s = np.asfortranarray(pa.astype(np.float32))
        piv, info = O.sgetrf(s)
        assert info == 0
        assert np.array_equal(piv, perm_piv), n
        assert np.array_equal(np.tril(s, -1) + np.eye(n), l.astype(np.float32)), n
        assert np.array_equal(np.triu(s), u.astype(np.float32)), n
        d = np.asfortranarray(pa.copy())
        piv64, _ = O.dgetrf(d)
        assert np.array_equal(piv64, perm_piv) and np.array_equal(d.astype(np.float32), s)


# ------------------------------------------------------------------ trsm ---
@pytest.mark.skipif(not O.have_ref(), reason="reference sources not built here")
def test_trsm_bit_exact_vs_compiled_reference():
    """orc_dtrsm (trsm_generic's order: per-row/column trsv with sequential
    dot products, alpha applied after the solve) equals the reference's
    cxxblas::trsm bit for bit, every side/uplo/trans/diag, both orders."""
    import itertools
    rng = np.random.default_rng(1)
    for order in "FC":
        for right, upper, trans, unit in itertools.product((0, 1), repeat=4):
            for (m, n) in ((5, 7), (13, 1), (1, 9), (33, 17)):
                na = n if right else m
                a = rng.uniform(-1, 1, (na, na)) + np.diag(2 + rng.uniform(0, 1, na))
                b = rng.uniform(-1, 1, (m, n))
                a, b = np.array(a, order=order), np.array(b, order=order)
                for alpha in (1.0, 0.7, -1.3):
                    b1, b2 = b.copy(order=order), b.copy(order=order)
                    O.dtrsm(order, right, upper, trans, unit, alpha, a, b1)
                    O.dtrsm(order, right, upper, trans, unit, alpha, a, b2, use_ref=True)
                    assert np.array_equal(b1, b2), (order, right, upper, trans, unit, m, n, alpha)


def test_trsm_closed_forms():
    """Closed forms that fix side, triangle, transpose and diag: against a
    solve in long-double-free exact arithmetic (integer triangles with unit
    or power-of-two diagonals solve exactly), the strict other triangle is
    never read (filled with NaN), and alpha scales the solution."""
    import itertools
    rng = np.random.default_rng(2)
    for order in "FC":
        for right, upper, trans, unit in itertools.product((0, 1), repeat=4):
            m, n = 9, 6
            na = n if right else m
            t = rng.integers(-2, 3, (na, na)).astype(float)
            t = np.triu(t, 1) if upper else np.tril(t, -1)
            dg = np.ones(na) if unit else 2.0 ** rng.integers(-2, 3, na)
            full = t + np.diag(dg)
            x = rng.integers(-4, 5, (m, n)).astype(float)
            op = full.T if trans else full
            b = (op @ x) if not right else (x @ op)          # exact: small integers/dyadics
            a = full.copy()
            a[np.tril_indices(na, -1) if upper else np.triu_indices(na, 1)] = np.nan
            if unit:
                a[np.diag_indices(na)] = np.nan               # never read either
            a, bb = np.array(a, order=order), np.array(b, order=order)
            O.dtrsm(order, right, upper, trans, unit, -2.0, a, bb)
            assert np.array_equal(bb, -2.0 * x), (order, right, upper, trans, unit)
