"""GPU parity for the CBLAS triangular-solve seam cblas_dtrsm against the
oracle's restatement of cxxblas trsm_generic (cxxblas/level3/trsm.tcc:40-68,
reached from blas::sm, flens/blas/level3/sm.tcc:63): every side, uplo,
transpose and diag, both storage orders, several alphas.

Bar: the forward-error bound of a triangular solve,
|X_gpu - X_orc| <= 4 (n+1) eps64 |alpha| |op(A)^-1| |op(A)| |X/alpha| + 2 eps64 |X|
elementwise (both results lie within gamma_n |T^-1||T||x| of the exact
solution; the last term covers the scaling by alpha after the solve), on
well-conditioned triangles (diagonal in [2,3), off-diagonal U[-1,1]/n).
"""
import ctypes as C
import itertools

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
DP = C.POINTER(C.c_double)


def tri(na, upper, rng):
    a = rng.uniform(-1.0, 1.0, (na, na)) / max(na, 1)
    a[np.diag_indices(na)] = 2.0 + rng.uniform(0.0, 1.0, na)
    # the other triangle holds garbage the solve must not read
    junk = np.tril(np.full((na, na), 1e300), -1) if upper else np.triu(np.full((na, na), 1e300), 1)
    return (np.triu(a) if upper else np.tril(a)) + junk


def bound(a, upper, trans, unit, right, alpha, x):
    na = a.shape[0]
    t = np.triu(a) if upper else np.tril(a)
    if unit:
        t = t - np.diag(np.diag(t)) + np.eye(na)
    op = t.T if trans else t
    inv = np.abs(np.linalg.inv(op))
    y = np.abs(x / alpha) if alpha != 0 else np.zeros_like(x)
    g = (inv @ (np.abs(op) @ y)) if not right else ((y @ np.abs(op)) @ inv)
    return 4 * (na + 1) * EPS * abs(alpha) * g + 2 * EPS * np.abs(x) + 1e-300


SIZES = [(1, 1), (5, 7), (33, 17), (64, 65), (130, 3)]


@pytest.mark.parametrize("order", [101, 102])
def test_cblas_dtrsm_all_combinations(gpu, order):
    lib = gpu.lib()
    rng = np.random.default_rng(order)
    npo = "C" if order == 101 else "F"
    worst = 0.0
    for right, upper, trans, unit in itertools.product((0, 1), repeat=4):
        for (m, n) in SIZES:
            na = n if right else m
            a = np.array(tri(na, upper, rng), order=npo)
            b = np.array(rng.uniform(-1.0, 1.0, (m, n)), order=npo)
            for alpha in (1.0, 0.7, -1.3):
                want = O.dtrsm(npo, right, upper, trans, unit, alpha, a, b.copy(order=npo))
                got = b.copy(order=npo)
                lda = na
                ldb = n if order == 101 else m
                lib.cblas_dtrsm(order, 142 if right else 141, 121 if upper else 122,
                                112 if trans else 111, 132 if unit else 131, m, n, alpha,
                                a.ctypes.data_as(DP), lda, got.ctypes.data_as(DP), ldb)
                bd = bound(a, upper, trans, unit, right, alpha, want)
                err = np.abs(got - want)
                case = (order, right, upper, trans, unit, m, n, alpha)
                assert np.all(np.isfinite(got)), case
                assert np.all(err <= bd), (case, float((err / bd).max()))
                worst = max(worst, float((err / bd).max()))
    print(f"order {order}: worst error / bound = {worst:.3g}")


def test_cblas_dtrsm_alpha_zero_and_exact(gpu):
    """alpha = 0 gives zeros after the solve, as the reference's gescal does;
    power-of-two diagonal triangles with integer data solve exactly, so the
    seam equals the oracle bit for bit."""
    lib = gpu.lib()
    m, n = 40, 6
    a = np.asfortranarray(np.diag(np.exp2(np.arange(m) % 5 - 2.0)))
    b = np.asfortranarray(np.arange(m * n, dtype=np.float64).reshape(m, n) - 50.0)
    for alpha in (0.0, 1.0, -2.0):
        got = b.copy(order="F")
        lib.cblas_dtrsm(102, 141, 121, 111, 131, m, n, alpha, a.ctypes.data_as(DP), m,
                        got.ctypes.data_as(DP), m)
        want = O.dtrsm("F", 0, 1, 0, 0, alpha, a, b.copy(order="F"))
        assert np.array_equal(got, want), alpha
