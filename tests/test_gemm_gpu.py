"""GPU parity for the FP64 MFMA GEMM (flx_dgemm / cblas_dgemm) against the
oracle's restatement of cxxblas gemm_generic (cxxblas/level3/gemm.tcc:149-299).

Bar: |C_gpu - C_ref| <= 4*K*eps64*(|alpha||A||B| +
|beta||C|) elementwise, eps64 = 2^-53; netlib DMMCH ratio < 16 over the
dblat3 grid; bit-exact where the arithmetic is exact (integer data,
alpha = 0); the reference's beta = 0 convention on the flx path.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
EPS = 2.0 ** -53
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bound(ta, tb, alpha, a, b, beta, c0):
    opa = a.T if ta else a
    opb = b.T if tb else b
    k = opa.shape[1]
    g = abs(alpha) * (np.abs(opa) @ np.abs(opb)) + abs(beta) * np.abs(c0)
    return 4 * max(k, 1) * EPS * g


def run_dev(F, ta, tb, alpha, a, b, beta, c):
    A = F.DeviceMatrix.from_numpy(a)
    B = F.DeviceMatrix.from_numpy(b)
    Cd = F.DeviceMatrix.from_numpy(c)
    F.blas.mm(ta, tb, alpha, A, B, beta, Cd)
    return Cd.numpy()


SHAPES = [(1, 1, 1), (5, 7, 3), (128, 128, 16), (129, 127, 17), (300, 200, 100),
          (64, 513, 40), (1000, 33, 257), (17, 400, 0), (256, 256, 256), (32, 2000, 128)]


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_vs_oracle_bound(gpu, ta, tb):
    for (m, n, k) in SHAPES:
        for (alpha, beta) in [(1.5, -0.5), (-1.0, 1.0), (0.7, 0.0)]:
            a = O.uniform(k if ta else m, m if ta else k, m * 7 + n)
            b = O.uniform(n if tb else k, k if tb else n, n * 5 + k)
            c0 = O.uniform(m, n, k + 3)
            want = O.dgemm(ta, tb, alpha, a, b, beta, c0.copy(order="F"))
            got = run_dev(gpu, ta, tb, alpha, a, b, beta, c0.copy(order="F"))
            err = np.abs(got - want)
            assert np.all(err <= bound(ta, tb, alpha, a, b, beta, c0) + 1e-300), (m, n, k, alpha, beta)


def test_narrow_gemm_matches_tiled_bitwise(gpu):
    """The one-wave-per-block kernel for narrow C (M <= 128, K <= 128, the LU's
    next-block update) runs the tiled kernel's MFMA sequence: the first 128
    rows of an M = 129 product (tiled) equal the M = 128 product (narrow)
    bit for bit, K a multiple of 16."""
    for (n, k) in [(300, 128), (1000, 64), (37, 16)]:
        a = O.uniform(129, k, 11 + k)
        b = O.uniform(k, n, 12 + n)
        c = O.uniform(129, n, 13 + n + k)
        tiled = run_dev(gpu, 0, 0, -1.0, a, b, 1.0, c.copy(order="F"))
        narrow = run_dev(gpu, 0, 0, -1.0, np.asfortranarray(a[:128]), b, 1.0,
                         np.asfortranarray(c[:128]))
        assert np.array_equal(tiled[:128].view(np.uint64), narrow.view(np.uint64)), (n, k)


def test_gemm_dblat3_grid_dmmch(gpu):
    with open(os.path.join(GOLD, "small.json")) as f:
        cases = json.load(f)["gemm"]
    worst = 0.0
    for c in cases:
        ta, tb, m, n, k = c["ta"], c["tb"], c["m"], c["n"], c["k"]
        if m == 0 or n == 0:
            continue
        s = c["seed"]
        a = O.uniform(k if ta else m, m if ta else k, s)
        b = O.uniform(n if tb else k, k if tb else n, s + 1)
        c0 = O.uniform(m, n, s + 2)
        got = run_dev(gpu, ta, tb, c["alpha"], a, b, c["beta"], c0.copy(order="F"))
        worst = max(worst, O.dmmch(ta, tb, c["alpha"], a, b, c["beta"], c0, got))
    assert worst < 16.0, worst


def test_gemm_integer_exact_and_alpha0(gpu):
    rng = np.random.default_rng(5)
    for (ta, tb) in [(0, 0), (1, 1)]:
        m, n, k = 150, 90, 70
        a = np.asfortranarray(rng.integers(-16, 17, (k, m) if ta else (m, k)).astype(float))
        b = np.asfortranarray(rng.integers(-16, 17, (n, k) if tb else (k, n)).astype(float))
        c = np.asfortranarray(rng.integers(-16, 17, (m, n)).astype(float))
        want = O.dgemm(ta, tb, 2.0, a, b, -1.0, c.copy(order="F"))
        got = run_dev(gpu, ta, tb, 2.0, a, b, -1.0, c.copy(order="F"))
        assert np.array_equal(got, want)
    c = O.uniform(70, 50, 9)
    got = run_dev(gpu, 0, 0, 0.0, O.uniform(70, 10, 1), O.uniform(10, 50, 2), 1.25, c.copy(order="F"))
    assert np.array_equal(got, c * 1.25)


def test_gemm_beta_zero_conventions(gpu):
    a, b = O.uniform(40, 30, 1), O.uniform(30, 20, 2)
    nan_c = np.asfortranarray(np.full((40, 20), np.nan))
    # flx path: the reference's generic order computes 0*C -> NaN propagates
    got = run_dev(gpu, 0, 0, 1.0, a, b, 0.0, nan_c.copy(order="F"))
    assert np.isnan(got).all()
    # CBLAS seam: BLAS convention, beta = 0 does not read C
    c = nan_c.copy(order="F")
    gpu.blas.mm(0, 0, 1.0, a, b, 0.0, c)
    want = O.dgemm(0, 0, 1.0, a, b, 0.0, np.zeros((40, 20), order="F"))
    assert np.all(np.abs(c - want) <= bound(0, 0, 1.0, a, b, 0.0, np.zeros((40, 20))))


def test_cblas_row_major_and_host_path(gpu):
    import ctypes as C
    m, n, k = 33, 45, 29
    a = O.uniform(m, k, 3)
    b = O.uniform(k, n, 4)
    c = O.uniform(m, n, 5)
    want = O.dgemm(0, 0, 0.5, a, b, 2.0, c.copy(order="F"))
    # row-major storage of the same matrices
    ar, br, cr = np.ascontiguousarray(a), np.ascontiguousarray(b), np.ascontiguousarray(c)
    dp = C.POINTER(C.c_double)
    gpu.lib().cblas_dgemm(101, 111, 111, m, n, k, 0.5, ar.ctypes.data_as(dp), k,
                          br.ctypes.data_as(dp), n, 2.0, cr.ctypes.data_as(dp), n)
    assert np.all(np.abs(cr - want) <= bound(0, 0, 0.5, a, b, 2.0, c))


def test_gemm_2048_config(gpu):
    """The DGEMM workload (M=N=K=2048, C = 1.5AB - 0.5C): full-matrix bound
    against an independent FP64 product (torch/rocBLAS), plus the sampled
    entries the compiled reference produced (tests/golden/gemm2048.json)."""
    import torch
    n = 2048
    A = gpu.DeviceMatrix.empty(n, n)
    B = gpu.DeviceMatrix.empty(n, n)
    Cd = gpu.DeviceMatrix.empty(n, n)
    h = gpu.default_handle()
    L = gpu.lib()
    for M, s in ((A, 11), (B, 12), (Cd, 13)):
        assert L.flx_fill_uniform(h.ptr, gpu._vp(M.ptr), n, n, n, s, 0) == 0
    a = A.storage.T.clone()
    b = B.storage.T.clone()
    c0 = Cd.storage.T.clone()
    # device generator == host generator, bit for bit (spot rows)
    assert np.array_equal(a[:3].cpu().numpy(), O.uniform(n, n, 11)[:3])
    gpu.blas.mm(0, 0, 1.5, A, B, -0.5, Cd)
    got = Cd.storage.T
    want = 1.5 * (a @ b) - 0.5 * c0
    g = 1.5 * (a.abs() @ b.abs()) + 0.5 * c0.abs()
    assert bool(((got - want).abs() <= 2 * 4 * n * EPS * g).all())
    path = os.path.join(GOLD, "gemm2048.json")
    if os.path.exists(path):
        meta = json.load(open(path))
        gc = got.cpu().numpy()
        gg = g.cpu().numpy()
        for i, j, v in meta["samples"]:
            assert abs(gc[i, j] - v) <= 4 * n * EPS * gg[i, j], (i, j)
