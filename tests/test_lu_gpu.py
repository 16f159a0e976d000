"""GPU parity for the FP64 LU path (flx_dgetrf / flx_dgetrs / flx_dgesv and
the dgetrf_ / dgetrs_ / dgesv_ / dlaswp_ seams) against the oracle and the
reference's own fixtures.

Bar: pivot vectors identical to the reference's on
columns without near-ties (the first mismatch must sit on a column whose
relative top-2 gap is below tau(n) = 64 n eps64: two summation orders of the
j-th column differ by about j eps64 relative), LAPACK dget01 ratio
||PA - LU||_1 / (n ||A||_1 eps) < 30, dget02 solve ratio < 30, INFO and
pivot conventions of flens/lapack/ge/tf2.tcc:101-126 on exact zeros.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EPS = 2.0 ** -53


def tau(n):
    """Relative top-2 pivot gap below which a summation-order flip is
    legitimate: the GPU and the reference sum a column's j updates in
    different orders, a relative difference of order j eps64."""
    return 64.0 * max(n, 1) * EPS


def first_mismatch_ok(piv_gpu, piv_ref, gap):
    """(ok, first mismatching column or -1).  ok: identical, or the first
    difference sits on a near-tie column (later columns then legitimately
    follow a different elimination)."""
    piv_gpu, piv_ref = np.asarray(piv_gpu), np.asarray(piv_ref)
    d = np.nonzero(piv_gpu != piv_ref)[0]
    if d.size == 0:
        return True, -1
    j = int(d[0])
    return bool(gap[j] < tau(len(piv_ref))), j


def dev_getrf(F, a, nb=None, ib=None):
    A = F.DeviceMatrix.from_numpy(a)
    h = F.default_handle()
    if nb is not None:
        h.set_lu_blocking(nb, ib or 32)
    piv, info = F.lapack.trf(A)
    out = A.numpy(), piv.cpu().numpy(), int(info.cpu().numpy()[0])
    if nb is not None:
        h.set_lu_blocking(256, 32)
    return out


def test_lu_grid_vs_reference(gpu):
    with open(os.path.join(GOLD, "small.json")) as f:
        cases = json.load(f)["lu"]
    for c in cases:
        m, n = c["m"], c["n"]
        if c["nb"] != 64:
            continue
        a0 = O.uniform(m, n, c["seed"])
        lu, piv, info = dev_getrf(gpu, a0)
        piv = piv[: min(m, n)]
        _, _, gap = O.dgetrf_gaps(a0.copy(order="F"))
        ok, j = first_mismatch_ok(piv, c["piv"], gap)
        assert ok, (m, n, j)
        assert info == c["info"]
        assert O.dget01(a0, lu, piv) < 30, (m, n)


@pytest.mark.parametrize("nb,ib", [(64, 32), (128, 32), (256, 32), (64, 64), (256, 64), (32, 32)])
@pytest.mark.parametrize("look", [1, 0])
def test_lu_blockings_multi_workgroup_panel(gpu, nb, ib, look):
    # heights above 512 rows use several cooperating workgroups per panel;
    # look = 0 runs the blocked LU without lookahead (no second stream)
    h = gpu.default_handle()
    h.set_lu_lookahead(look)
    try:
        for (m, n, seed) in [(700, 700, 1), (1030, 1030, 2), (600, 300, 3), (300, 650, 4)]:
            a0 = O.uniform(m, n, 40 + seed)
            lu, piv, info = dev_getrf(gpu, a0, nb, ib)
            piv = piv[: min(m, n)]
            ref = a0.copy(order="F")
            rpiv, rinfo, gap = O.dgetrf_gaps(ref, nb=64)
            ok, j = first_mismatch_ok(piv, rpiv, gap)
            assert ok and info == rinfo == 0, (m, n, nb, ib, look, j)
            assert O.dget01(a0, lu, piv) < 30
    finally:
        h.set_lu_lookahead(1)


def test_lu_nan_pivot_rule(gpu):
    """iamax's NaN rule (cxxblas/level1/iamax.tcc:42-58): a NaN is the pivot
    only at the first live position of its column and is skipped elsewhere.
    Pivots, INFO and the NaN pattern of L\\U must equal the oracle's, which
    follows the reference's loop; the finite entries agree to rounding."""
    n = 1100
    cases = []
    a = O.uniform(n, n, 77); a[700, 0] = np.nan; cases.append(a)      # skipped, spreads along a row
    a = O.uniform(n, n, 78); a[0, 0] = np.nan; cases.append(a)        # first position: chosen
    a = O.uniform(n, n, 79); a[5, 300] = np.nan; a[900, 300] = np.nan; cases.append(a)
    a = O.uniform(300, 300, 80); a[:, 10] = np.nan; cases.append(a)  # a whole NaN column
    for k, a0 in enumerate(cases):
        lu, piv, info = dev_getrf(gpu, a0)
        ref = a0.copy(order="F")
        rpiv, rinfo = O.dgetrf(ref)
        m = a0.shape[0]
        assert info == rinfo, (k, info, rinfo)
        d = np.nonzero(piv[:m] != rpiv)[0]
        assert d.size == 0, (k, int(d[0]), int(piv[d[0]]), int(rpiv[d[0]]))
        assert np.array_equal(np.isnan(lu), np.isnan(ref)), k
        fin = np.isfinite(ref)
        assert np.allclose(lu[fin], ref[fin], rtol=1e-9, atol=1e-9), k


def test_lu_exact_zero_info_and_ties(gpu):
    sp = json.load(open(os.path.join(GOLD, "small.json")))["special"]
    for key in ("zero_col", "rank_def", "zero_row", "tie"):
        flat = np.array(sp[key]["a"], float)
        k = int(round(len(flat) ** 0.5))
        a = np.array(flat.reshape(k, k).T, order="F")   # a copy, not a view of flat
        piv, info = gpu.lapack.trf(a)          # host seam dgetrf_
        assert piv.tolist() == sp[key]["piv"] and info == sp[key]["info"], key
        lu, dpiv, dinfo = dev_getrf(gpu, np.array(flat.reshape(k, k).T, order="F"))
        if not (dpiv[:k].tolist() == sp[key]["piv"] and dinfo == sp[key]["info"]):
            again = dev_getrf(gpu, np.array(flat.reshape(k, k).T, order="F"))
            raise AssertionError(f"{key}: piv {dpiv.tolist()} info {dinfo} lu {lu.tolist()}; "
                                 f"retry piv {again[1].tolist()} info {again[2]}")
    # a zero column deep inside a multi-workgroup panel
    a = O.uniform(900, 900, 8)
    a[:, 37] = 0.0
    lu, piv, info = dev_getrf(gpu, a)
    ref = a.copy(order="F")
    rpiv, rinfo = O.dgetrf(ref)
    assert info == rinfo == 38


def test_known_answer_host_seam(gpu):
    ab = np.asfortranarray(np.array([[2, 3, -1, 0, 20], [-6, -5, 0, 2, -33],
                                     [2, -5, 6, -6, -43], [4, 6, 2, -3, 49]], float))
    piv, info = gpu.lapack.trf(ab)
    assert piv.tolist() == [2, 3, 4, 4] and info == 0
    x = np.linalg.solve(np.triu(ab[:, :4]), ab[:, 4])
    assert np.allclose(x, [1, 9, 9, 9], atol=1e-12)
    # dgesv_ on the same system
    a = np.asfortranarray(np.array([[2, 3, -1, 0], [-6, -5, 0, 2], [2, -5, 6, -6], [4, 6, 2, -3]], float))
    b = np.asfortranarray(np.array([[20], [-33], [-43], [49]], float))
    piv, info = gpu.lapack.sv(a, b)
    assert info == 0 and piv.tolist() == [2, 3, 4, 4]
    assert np.allclose(b[:, 0], [1, 9, 9, 9], atol=1e-12)


@pytest.mark.parametrize("n", [1, 7, 64, 200, 513, 1100])
def test_getrs_and_gesv(gpu, n):
    a0 = O.uniform(n, n, 300 + n)
    for nrhs in (1, 2, 15, 16, 37):
        b0 = O.uniform(n, nrhs, 17 + n)
        # device trf + trs (both transposes)
        A = gpu.DeviceMatrix.from_numpy(a0)
        piv, info = gpu.lapack.trf(A)
        for trans in (0, 1):
            B = gpu.DeviceMatrix.from_numpy(b0)
            gpu.lapack.trs(trans, A, piv, B)
            x = B.numpy()
            at = a0 if trans == 0 else np.asfortranarray(a0.T)
            assert O.dget02(at, x, b0) < 30, (n, nrhs, trans)
        # device gesv
        A = gpu.DeviceMatrix.from_numpy(a0)
        B = gpu.DeviceMatrix.from_numpy(b0)
        piv2, info2 = gpu.lapack.sv(A, B)
        assert info2 == 0 and O.dget02(a0, B.numpy(), b0) < 30
        # host seams
        a = a0.copy(order="F")
        b = b0.copy(order="F")
        hp, hi = gpu.lapack.sv(a, b)
        assert hi == 0 and O.dget02(a0, b, b0) < 30
        b = b0.copy(order="F")
        gpu.lapack.trs(1, a, hp, b)
        assert O.dget02(np.asfortranarray(a0.T), b, b0) < 30


def test_gesv_singular_skips_solve(gpu):
    a = O.uniform(50, 50, 3)
    a[:, 10] = 0.0
    b = O.uniform(50, 2, 4)
    b_before = b.copy(order="F")
    piv, info = gpu.lapack.sv(a.copy(order="F"), b)
    assert info == 11
    assert np.array_equal(b, b_before)


@pytest.mark.parametrize("n,nrhs", [(300, 3), (1100, 20)])
def test_device_gesv_padded_leading_dimensions(gpu, n, nrhs):
    """flx_dgesv on views with lda > n and ldb > n: the factors and pivots
    equal flx_dgetrf's bit for bit (B is carried through the same
    factorisation as extra columns), X solves the system (dget02 < 30), and
    the padding rows are never written; a singular A leaves B untouched."""
    import ctypes as C
    import torch
    lda, ldb = n + 5, n + 3
    a0 = O.uniform(n, n, 900 + n)
    b0 = O.uniform(n, nrhs, 901 + n)

    def padded(x, ld, fill):
        buf = np.full((ld, x.shape[1]), fill, order="F")
        buf[: x.shape[0], :] = x
        return torch.from_numpy(np.ascontiguousarray(buf.T)).to("cuda")

    def unpad(t, rows):
        return np.asfortranarray(t.cpu().numpy().T[:rows, :])

    L = gpu.lib()
    h = gpu.default_handle()
    tA, tB = padded(a0, lda, 7.5), padded(b0, ldb, -3.25)
    ipiv = gpu._int_buffer(n)
    info = C.c_int(-1)
    assert L.flx_dgesv(h.ptr, n, nrhs, gpu._vp(tA.data_ptr()), lda, gpu._vp(ipiv.data_ptr()),
                       gpu._vp(tB.data_ptr()), ldb, C.byref(info)) == 0
    assert info.value == 0
    full_a, full_b = tA.cpu().numpy().T, tB.cpu().numpy().T
    assert np.all(full_a[n:, :] == 7.5) and np.all(full_b[n:, :] == -3.25)
    x = unpad(tB, n)
    assert O.dget02(a0, x, b0) < 30
    # the factors and pivots are flx_dgetrf's
    A2 = gpu.DeviceMatrix.from_numpy(a0)
    piv2, info2 = gpu.lapack.trf(A2)
    assert int(info2.cpu()[0]) == 0
    assert np.array_equal(ipiv.cpu().numpy()[:n], piv2.cpu().numpy()[:n])
    assert np.array_equal(unpad(tA, n), A2.numpy())
    # singular: INFO > 0, B left as it was
    a1 = a0.copy(order="F")
    a1[:, 3] = 0.0
    tA, tB = padded(a1, lda, 0.0), padded(b0, ldb, 0.0)
    assert L.flx_dgesv(h.ptr, n, nrhs, gpu._vp(tA.data_ptr()), lda, gpu._vp(ipiv.data_ptr()),
                       gpu._vp(tB.data_ptr()), ldb, C.byref(info)) == 0
    assert info.value == 4
    assert np.array_equal(unpad(tB, n), b0)


def test_laswp_seam(gpu):
    a0 = O.uniform(30, 7, 5)
    piv = np.array([3, 5, 3, 30, 5, 6, 7, 29], np.int32)
    for incx in (1, -1):
        a, r = a0.copy(order="F"), a0.copy(order="F")
        gpu.lapack.laswp(a, 1, 8, piv, incx)
        O.dlaswp(r, 1, 8, piv, incx)
        assert np.array_equal(a, r)


def test_invalid_arguments_report_like_lapack(gpu, capfd):
    import ctypes as C
    L = gpu.lib()
    info = C.c_int(0)
    a = np.zeros((4, 4), order="F")
    piv = np.zeros(4, np.int32)
    L.dgetrf_(C.byref(C.c_int(4)), C.byref(C.c_int(4)), a.ctypes.data_as(gpu._dp),
              C.byref(C.c_int(2)), piv.ctypes.data_as(gpu._ip), C.byref(info))
    assert info.value == -4
    assert "DGETRF" in capfd.readouterr().err


@pytest.mark.parametrize("nb", [64, 128, 256])
def test_lu8192_vs_reference_fixture(gpu, nb):
    """The FP64 LU workload (n = 8192, A ~ U[-1,1) seed 2, nrhs = 16 seed 3):
    pivots against the compiled reference's, backward error by an
    independent FP64 product (torch/rocBLAS) on the device."""
    import torch
    path = os.path.join(GOLD, f"lu8192_nb{nb}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not generated")
    z = np.load(path)
    n = 8192
    h = gpu.default_handle()
    L = gpu.lib()
    A = gpu.DeviceMatrix.empty(n, n)
    assert L.flx_fill_uniform(h.ptr, gpu._vp(A.ptr), n, n, n, 2, 0) == 0
    a0 = A.storage.T.clone()
    h.set_lu_blocking(nb, 32)
    pivd, info = gpu.lapack.trf(A)
    h.set_lu_blocking(256, 32)
    piv = pivd.cpu().numpy()
    assert int(info.cpu()[0]) == 0
    ok, j = first_mismatch_ok(piv, z["piv"], z["gap"])
    what = "identical" if j < 0 else f"first differ at column {j} (gap {float(z['gap'][j]):.3g})"
    print(f"nb={nb}: pivots {what}")
    assert ok, (j, float(z["gap"][j]))
    # the solve of the same workload (B ~ U[-1,1) seed 3, nrhs = 16) against
    # the reference's first rows of X, LAPACK dget04's criterion:
    # ||x - x_ref|| / (||x_ref|| kappa eps) < 30 for each of the two
    # solutions (kappa from flx_dgecon on the factor)
    nrhs = 16
    B = gpu.DeviceMatrix.empty(n, nrhs)
    assert L.flx_fill_uniform(h.ptr, gpu._vp(B.ptr), n, nrhs, n, 3, 0) == 0
    gpu.lapack.trs(0, A, pivd, B)
    x = B.numpy()
    anorm = float(a0.abs().sum(1).max().item())
    rcond = gpu.lapack.con("I", A, anorm)
    xr = z["x_rows"]
    for c in range(nrhs):
        num = np.abs(x[: xr.shape[0], c] - xr[:, c]).max()
        den = np.abs(x[:, c]).max() * (1.0 / rcond) * 2.0 ** -53
        assert num / den < 60, (c, num / den, 1.0 / rcond)
    lu = A.storage.T
    l = torch.tril(lu, -1) + torch.eye(n, dtype=lu.dtype, device=lu.device)
    u = torch.triu(lu)
    prod = l @ u
    # P A: apply the interchanges to the rows of a0
    perm = np.arange(n)
    for i, q in enumerate(piv - 1):
        perm[i], perm[q] = perm[q], perm[i]
    pa = a0[torch.from_numpy(perm).to(a0.device)]
    r1 = (prod - pa).abs().sum(0).max().item()
    a1 = a0.abs().sum(0).max().item()
    ratio = r1 / (n * a1 * 2.0 ** -53)
    assert ratio < 30, ratio


def test_flens_program_runs_on_libflx(gpu):
    """The FLENS program of tests/dropin (FLENS API only, external
    BLAS/LAPACK switches, linked to libflx) runs its mm / trf / trs / sv on
    the GPU and passes its own checks (KAT pivots, GEMM bound, dget02)."""
    import subprocess
    exe = os.path.join(os.path.dirname(GOLD), "..", "oracle", "_ref", "flens_dropin")
    if not os.path.exists(exe):
        pytest.skip("drop-in program not built (needs the reference headers)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "DROPIN PASSED" in r.stdout, r.stdout + r.stderr


def test_lu_n20000_device_and_host_seam(gpu):
    """An order above 16384 (40 cooperating panel workgroups): the device
    LU's dget01 ratio ||P A - L U||_1 / (n ||A||_1 eps) is computed on the
    device in FP64 (the n^3 product is out of the oracle's reach), and the
    FLENS dgetrf_ seam on host memory computes the same factors and pivots
    instead of refusing the size."""
    import ctypes as C
    import torch
    n = 20000
    L = gpu.lib()
    h = gpu.default_handle()
    A0 = gpu.DeviceMatrix.empty(n, n)
    assert L.flx_fill_uniform(h.ptr, gpu._vp(A0.ptr), n, n, n, 2, 0) == 0
    A = gpu.DeviceMatrix.empty(n, n)
    A.storage.copy_(A0.storage)
    piv, info = gpu.lapack.trf(A)
    torch.cuda.synchronize()
    assert info == 0
    ipiv = piv.cpu().numpy() if hasattr(piv, "cpu") else np.asarray(piv)
    assert ipiv.min() >= 1 and ipiv.max() <= n and (ipiv[:n] >= np.arange(1, n + 1)).all()
    perm = np.arange(n)
    for i in range(n):
        k = ipiv[i] - 1
        perm[i], perm[k] = perm[k], perm[i]
    lu = A.storage.T
    Lm = torch.tril(lu, -1)
    Lm.diagonal().fill_(1.0)
    R = A0.storage.T[torch.from_numpy(perm).cuda()]
    R -= Lm @ torch.triu(lu)
    del Lm
    ratio = (R.abs().sum(0).max() / (n * A0.storage.T.abs().sum(0).max() * 2.0 ** -53)).item()
    del R
    print(f"n = {n}: dget01 = {ratio:.3f}")
    assert ratio < 30
    # the host seam: same factors and pivots, bit for bit
    a_host = np.asfortranarray(A0.storage.cpu().numpy().T)
    ip = np.zeros(n, dtype=np.int32)
    inf = C.c_int(-1)
    L.dgetrf_(C.byref(C.c_int(n)), C.byref(C.c_int(n)), a_host.ctypes.data_as(gpu._dp),
              C.byref(C.c_int(n)), ip.ctypes.data_as(gpu._ip), C.byref(inf))
    assert inf.value == 0
    assert np.array_equal(ip, ipiv[:n])
    assert np.array_equal(a_host, A.storage.cpu().numpy().T)


def test_two_handles_factor_concurrently(gpu):
    """Two factorisations in flight at once on two handles and streams.  A
    panel's cooperating workgroups wait on one another (bounded waits), so
    two launches sharing the device must both complete without a failed
    wait (flx_sync reports one as an error) and give the pivots and
    backward error of a lone run."""
    import torch
    n = 3000
    a = [O.uniform(n, n, 201), O.uniform(n, n, 202)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    handles = [gpu.Handle(s) for s in streams]
    pristine = [gpu.DeviceMatrix.from_numpy(x) for x in a]
    work = [gpu.DeviceMatrix.from_numpy(x) for x in a]
    torch.cuda.synchronize()
    for rep in range(3):
        out = []
        for k in range(2):
            with torch.cuda.stream(streams[k]):
                work[k].storage.copy_(pristine[k].storage)
            out.append(gpu.lapack.trf(work[k], handle=handles[k]))
        for h in handles:
            h.sync()
        torch.cuda.synchronize()
        for k in range(2):
            piv = out[k][0].cpu().numpy()[:n]
            assert int(out[k][1].cpu().numpy()[0]) == 0
            ref = a[k].copy(order="F")
            rpiv, rinfo, gap = O.dgetrf_gaps(ref)
            ok, j = first_mismatch_ok(piv, rpiv, gap)
            assert ok, (rep, k, j)
            assert O.dget01(a[k], work[k].numpy(), piv) < 30, (rep, k)
