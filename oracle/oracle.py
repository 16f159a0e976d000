"""CPU oracle bindings -- TEST INFRASTRUCTURE ONLY.

ctypes front-end over ``oracle/liboracle.so`` (the plain-C restatement of the
FLENS generic GEMM / LU path in ``flens_oracle.c``) and, where it was built,
``oracle/_ref/libflensref.so`` (the reference itself, compiled from its own
headers).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module; the product package never does.

All matrices are numpy float64 arrays in Fortran (column-major) order;
pivot vectors are 1-based int32, as in LAPACK/FLENS.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORC = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libflensref.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_i, _d, _u64 = C.c_int, C.c_double, C.c_uint64


def build() -> None:
    """Compile liboracle.so (and _ref where the reference is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path: str):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_ORC):
            build()
        _lib = C.CDLL(_ORC)
        _lib.orc_fill_uniform.argtypes = [_dp, _i, _i, _i, _u64, _u64]
        _lib.orc_dgemm.argtypes = [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]
        _lib.orc_dgetrf_nb.argtypes = [_i, _i, _dp, _i, _ip, _i]
        _lib.orc_dgetrf_nb.restype = _i
        _lib.orc_dgetrf_gaps.argtypes = [_i, _i, _dp, _i, _ip, _i, _dp]
        _lib.orc_dgetrf_gaps.restype = _i
        _lib.orc_sgetrf_nb.argtypes = [_i, _i, _fp, _i, _ip, _i]
        _lib.orc_sgetrf_nb.restype = _i
        _lib.orc_dgetrs_c.argtypes = [_i, _i, _i, _dp, _i, _ip, _dp, _i]
        _lib.orc_sgetrs_c.argtypes = [_i, _i, _i, _fp, _i, _ip, _fp, _i]
        _lib.orc_dgesv_nb.argtypes = [_i, _i, _dp, _i, _ip, _dp, _i, _i]
        _lib.orc_dgesv_nb.restype = _i
        _lib.orc_dlaswp.argtypes = [_i, _dp, _i, _i, _i, _ip, _i]
        _lib.orc_dget01.argtypes = [_i, _i, _dp, _i, _dp, _i, _ip]
        _lib.orc_dget01.restype = _d
        _lib.orc_dget02.argtypes = [_i, _i, _dp, _i, _dp, _i, _dp, _i]
        _lib.orc_dget02.restype = _d
        _lib.orc_dmmch.argtypes = [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _dp, _i]
        _lib.orc_dmmch.restype = _d
        _lib.orc_dsgesv.argtypes = [_i, _i, _dp, _i, _ip, _dp, _i, _dp, _i, _i, _ip]
        _lib.orc_dsgesv.restype = _i
        _lib.orc_dgetrf_steps.argtypes = [_i, _i, _dp, _i, _i, _i]
        _lib.orc_dlange.argtypes = [C.c_char, _i, _i, _dp, _i]
        _lib.orc_dlange.restype = _d
        _lib.orc_dgeequ.argtypes = [_i, _i, _dp, _i, _dp, _dp, _dp, _dp, _dp]
        _lib.orc_dgeequ.restype = _i
        _lib.orc_dlaqge.argtypes = [_i, _i, _dp, _i, _dp, _dp, _d, _d, _d]
        _lib.orc_dlaqge.restype = _i
        _lib.orc_dgecon.argtypes = [C.c_char, _i, _dp, _i, _d]
        _lib.orc_dgecon.restype = _d
        _lib.orc_dgerfs.argtypes = [_i, _i, _i, _dp, _i, _dp, _i, _ip, _dp, _i, _dp, _i, _dp,
                                    _dp]
        _lib.orc_dgesvx.argtypes = [C.c_char, _i, _i, _i, _dp, _i, _dp, _i, _ip,
                                    C.POINTER(C.c_char), _dp, _dp, _dp, _i, _dp, _i, _dp, _dp,
                                    _dp, _dp, _i]
        _lib.orc_dgesvx.restype = _i
        _lib.orc_dgetri.argtypes = [_i, _dp, _i, _ip]
        _lib.orc_dgetri.restype = _i
        _lib.orc_dtrsm.argtypes = [_i, _i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]
    return _lib


def have_ref() -> bool:
    """Whether the compiled reference exists here (checked without loading
    it, so collecting a test module never maps it into the process)."""
    return os.path.exists(_REF)


def ref():
    """The compiled reference (None where it was not built)."""
    global _ref
    if _ref is None:
        _ref = _load(_REF)
        if _ref is None:
            return None
        _ref.ref_dgemm.argtypes = [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]
        _ref.ref_dgetrf.argtypes = [_i, _i, _dp, _i, _ip, _i]
        _ref.ref_dgetrf.restype = _i
        _ref.ref_dgetrs.argtypes = [_i, _i, _i, _dp, _i, _ip, _dp, _i]
        _ref.ref_dgesv.argtypes = [_i, _i, _dp, _i, _ip, _dp, _i, _i]
        _ref.ref_dgesv.restype = _i
        _ref.ref_dlaswp.argtypes = [_i, _i, _dp, _i, _i, _i, _ip, _i]
        _ref.ref_dlange.argtypes = [C.c_char, _i, _i, _dp, _i]
        _ref.ref_dlange.restype = _d
        _ref.ref_dgeequ.argtypes = [_i, _i, _dp, _i, _dp, _dp, _dp, _dp, _dp]
        _ref.ref_dgeequ.restype = _i
        _ref.ref_dlaqge.argtypes = [_i, _i, _dp, _i, _dp, _dp, _d, _d, _d]
        _ref.ref_dlaqge.restype = _i
        _ref.ref_dgecon.argtypes = [C.c_char, _i, _dp, _i, _d]
        _ref.ref_dgecon.restype = _d
        _ref.ref_dgerfs.argtypes = [_i, _i, _i, _dp, _i, _dp, _i, _ip, _dp, _i, _dp, _i, _dp,
                                    _dp]
        _ref.ref_dgesvx.argtypes = [C.c_char, _i, _i, _i, _dp, _i, _dp, _i, _ip, C.c_char,
                                    _dp, _dp, _dp, _i, _dp, _i, _dp, _dp, _dp, _dp, _i]
        _ref.ref_dgesvx.restype = _i
        _ref.ref_dgetri.argtypes = [_i, _dp, _i, _ip]
        _ref.ref_dgetri.restype = _i
        _ref.ref_dtrsm.argtypes = [_i, _i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]
    return _ref


def _p(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_fp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def _ld(a):
    """Leading dimension of a Fortran-ordered 2-D array."""
    assert a.flags.f_contiguous or a.ndim == 1
    return max(1, a.shape[0])


# ------------------------------------------------------------------- RNG ---
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(m: int, n: int, seed: int, offset: int = 0) -> np.ndarray:
    """m x n column-major matrix, element (i,j) = stream[offset + j*m + i],
    exact dyadic values in [-1,1).  Vectorised numpy twin of
    orc_fill_uniform and of the library's device generator."""
    idx = np.arange(m * n, dtype=np.uint64) + np.uint64(offset)
    with np.errstate(over="ignore"):
        z = _mix64(np.uint64(seed) + (idx + np.uint64(1)) * _GOLDEN)
    k = (z >> np.uint64(11)).astype(np.int64) - (np.int64(1) << np.int64(52))
    v = k.astype(np.float64) * 2.0 ** -52
    return np.asfortranarray(v.reshape(n, m).T)


# ------------------------------------------------------------ oracle API ---
def dgemm(transa, transb, alpha, a, b, beta, c, use_ref=False):
    """C <- beta*C + alpha*op(A)*op(B) in place (gemm_generic order)."""
    L = ref() if use_ref else lib()
    m, n = c.shape
    k = a.shape[0] if transa else a.shape[1]
    L_fn = L.ref_dgemm if use_ref else L.orc_dgemm
    L_fn(int(transa), int(transb), m, n, k, float(alpha), _p(a), _ld(a), _p(b), _ld(b),
         float(beta), _p(c), _ld(c))
    return c


def dgetrf(a, nb=64, use_ref=False):
    """In-place LU (trf); returns (ipiv 1-based int32, info)."""
    m, n = a.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    if use_ref:
        info = ref().ref_dgetrf(m, n, _p(a), _ld(a), _p(ipiv), nb)
    else:
        info = lib().orc_dgetrf_nb(m, n, _p(a), _ld(a), _p(ipiv), nb)
    return ipiv[: min(m, n)], info


def dgetrf_steps(a, nb, nsteps):
    """Run only the first nsteps block steps of the blocked getrf."""
    m, n = a.shape
    lib().orc_dgetrf_steps(m, n, _p(a), _ld(a), nb, nsteps)
    return a


def dgetrf_step_at(a, i0: int, j0: int, rows: int, cols: int, nb: int):
    """One block step of the blocked getrf on the sub-block
    a[i0:i0+rows, j0:j0+cols] of the Fortran-ordered square array a, with
    a's own leading dimension (timing samples of the step at depth j0)."""
    assert a.flags.f_contiguous and a.dtype == np.float64
    n = a.shape[0]
    assert 0 <= i0 and 0 <= j0 and rows > 0 and cols > 0
    assert i0 + rows <= n and j0 + cols <= a.shape[1]
    p = C.cast(C.c_void_p(a.ctypes.data + 8 * (i0 + j0 * n)), _dp)
    lib().orc_dgetrf_steps(rows, cols, p, n, nb, 1)


def getrf_steps_flops(m: int, n: int, nb: int, nsteps: int) -> float:
    """Exact flop count of the first nsteps block steps (panel getf2 with
    one reciprocal-multiply per subdiagonal entry, unit-lower trsm, GEMM)."""
    mn = min(m, n)
    fl = 0.0
    j = 0
    for _ in range(nsteps):
        if j >= mn:
            break
        jb = min(nb, mn - j)
        mp = m - j
        for c in range(jb):
            below = mp - c - 1
            fl += below + 2.0 * below * (jb - c - 1)
        nr = n - j - jb
        if nr > 0:
            fl += float(jb) * (jb - 1) * nr
            fl += 2.0 * (m - j - jb) * nr * jb
        j += jb
    return fl


def dgetrf_gaps(a, nb=64):
    m, n = a.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    gap = np.zeros(max(1, min(m, n)), dtype=np.float64)
    info = lib().orc_dgetrf_gaps(m, n, _p(a), _ld(a), _p(ipiv), nb, _p(gap))
    return ipiv[: min(m, n)], info, gap[: min(m, n)]


def sgetrf(a, nb=64):
    m, n = a.shape
    ipiv = np.zeros(max(1, min(m, n)), dtype=np.int32)
    info = lib().orc_sgetrf_nb(m, n, _p(a), _ld(a), _p(ipiv), nb)
    return ipiv[: min(m, n)], info


def dgetrs(trans, a, ipiv, b, use_ref=False):
    n = a.shape[0]
    nrhs = b.shape[1]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    if use_ref:
        ref().ref_dgetrs(int(trans), n, nrhs, _p(a), _ld(a), _p(ipiv), _p(b), _ld(b))
    else:
        lib().orc_dgetrs_c(int(trans), n, nrhs, _p(a), _ld(a), _p(ipiv), _p(b), _ld(b))
    return b


def sgetrs(trans, a, ipiv, b):
    n = a.shape[0]
    nrhs = b.shape[1]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_sgetrs_c(int(trans), n, nrhs, _p(a), _ld(a), _p(ipiv), _p(b), _ld(b))
    return b


def dgesv(a, b, nb=64, use_ref=False):
    n = a.shape[0]
    ipiv = np.zeros(max(1, n), dtype=np.int32)
    if use_ref:
        info = ref().ref_dgesv(n, b.shape[1], _p(a), _ld(a), _p(ipiv), _p(b), _ld(b), nb)
    else:
        info = lib().orc_dgesv_nb(n, b.shape[1], _p(a), _ld(a), _p(ipiv), _p(b), _ld(b), nb)
    return ipiv[:n], info


def dlaswp(a, k1, k2, ipiv, incx, use_ref=False):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    m, n = a.shape
    if use_ref:
        ref().ref_dlaswp(m, n, _p(a), _ld(a), k1, k2, _p(ipiv), incx)
    else:
        lib().orc_dlaswp(n, _p(a), _ld(a), k1, k2, _p(ipiv), incx)
    return a


def dget01(a0, lu, ipiv):
    m, n = a0.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return lib().orc_dget01(m, n, _p(a0), _ld(a0), _p(lu), _ld(lu), _p(ipiv))


def dget02(a, x, b):
    n = a.shape[0]
    return lib().orc_dget02(n, x.shape[1], _p(a), _ld(a), _p(x), _ld(x), _p(b), _ld(b))


def dmmch(transa, transb, alpha, a, b, beta, c0, c):
    m, n = c.shape
    k = a.shape[0] if transa else a.shape[1]
    return lib().orc_dmmch(int(transa), int(transb), m, n, k, float(alpha), _p(a), _ld(a),
                           _p(b), _ld(b), float(beta), _p(c0), _p(c), _ld(c))


def dtrsm(order, right, upper, trans, unit, alpha, a, b, use_ref=False):
    """B <- alpha * op(A)^{-1} B (left) or alpha * B op(A)^{-1} (right), in
    place, trsm_generic's order.  order 'F': a, b column-major arrays;
    order 'C': row-major (C-ordered) arrays, as a CBLAS RowMajor caller
    holds them."""
    if order == "F":
        assert a.flags.f_contiguous and b.flags.f_contiguous
        m, n = b.shape
        lda, ldb, o = max(1, a.shape[0]), max(1, m), 1
    else:
        assert a.flags.c_contiguous and b.flags.c_contiguous
        m, n = b.shape
        lda, ldb, o = max(1, a.shape[1]), max(1, n), 0
    fn = ref().ref_dtrsm if use_ref else lib().orc_dtrsm
    fn(o, int(right), int(upper), int(trans), int(unit), m, n, float(alpha), _p(a), lda,
       _p(b), ldb)
    return b


def dsgesv(a, b, itermax=30):
    """netlib dsgesv semantics; returns (x, ipiv, info, iter)."""
    n = a.shape[0]
    nrhs = b.shape[1]
    x = np.zeros((n, nrhs), order="F")
    ipiv = np.zeros(max(1, n), dtype=np.int32)
    it = C.c_int(0)
    info = lib().orc_dsgesv(n, nrhs, _p(a), _ld(a), _p(ipiv), _p(b), _ld(b), _p(x), _ld(x),
                            itermax, C.byref(it))
    return x, ipiv[:n], info, it.value


# ------------------------------------------ norms, equilibration, rcond ---
def _impl(use_ref):
    return (ref(), "ref_") if use_ref else (lib(), "orc_")


def dlange(norm, a, use_ref=False):
    """flens::lapack::lan (GeMatrix): norm 'M', 'O', 'I' or 'F'."""
    L, p = _impl(use_ref)
    m, n = a.shape
    return getattr(L, p + "dlange")(norm.encode(), m, n, _p(a), _ld(a))


def dgeequ(a, use_ref=False):
    """flens::lapack::equ -> (r, c, rowcnd, colcnd, amax, info); rowcnd and
    colcnd are NaN where the routine leaves them unset."""
    L, p = _impl(use_ref)
    m, n = a.shape
    r = np.zeros(max(1, m))
    c = np.zeros(max(1, n))
    rc, cc, am = C.c_double(np.nan), C.c_double(np.nan), C.c_double(np.nan)
    info = getattr(L, p + "dgeequ")(m, n, _p(a), _ld(a), _p(r), _p(c), C.byref(rc),
                                    C.byref(cc), C.byref(am))
    return r[:m], c[:n], rc.value, cc.value, am.value, info


def dlaqge(a, r, c, rowcnd, colcnd, amax, use_ref=False):
    """flens::lapack::laq: scales a in place, returns 'N', 'R', 'C' or 'B'."""
    L, p = _impl(use_ref)
    m, n = a.shape
    r = np.ascontiguousarray(r, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    return chr(getattr(L, p + "dlaqge")(m, n, _p(a), _ld(a), _p(r), _p(c), float(rowcnd),
                                        float(colcnd), float(amax)))


def dgecon(norm, lu, anorm, use_ref=False):
    """flens::lapack::con on an LU factor: reciprocal condition number."""
    L, p = _impl(use_ref)
    n = lu.shape[0]
    return getattr(L, p + "dgecon")(norm.encode(), n, _p(lu), _ld(lu), float(anorm))


def dgerfs(trans, a, af, ipiv, b, x, use_ref=False):
    """flens::lapack::rfs: refines x in place; returns (ferr, berr)."""
    L, p = _impl(use_ref)
    n, nrhs = b.shape
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    getattr(L, p + "dgerfs")(int(trans), n, nrhs, _p(a), _ld(a), _p(af), _ld(af), _p(ipiv),
                             _p(b), _ld(b), _p(x), _ld(x), _p(ferr), _p(berr))
    return ferr[:nrhs], berr[:nrhs]


def dgesvx(fact, trans, a, af, ipiv, equed, r, c, b, nb=64, use_ref=False):
    """flens::lapack::svx.  a, af, ipiv, r, c, b are updated in place.
    Returns dict(x, rcond, ferr, berr, rpvgrw, equed, info); the compiled
    reference cannot report equed (it takes it by value)."""
    n, nrhs = b.shape
    x = np.zeros((n, nrhs), order="F")
    ferr = np.zeros(max(1, nrhs))
    berr = np.zeros(max(1, nrhs))
    rc, g = C.c_double(0.0), C.c_double(0.0)
    if use_ref:
        info = ref().ref_dgesvx(fact.encode(), int(trans), n, nrhs, _p(a), _ld(a), _p(af),
                                _ld(af), _p(ipiv), equed.encode(), _p(r), _p(c), _p(b), _ld(b),
                                _p(x), _ld(x), C.byref(rc), _p(ferr), _p(berr), C.byref(g), nb)
        eq = None
    else:
        e = C.c_char(equed.encode())
        info = lib().orc_dgesvx(fact.encode(), int(trans), n, nrhs, _p(a), _ld(a), _p(af),
                                _ld(af), _p(ipiv), C.byref(e), _p(r), _p(c), _p(b), _ld(b),
                                _p(x), _ld(x), C.byref(rc), _p(ferr), _p(berr), C.byref(g), nb)
        eq = e.value.decode()
    return dict(x=x, rcond=rc.value, ferr=ferr[:nrhs], berr=berr[:nrhs], rpvgrw=g.value,
                equed=eq, info=info)


def dgetri(a, ipiv, use_ref=False):
    """flens::lapack::tri (GeMatrix): inv(A) in place from the LU factor;
    returns info."""
    L, p = _impl(use_ref)
    n = a.shape[0]
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return getattr(L, p + "dgetri")(n, _p(a), _ld(a), _p(ipiv))
