"""flens_amd -- MI355X-native dense GEMM and LU solve behind FLENS's call surface.

The compute lives in ``libflx.so`` (hand-written HIP for gfx950, built in-tree
by ``flens_amd.build``); this module is a thin ctypes mirror of its C ABI
(``include/flx.h``) with the reference's names:

* ``blas.mm``       -- flens::blas::mm        (flens/blas/level3/mm.tcc:380-479)
* ``lapack.trf``    -- flens::lapack::trf     (flens/lapack/ge/trf.tcc:183-261)
* ``lapack.trs``    -- flens::lapack::trs     (flens/lapack/ge/trs.tcc:151-252)
* ``lapack.sv``     -- flens::lapack::sv      (flens/lapack/ge/sv.tcc:115-219)
* ``lapack.laswp``  -- flens::lapack::laswp   (flens/lapack/la/laswp.tcc:126-186)
* ``lapack.dsgesv`` -- cxxlapack::sgesv -> dsgesv_ (cxxlapack/netlib/interface/
  lapack.in.h:10611-10623): FP32 factor + FP64 refinement, LAPACK semantics
* ``lapack.sv_mixed`` -- the mixed-precision solver with a choice of factor
  precision (FP32 / BF16 / FP16-scaled) and refinement (classic / GMRES-IR)

Two data paths, exactly as in the C ABI:

* host numpy arrays (Fortran order) go through the FLENS plugin seams
  (``cblas_dgemm``, ``dgetrf_``, ``dgetrs_``, ``dgesv_``, ``dlaswp_``);
* device arrays (:class:`DeviceMatrix`, column-major views of torch CUDA
  storage) go through the asynchronous ``flx_*`` entry points.

There is no CPU fallback: importing works anywhere, but every compute call
raises if ``libflx.so`` is missing or no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflx.so")

# every symbol include/flx.h declares
EXPORTS = [
    "flx_create", "flx_destroy", "flx_set_stream", "flx_sync", "flx_status_string",
    "flx_set_lu_blocking", "flx_set_lu_lookahead", "flx_set_lu_side_groups", "flx_dgemm", "flx_dgetrf", "flx_dgetrs", "flx_dgesv",
    "flx_dlaswp", "flx_fill_uniform", "flx_profile_start", "flx_profile_stop",
    "flx_profile_classes", "flx_test_raise_sync_flag", "flx_dgetrf_noalloc", "flx_noalloc_bytes",
    "flx_probe_mfma", "flx_probe_copy", "flx_batch_bin_tiles",
    "flx_dgesv_mixed", "flx_dsgesv",
    "flx_batch_gesv_create", "flx_batch_gemm_create", "flx_batch_execute", "flx_batch_config",
    "flx_batch_destroy", "flx_mlp_forward", "flx_mlp_grad", "flx_mlp_adam",
    "flx_dlange", "flx_dgeequ", "flx_dlaqge", "flx_dgecon", "flx_dgerfs", "flx_dgesvx",
    "flx_dgetri", "flx_save_gematrix", "flx_load_gematrix_header", "flx_load_gematrix",
    "flx_save_densevector", "flx_load_densevector",
    "cblas_dgemm", "cblas_dtrsm", "dgemm_", "dgetrf_", "dgetrs_", "dgesv_", "dlaswp_", "dsgesv_",
    "dlange_", "dgeequ_", "dgecon_", "dgerfs_", "dgesvx_", "dgetri_",
    "xerbla_",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i, _d, _vp, _u64 = C.c_int, C.c_double, C.c_void_p, C.c_uint64

NOALLOC_SCRATCH_MIB = 36   # FLX_NOALLOC_SCRATCH_MIB of include/flx.h

PREC = {"fp32": 0, "bf16": 1, "fp16": 2}
REFINE = {"classic": 0, "gmres": 1, "auto": 2}
PIVOT = {"tournament": 0, "rbt": 1}


class MixedOpts(C.Structure):
    _fields_ = [("prec", C.c_int), ("refine", C.c_int), ("itermax", C.c_int),
                ("gmres_m", C.c_int), ("tol", C.c_double), ("gmres_tol", C.c_double),
                ("fallback", C.c_int), ("nb", C.c_int), ("graph", C.c_int), ("pivot", C.c_int)]


class BatchExtra(C.Structure):
    _fields_ = [("selector", _vp), ("features", _vp), ("ticks", _vp), ("choice", _vp),
                ("cfg", _vp), ("inner", _vp), ("used", _vp)]


MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 5, 64, 8, 5064
# solve plan modes (FLX_BATCH_*) and the ladders of a class (ladder + 4 x tile)
BATCH_FP64, BATCH_FP32_IR, BATCH_LEARNED, BATCH_BF16_GMRES, BATCH_FP16_GMRES, BATCH_RULE, \
    BATCH_TABLE = range(7)
LADDERS = ("fp64", "fp32_ir", "bf16_gmres", "fp16_gmres")


class MixedStats(C.Structure):
    _fields_ = [("iter", C.c_int), ("inner_iters", C.c_int), ("fallback", C.c_int),
                ("info", C.c_int), ("backward_error", C.c_double), ("anorm", C.c_double),
                ("kappa_est", C.c_double),
                ("factor_ms", C.c_double), ("refine_ms", C.c_double),
                ("fallback_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class FlxError(RuntimeError):
    pass


def lib():
    """Load libflx.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FlxError(f"{LIB_PATH} is missing: run `python -m flens_amd.build` first")
    L = C.CDLL(LIB_PATH)
    L.flx_create.argtypes = [C.POINTER(_vp)]
    L.flx_destroy.argtypes = [_vp]
    L.flx_set_stream.argtypes = [_vp, _vp]
    L.flx_sync.argtypes = [_vp]
    L.flx_status_string.argtypes = [_i]
    L.flx_status_string.restype = C.c_char_p
    L.flx_set_lu_blocking.argtypes = [_vp, _i, _i]
    _c, _cp = C.c_char, C.c_char_p
    L.flx_dlange.argtypes = [_vp, _c, _i, _i, _vp, _i, _dp]
    L.flx_dgeequ.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp, _dp, _dp, _dp, _ip]
    L.flx_dlaqge.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp, _d, _d, _d, _cp]
    L.flx_dgecon.argtypes = [_vp, _c, _i, _vp, _i, _d, _dp]
    L.flx_dgerfs.argtypes = [_vp, _i, _i, _i, _vp, _i, _vp, _i, _vp, _vp, _i, _vp, _i, _dp, _dp]
    L.flx_dgesvx.argtypes = [_vp, _c, _i, _i, _i, _vp, _i, _vp, _i, _vp, _cp, _vp, _vp, _vp, _i,
                             _vp, _i, _dp, _dp, _dp, _dp, _ip]
    L.flx_dgetri.argtypes = [_vp, _i, _vp, _i, _vp, _ip]
    L.flx_save_gematrix.argtypes = [_vp, _cp, _i, _i, _vp, _i, _i, _i]
    L.flx_load_gematrix_header.argtypes = [_cp, _ip, _ip, _ip, _ip]
    L.flx_load_gematrix.argtypes = [_vp, _cp, _vp, _i]
    L.flx_save_densevector.argtypes = [_vp, _cp, _i, _vp, _i, _i]
    L.flx_load_densevector.argtypes = [_vp, _cp, _vp, _i, _i, _ip, _ip]
    L.dgetri_.argtypes = [_ip, _dp, _ip, _ip, _dp, _ip, _ip]
    L.dlange_.argtypes = [_cp, _ip, _ip, _dp, _ip, _dp]
    L.dlange_.restype = _d
    L.dgeequ_.argtypes = [_ip, _ip, _dp, _ip, _dp, _dp, _dp, _dp, _dp, _ip]
    L.dgecon_.argtypes = [_cp, _ip, _dp, _ip, _dp, _dp, _dp, _ip, _ip]
    L.dgerfs_.argtypes = [_cp, _ip, _ip, _dp, _ip, _dp, _ip, _ip, _dp, _ip, _dp, _ip, _dp, _dp,
                          _dp, _ip, _ip]
    L.dgesvx_.argtypes = [_cp, _cp, _ip, _ip, _dp, _ip, _dp, _ip, _ip, _cp, _dp, _dp, _dp, _ip,
                          _dp, _ip, _dp, _dp, _dp, _dp, _ip, _ip]
    L.flx_set_lu_lookahead.argtypes = [_vp, _i]
    L.flx_test_raise_sync_flag.argtypes = [_vp]
    L.flx_dgetrf_noalloc.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp]
    L.flx_noalloc_bytes.argtypes = [_i, _i]
    L.flx_noalloc_bytes.restype = C.c_size_t
    L.flx_probe_mfma.argtypes = [_vp, _i, _i, _dp]
    L.flx_probe_copy.argtypes = [_vp, _vp, _vp, C.c_size_t, _dp]
    L.flx_batch_bin_tiles.argtypes = [_vp, _ip, _i]
    L.flx_set_lu_side_groups.argtypes = [_vp, _i, _i]
    L.flx_dgemm.argtypes = [_vp, _i, _i, _i, _i, _i, _d, _vp, _i, _vp, _i, _d, _vp, _i]
    L.flx_dgetrf.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp]
    L.flx_dgetrs.argtypes = [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _i]
    L.flx_dgesv.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _ip]
    L.flx_dlaswp.argtypes = [_vp, _i, _vp, _i, _i, _i, _vp, _i]
    L.flx_fill_uniform.argtypes = [_vp, _vp, _i, _i, _i, _u64, _u64]
    L.flx_profile_start.argtypes = [_vp]
    L.flx_profile_stop.argtypes = [_vp, _dp]
    L.cblas_dgemm.argtypes = [_i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]
    L.cblas_dgemm.restype = None
    L.cblas_dtrsm.argtypes = [_i, _i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]
    L.cblas_dtrsm.restype = None
    L.dgetrf_.argtypes = [_ip, _ip, _dp, _ip, _ip, _ip]
    L.dgetrf_.restype = None
    L.dgetrs_.argtypes = [C.c_char_p, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]
    L.dgetrs_.restype = None
    L.dgesv_.argtypes = [_ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]
    L.dgesv_.restype = None
    L.dlaswp_.argtypes = [_ip, _dp, _ip, _ip, _ip, _ip, _ip]
    L.dlaswp_.restype = None
    L.flx_dgesv_mixed.argtypes = [_vp, _i, _i, _vp, _i, _vp, _i, _vp, _i,
                                  C.POINTER(MixedOpts), C.POINTER(MixedStats)]
    L.flx_dsgesv.argtypes = [_vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp, _i, _ip, _ip]
    L.dsgesv_.argtypes = [_ip, _ip, _dp, _ip, _ip, _dp, _ip, _dp, _ip, _dp,
                          C.POINTER(C.c_float), _ip, _ip]
    L.dsgesv_.restype = None
    L.flx_batch_gesv_create.argtypes = [_vp, C.POINTER(_vp), _i, _ip, _ip, _vp, _vp, _vp, _vp,
                                        _vp, _i, _i, _i, C.POINTER(BatchExtra)]
    L.flx_mlp_forward.argtypes = [_vp, _vp, _vp, _i, _vp]
    L.flx_mlp_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, _vp, _vp]
    L.flx_mlp_adam.argtypes = [_vp, _vp, _vp, _vp, _vp, _i, C.c_float, C.c_float]
    L.flx_batch_gemm_create.argtypes = [_vp, C.POINTER(_vp), _i, _ip, _ip, _ip, _dp, _dp, _vp,
                                        _vp, _vp, _i]
    L.flx_batch_execute.argtypes = [_vp]
    L.flx_batch_config.argtypes = [_vp, _ip, _dp]
    L.flx_batch_destroy.argtypes = [_vp]
    _lib = L
    return L


def _check(status: int, what: str):
    if status != 0:
        msg = lib().flx_status_string(status).decode()
        raise FlxError(f"{what}: {msg} (status {status})")


class Handle:
    """flx handle bound to a HIP stream (torch's current stream by default)."""

    def __init__(self, stream=None):
        self._h = _vp()
        _check(lib().flx_create(C.byref(self._h)), "flx_create")
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        ptr = getattr(stream, "cuda_stream", stream)
        _check(lib().flx_set_stream(self._h, _vp(ptr)), "flx_set_stream")

    def set_lu_blocking(self, nb: int, ib: int = 32):
        _check(lib().flx_set_lu_blocking(self._h, nb, ib), "flx_set_lu_blocking")

    def set_lu_lookahead(self, depth: int):
        """0: each trailing update finishes before the next panel; 1: the
        next block's columns are updated first and the rest of the update
        runs on a second stream beside the next block's panels."""
        _check(lib().flx_set_lu_lookahead(self._h, depth), "flx_set_lu_lookahead")

    def set_lu_side_groups(self, groups: int, mixed: bool = False):
        """Workgroups the lookahead's second stream may use in the FP64 LU
        (or, mixed=True, in the mixed-precision factor); 0: the default."""
        _check(lib().flx_set_lu_side_groups(self._h, int(bool(mixed)), groups),
               "flx_set_lu_side_groups")

    def sync(self):
        _check(lib().flx_sync(self._h), "flx_sync")

    PROF_CLASSES = ["gemm", "panel", "rowperm", "trsm_rm", "transpose", "trs", "other"]

    def profile_start(self, classes=None):
        """Start the launch profiler; classes limits it to some kernel
        classes (names of PROF_CLASSES)."""
        _check(lib().flx_profile_start(self._h), "flx_profile_start")
        if classes is not None:
            mask = 0
            for c in classes:
                mask |= 1 << self.PROF_CLASSES.index(c)
            lib().flx_profile_classes.argtypes = [_vp, C.c_uint]
            _check(lib().flx_profile_classes(self._h, mask), "flx_profile_classes")

    def profile_stop(self) -> dict:
        out = (C.c_double * (4 * len(self.PROF_CLASSES)))()
        _check(lib().flx_profile_stop(self._h, out), "flx_profile_stop")
        res = {}
        for i, name in enumerate(self.PROF_CLASSES):
            cnt, ms, fl, by = out[4 * i: 4 * i + 4]
            res[name] = dict(launches=int(cnt), ms=ms, flops=fl, bytes=by)
        return res

    @property
    def ptr(self):
        return self._h

    def close(self):
        if self._h:
            lib().flx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class DeviceMatrix:
    """Column-major m x n view (leading dimension ld) of device memory.

    ``storage`` keeps the owning torch tensor alive."""
    ptr: int
    m: int
    n: int
    ld: int
    storage: object = None

    @staticmethod
    def empty(m: int, n: int, dtype="float64", device="cuda"):
        import torch
        t = torch.empty((max(n, 1), max(m, 1)), dtype=getattr(torch, dtype), device=device)
        return DeviceMatrix(t.data_ptr(), m, n, max(m, 1), t)

    @staticmethod
    def from_numpy(a: np.ndarray, device="cuda"):
        import torch
        a = np.asfortranarray(a)
        m, n = a.shape
        t = torch.from_numpy(np.ascontiguousarray(a.T)).to(device)
        return DeviceMatrix(t.data_ptr(), m, n, max(m, 1), t)

    def numpy(self) -> np.ndarray:
        t = self.storage
        assert t is not None and self.ld == self.m
        return np.asfortranarray(t.cpu().numpy().T[: self.m, : self.n])


def _int_buffer(n: int, device="cuda"):
    import torch
    return torch.zeros(max(n, 1), dtype=torch.int32, device=device)


# ---------------------------------------------------------------------------
# blas / lapack mirrors
# ---------------------------------------------------------------------------
class blas:
    NoTrans, Trans = 0, 1

    @staticmethod
    def mm(transA, transB, alpha, A, B, beta, C, handle: Handle | None = None):
        """C <- beta*C + alpha*op(A)*op(B).

        numpy (Fortran order) operands go through cblas_dgemm (synchronous,
        BLAS beta=0 convention); DeviceMatrix operands through flx_dgemm."""
        ta, tb = int(bool(transA)), int(bool(transB))
        if isinstance(C, np.ndarray):
            for x in (A, B, C):
                assert x.dtype == np.float64 and x.flags.f_contiguous
            m, n = C.shape
            k = A.shape[0] if ta else A.shape[1]
            lib().cblas_dgemm(102, 112 if ta else 111, 112 if tb else 111, m, n, k,
                              float(alpha), A.ctypes.data_as(_dp), max(1, A.shape[0]),
                              B.ctypes.data_as(_dp), max(1, B.shape[0]), float(beta),
                              C.ctypes.data_as(_dp), max(1, m))
            return C
        h = handle or default_handle()
        k = A.m if ta else A.n
        _check(lib().flx_dgemm(h.ptr, ta, tb, C.m, C.n, k, float(alpha), _vp(A.ptr), A.ld,
                               _vp(B.ptr), B.ld, float(beta), _vp(C.ptr), C.ld), "flx_dgemm")
        return C


def _iref(v: int):
    return C.byref(C.c_int(int(v)))


def _vec(t, n):
    """Device float64 vector of length n (a new one when t is None)."""
    import torch
    if t is None:
        t = torch.zeros(max(1, n), dtype=torch.float64, device="cuda")
    return t


class io:
    """flens::save / flens::load binary files (IndexType int, first index 1
    by default).  Device DeviceMatrix / torch vectors go through libflx
    (row order produced on the device); numpy arrays are read and written
    on the host in the same format."""

    @staticmethod
    def save(path: str, x, first: int = 1, handle: Handle | None = None) -> None:
        if isinstance(x, np.ndarray):
            with open(path, "wb") as f:
                if x.ndim == 2:
                    np.array([x.shape[0], x.shape[1], first, first], dtype=np.int32).tofile(f)
                    np.ascontiguousarray(x, dtype=np.float64).tofile(f)   # row by row
                else:
                    np.array([x.shape[0], first], dtype=np.int32).tofile(f)
                    x.tofile(f)
            return
        h = handle or default_handle()
        if isinstance(x, DeviceMatrix):
            _check(lib().flx_save_gematrix(h.ptr, path.encode(), x.m, x.n, _vp(x.ptr), x.ld, first,
                                           first), "flx_save_gematrix")
            return
        import torch
        elem = 4 if x.dtype == torch.int32 else 8
        _check(lib().flx_save_densevector(h.ptr, path.encode(), x.numel(), _vp(x.data_ptr()), elem,
                                          first), "flx_save_densevector")

    @staticmethod
    def load_numpy(path: str) -> np.ndarray:
        """Host read of a matrix file: (m, n) float64, Fortran order."""
        with open(path, "rb") as f:
            hdr = np.fromfile(f, dtype=np.int32, count=4)
            m, n = int(hdr[0]), int(hdr[1])
            data = np.fromfile(f, dtype=np.float64, count=m * n)
        return np.asfortranarray(data.reshape(m, n))

    @staticmethod
    def load_vector_numpy(path: str, dtype=np.float64) -> np.ndarray:
        """Host read of a vector file (float64, or int32 for pivots)."""
        with open(path, "rb") as f:
            hdr = np.fromfile(f, dtype=np.int32, count=2)
            return np.fromfile(f, dtype=dtype, count=int(hdr[0]))

    @staticmethod
    def load(path: str, handle: Handle | None = None) -> "DeviceMatrix":
        """Device read of a matrix file into a new DeviceMatrix."""
        h = handle or default_handle()
        m, n = C.c_int(0), C.c_int(0)
        _check(lib().flx_load_gematrix_header(path.encode(), C.byref(m), C.byref(n), None, None),
               "flx_load_gematrix_header")
        A = DeviceMatrix.empty(m.value, n.value)
        _check(lib().flx_load_gematrix(h.ptr, path.encode(), _vp(A.ptr), A.ld), "flx_load_gematrix")
        return A


class lapack:
    # ---- norms, equilibration, condition estimate, refinement -----------
    # Host numpy arrays go through the Fortran seams (dlange_ ...), device
    # DeviceMatrix arguments through the C ABI (flx_dlange ...).
    @staticmethod
    def lan(norm: str, A, handle: Handle | None = None) -> float:
        """flens::lapack::lan (GeMatrix): norm 'M', 'O', 'I' or 'F'."""
        if isinstance(A, np.ndarray):
            m, n = A.shape
            return lib().dlange_(norm.encode(), _iref(m), _iref(n), A.ctypes.data_as(_dp),
                                 _iref(max(1, m)), None)
        h = handle or default_handle()
        v = C.c_double(0.0)
        _check(lib().flx_dlange(h.ptr, norm.encode(), A.m, A.n, _vp(A.ptr), A.ld, C.byref(v)),
               "flx_dlange")
        return v.value

    @staticmethod
    def equ(A, handle: Handle | None = None):
        """flens::lapack::equ -> (r, c, rowcnd, colcnd, amax, info)."""
        rc, cc, am = C.c_double(np.nan), C.c_double(np.nan), C.c_double(np.nan)
        info = C.c_int(0)
        if isinstance(A, np.ndarray):
            m, n = A.shape
            r, c = np.zeros(max(1, m)), np.zeros(max(1, n))
            lib().dgeequ_(_iref(m), _iref(n), A.ctypes.data_as(_dp), _iref(max(1, m)),
                          r.ctypes.data_as(_dp), c.ctypes.data_as(_dp), C.byref(rc),
                          C.byref(cc), C.byref(am), C.byref(info))
            return r[:m], c[:n], rc.value, cc.value, am.value, info.value
        h = handle or default_handle()
        r, c = _vec(None, A.m), _vec(None, A.n)
        _check(lib().flx_dgeequ(h.ptr, A.m, A.n, _vp(A.ptr), A.ld, _vp(r.data_ptr()),
                                _vp(c.data_ptr()), C.byref(rc), C.byref(cc), C.byref(am),
                                C.byref(info)), "flx_dgeequ")
        return r[: A.m], c[: A.n], rc.value, cc.value, am.value, info.value

    @staticmethod
    def laq(A, r, c, rowcnd, colcnd, amax, handle: Handle | None = None) -> str:
        """flens::lapack::laq on device A; returns 'N', 'R', 'C' or 'B'."""
        h = handle or default_handle()
        eq = C.create_string_buffer(1)
        _check(lib().flx_dlaqge(h.ptr, A.m, A.n, _vp(A.ptr), A.ld, _vp(r.data_ptr()),
                                _vp(c.data_ptr()), float(rowcnd), float(colcnd), float(amax),
                                eq), "flx_dlaqge")
        return eq.raw.decode()

    @staticmethod
    def con(norm: str, LU, anorm: float, handle: Handle | None = None) -> float:
        """flens::lapack::con on an LU factor: reciprocal condition number."""
        rc = C.c_double(0.0)
        if isinstance(LU, np.ndarray):
            n = LU.shape[0]
            info = C.c_int(0)
            lib().dgecon_(norm.encode(), _iref(n), LU.ctypes.data_as(_dp), _iref(max(1, n)),
                          C.byref(C.c_double(anorm)), C.byref(rc), None, None, C.byref(info))
            if info.value:
                raise FlxError(f"dgecon_ info={info.value}")
            return rc.value
        h = handle or default_handle()
        _check(lib().flx_dgecon(h.ptr, norm.encode(), LU.n, _vp(LU.ptr), LU.ld, float(anorm),
                                C.byref(rc)), "flx_dgecon")
        return rc.value

    @staticmethod
    def rfs(trans, A, AF, piv, B, X, handle: Handle | None = None):
        """flens::lapack::rfs: refines X in place; returns (ferr, berr)."""
        if isinstance(A, np.ndarray):
            n, nrhs = B.shape
            piv = np.ascontiguousarray(piv, dtype=np.int32)
            ferr, berr = np.zeros(max(1, nrhs)), np.zeros(max(1, nrhs))
            info = C.c_int(0)
            ln = _iref(max(1, n))
            lib().dgerfs_(b"T" if trans else b"N", _iref(n), _iref(nrhs), A.ctypes.data_as(_dp),
                          ln, AF.ctypes.data_as(_dp), ln, piv.ctypes.data_as(_ip),
                          B.ctypes.data_as(_dp), ln, X.ctypes.data_as(_dp), ln,
                          ferr.ctypes.data_as(_dp), berr.ctypes.data_as(_dp), None, None,
                          C.byref(info))
            if info.value:
                raise FlxError(f"dgerfs_ info={info.value}")
            return ferr[:nrhs], berr[:nrhs]
        h = handle or default_handle()
        nrhs = B.n
        ferr, berr = np.zeros(max(1, nrhs)), np.zeros(max(1, nrhs))
        _check(lib().flx_dgerfs(h.ptr, int(bool(trans)), A.n, nrhs, _vp(A.ptr), A.ld,
                                _vp(AF.ptr), AF.ld, _vp(piv.data_ptr()), _vp(B.ptr), B.ld,
                                _vp(X.ptr), X.ld, ferr.ctypes.data_as(_dp),
                                berr.ctypes.data_as(_dp)), "flx_dgerfs")
        return ferr[:nrhs], berr[:nrhs]

    @staticmethod
    def svx(fact: str, trans, A, AF, piv, equed: str, r, c, B, X,
            handle: Handle | None = None):
        """flens::lapack::svx.  Device arguments are updated in place (host
        numpy ones through dgesvx_).  Returns dict(rcond, ferr, berr, rpvgrw,
        equed, info)."""
        eq = C.create_string_buffer(equed.encode(), 1)
        rc, g = C.c_double(0.0), C.c_double(0.0)
        info = C.c_int(0)
        if isinstance(A, np.ndarray):
            n, nrhs = B.shape
            ferr, berr = np.zeros(max(1, nrhs)), np.zeros(max(1, nrhs))
            work = np.zeros(max(1, 4 * n))
            ln = _iref(max(1, n))
            lib().dgesvx_(fact.encode(), b"T" if trans else b"N", _iref(n), _iref(nrhs),
                          A.ctypes.data_as(_dp), ln, AF.ctypes.data_as(_dp), ln,
                          piv.ctypes.data_as(_ip), eq, r.ctypes.data_as(_dp),
                          c.ctypes.data_as(_dp), B.ctypes.data_as(_dp), ln,
                          X.ctypes.data_as(_dp), ln, C.byref(rc), ferr.ctypes.data_as(_dp),
                          berr.ctypes.data_as(_dp), work.ctypes.data_as(_dp), None,
                          C.byref(info))
            return dict(rcond=rc.value, ferr=ferr[:nrhs], berr=berr[:nrhs], rpvgrw=work[0],
                        equed=eq.raw.decode(), info=info.value)
        h = handle or default_handle()
        nrhs = B.n
        ferr, berr = np.zeros(max(1, nrhs)), np.zeros(max(1, nrhs))
        _check(lib().flx_dgesvx(h.ptr, fact.encode(), int(bool(trans)), A.n, nrhs, _vp(A.ptr),
                                A.ld, _vp(AF.ptr), AF.ld, _vp(piv.data_ptr()), eq,
                                _vp(r.data_ptr()), _vp(c.data_ptr()), _vp(B.ptr), B.ld,
                                _vp(X.ptr), X.ld, C.byref(rc), ferr.ctypes.data_as(_dp),
                                berr.ctypes.data_as(_dp), C.byref(g), C.byref(info)),
               "flx_dgesvx")
        return dict(rcond=rc.value, ferr=ferr[:nrhs], berr=berr[:nrhs], rpvgrw=g.value,
                    equed=eq.raw.decode(), info=info.value)

    @staticmethod
    def tri(A, piv, handle: Handle | None = None) -> int:
        """flens::lapack::tri (GeMatrix): inv(A) in place from its LU factor
        and pivots; returns info (i > 0: U(i,i) = 0, A untouched)."""
        info = C.c_int(0)
        if isinstance(A, np.ndarray):
            n = A.shape[0]
            piv = np.ascontiguousarray(piv, dtype=np.int32)
            work = np.zeros(max(1, n))
            lib().dgetri_(_iref(n), A.ctypes.data_as(_dp), _iref(max(1, n)),
                          piv.ctypes.data_as(_ip), work.ctypes.data_as(_dp), _iref(max(1, n)),
                          C.byref(info))
            return info.value
        h = handle or default_handle()
        _check(lib().flx_dgetri(h.ptr, A.n, _vp(A.ptr), A.ld, _vp(piv.data_ptr()),
                                C.byref(info)), "flx_dgetri")
        return info.value

    @staticmethod
    def trf(A, piv=None, handle: Handle | None = None):
        """LU with partial pivoting in place.  Returns (piv, info): numpy
        int32 1-based pivots for host A; a device int32 tensor for device A
        (info is then a device tensor of one int)."""
        if isinstance(A, np.ndarray):
            assert A.dtype == np.float64 and A.flags.f_contiguous
            m, n = A.shape
            piv = np.zeros(max(1, min(m, n)), dtype=np.int32)
            info = C.c_int(0)
            lib().dgetrf_(C.byref(C.c_int(m)), C.byref(C.c_int(n)), A.ctypes.data_as(_dp),
                          C.byref(C.c_int(max(1, m))), piv.ctypes.data_as(_ip), C.byref(info))
            return piv[: min(m, n)], info.value
        h = handle or default_handle()
        piv = _int_buffer(min(A.m, A.n)) if piv is None else piv
        info = _int_buffer(1)
        _check(lib().flx_dgetrf(h.ptr, A.m, A.n, _vp(A.ptr), A.ld, _vp(piv.data_ptr()),
                                _vp(info.data_ptr())), "flx_dgetrf")
        return piv, info

    @staticmethod
    def trs(trans, A, piv, B, handle: Handle | None = None):
        if isinstance(B, np.ndarray):
            n = A.shape[0]
            nrhs = B.shape[1]
            piv = np.ascontiguousarray(piv, dtype=np.int32)
            info = C.c_int(0)
            lib().dgetrs_(b"T" if trans else b"N", C.byref(C.c_int(n)), C.byref(C.c_int(nrhs)),
                          A.ctypes.data_as(_dp), C.byref(C.c_int(max(1, n))),
                          piv.ctypes.data_as(_ip), B.ctypes.data_as(_dp),
                          C.byref(C.c_int(max(1, n))), C.byref(info))
            if info.value != 0:
                raise FlxError(f"dgetrs_ info={info.value}")
            return B
        h = handle or default_handle()
        _check(lib().flx_dgetrs(h.ptr, int(bool(trans)), A.n, B.n, _vp(A.ptr), A.ld,
                                _vp(piv.data_ptr()), _vp(B.ptr), B.ld), "flx_dgetrs")
        return B

    @staticmethod
    def sv(A, B, handle: Handle | None = None):
        """Solve A X = B; returns (piv, info).  B is overwritten with X."""
        if isinstance(A, np.ndarray):
            n = A.shape[0]
            piv = np.zeros(max(1, n), dtype=np.int32)
            info = C.c_int(0)
            lib().dgesv_(C.byref(C.c_int(n)), C.byref(C.c_int(B.shape[1])),
                         A.ctypes.data_as(_dp), C.byref(C.c_int(max(1, n))),
                         piv.ctypes.data_as(_ip), B.ctypes.data_as(_dp),
                         C.byref(C.c_int(max(1, n))), C.byref(info))
            return piv[:n], info.value
        h = handle or default_handle()
        piv = _int_buffer(A.n)
        info = C.c_int(0)
        _check(lib().flx_dgesv(h.ptr, A.n, B.n, _vp(A.ptr), A.ld, _vp(piv.data_ptr()),
                               _vp(B.ptr), B.ld, C.byref(info)), "flx_dgesv")
        return piv, info.value

    @staticmethod
    def dsgesv(A, B, X, handle: Handle | None = None):
        """LAPACK dsgesv: X = A^{-1} B by an FP32 factor and FP64 classic
        refinement, FP64 LU as the fallback (A is overwritten only then).
        Returns (piv, iter, info)."""
        if isinstance(A, np.ndarray):
            n, nrhs = A.shape[0], B.shape[1]
            piv = np.zeros(max(1, n), dtype=np.int32)
            it, info = C.c_int(0), C.c_int(0)
            lib().dsgesv_(C.byref(C.c_int(n)), C.byref(C.c_int(nrhs)), A.ctypes.data_as(_dp),
                          C.byref(C.c_int(max(1, n))), piv.ctypes.data_as(_ip),
                          B.ctypes.data_as(_dp), C.byref(C.c_int(max(1, n))),
                          X.ctypes.data_as(_dp), C.byref(C.c_int(max(1, n))), None, None,
                          C.byref(it), C.byref(info))
            return piv[:n], it.value, info.value
        h = handle or default_handle()
        piv = _int_buffer(A.n)
        it, info = C.c_int(0), C.c_int(0)
        _check(lib().flx_dsgesv(h.ptr, A.n, B.n, _vp(A.ptr), A.ld, _vp(piv.data_ptr()),
                                _vp(B.ptr), B.ld, _vp(X.ptr), X.ld, C.byref(it), C.byref(info)),
               "flx_dsgesv")
        return piv, it.value, info.value

    @staticmethod
    def sv_mixed(A, B, X, prec="fp32", refine="classic", itermax=30, gmres_m=40, tol=0.0,
                 gmres_tol=0.0, fallback=True, nb=256, graph=True, pivot="tournament",
                 handle: Handle | None = None) -> dict:
        """Mixed-precision solve of A X = B on device matrices; returns the
        solver's statistics (iterations, fallback, backward error, times).
        pivot: "tournament" (CALU panels) or "rbt" (random butterflies, no
        pivoting)."""
        h = handle or default_handle()
        o = MixedOpts(PREC[prec], REFINE[refine], itermax, gmres_m, tol, gmres_tol,
                      int(bool(fallback)), nb, int(bool(graph)), PIVOT[pivot])
        st = MixedStats()
        _check(lib().flx_dgesv_mixed(h.ptr, A.n, B.n, _vp(A.ptr), A.ld, _vp(B.ptr), B.ld,
                                     _vp(X.ptr), X.ld, C.byref(o), C.byref(st)),
               "flx_dgesv_mixed")
        return st.as_dict()

    @staticmethod
    def laswp(A: np.ndarray, k1: int, k2: int, piv, incx: int = 1):
        piv = np.ascontiguousarray(piv, dtype=np.int32)
        m, n = A.shape
        lib().dlaswp_(C.byref(C.c_int(n)), A.ctypes.data_as(_dp), C.byref(C.c_int(max(1, m))),
                      C.byref(C.c_int(k1)), C.byref(C.c_int(k2)), piv.ctypes.data_as(_ip),
                      C.byref(C.c_int(incx)))
        return A


class BatchPlan:
    """A batch of solves or GEMMs captured as one persistent launch
    (flx_batch_*).  Keeps the device buffers it was built on alive."""

    def __init__(self, ptr, keep, handle):
        self._p, self._keep, self._h = ptr, keep, handle

    def execute(self):
        _check(lib().flx_batch_execute(self._p), "flx_batch_execute")

    def config(self):
        cfg = (C.c_int * 2)()
        ms = C.c_double(0.0)
        _check(lib().flx_batch_config(self._p, cfg, C.byref(ms)), "flx_batch_config")
        return {"tile_cfg": cfg[0], "grid": cfg[1], "tuned_ms": ms.value}

    def bin_tiles(self):
        """Solve plans: the autotuned tile of each size bin (n in
        [8 2^b, 16 2^b), b = 0..5)."""
        t = (C.c_int * 6)()
        nb = lib().flx_batch_bin_tiles(self._p, t, 6)
        if nb < 0:
            _check(nb, "flx_batch_bin_tiles")
        return list(t)[:nb]

    def close(self):
        if self._p:
            lib().flx_batch_destroy(self._p)
            self._p = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class mlp:
    """The selector network's device kernels (flx_mlp_*) on torch tensors."""

    @staticmethod
    def forward(params, X, handle: Handle | None = None):
        import torch
        h = handle or default_handle()
        out = torch.empty((X.shape[0], MLP_OUT), dtype=torch.float32, device=X.device)
        _check(lib().flx_mlp_forward(h.ptr, _vp(params.data_ptr()), _vp(X.data_ptr()),
                                     X.shape[0], _vp(out.data_ptr())), "flx_mlp_forward")
        return out

    @staticmethod
    def grad(params, X, label, weight=None, handle: Handle | None = None):
        import torch
        h = handle or default_handle()
        g = torch.empty(MLP_PARAMS, dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        _check(lib().flx_mlp_grad(h.ptr, _vp(params.data_ptr()), _vp(X.data_ptr()),
                                  _vp(label.data_ptr()),
                                  _vp(weight.data_ptr()) if weight is not None else None,
                                  X.shape[0], _vp(g.data_ptr()), _vp(loss.data_ptr())),
               "flx_mlp_grad")
        return g, loss

    @staticmethod
    def adam(params, grad, m, v, step, lr=1e-3, weight_decay=0.0, handle: Handle | None = None):
        h = handle or default_handle()
        _check(lib().flx_mlp_adam(h.ptr, _vp(params.data_ptr()), _vp(grad.data_ptr()),
                                  _vp(m.data_ptr()), _vp(v.data_ptr()), step, lr, weight_decay),
               "flx_mlp_adam")


class batched:
    """Packed variable-size batches (see include/flx.h for the layout)."""

    @staticmethod
    def gesv_plan(n, nrhs, A, B, ipiv, info, iters=None, mode=0, itmax=5, tune=True,
                  selector=None, features=None, ticks=None, choice=None, cfg=None,
                  inner=None, used=None, handle: Handle | None = None) -> BatchPlan:
        """A, B: device float64 tensors (packed); ipiv, info, iters, choice,
        cfg, inner, used: int32; selector, features: float32; ticks: int64
        (all on the device).  mode: BATCH_*."""
        h = handle or default_handle()
        n = np.ascontiguousarray(n, dtype=np.int32)
        nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        ptr = (lambda t: _vp(t.data_ptr()) if t is not None else None)
        ex = BatchExtra(ptr(selector), ptr(features), ptr(ticks), ptr(choice), ptr(cfg),
                        ptr(inner), ptr(used))
        p = _vp()
        _check(lib().flx_batch_gesv_create(
            h.ptr, C.byref(p), len(n), n.ctypes.data_as(_ip), nrhs.ctypes.data_as(_ip),
            _vp(A.data_ptr()), _vp(B.data_ptr()), _vp(ipiv.data_ptr()), _vp(info.data_ptr()),
            ptr(iters), mode, itmax, int(bool(tune)), C.byref(ex)), "flx_batch_gesv_create")
        return BatchPlan(p, (A, B, ipiv, info, iters, selector, features, ticks, choice, cfg,
                             inner, used), h)

    @staticmethod
    def gemm_plan(m, n, k, alpha, beta, A, B, Cm, tune=True,
                  handle: Handle | None = None) -> BatchPlan:
        h = handle or default_handle()
        m = np.ascontiguousarray(m, dtype=np.int32)
        n = np.ascontiguousarray(n, dtype=np.int32)
        k = np.ascontiguousarray(k, dtype=np.int32)
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        beta = np.ascontiguousarray(beta, dtype=np.float64)
        p = _vp()
        _check(lib().flx_batch_gemm_create(
            h.ptr, C.byref(p), len(m), m.ctypes.data_as(_ip), n.ctypes.data_as(_ip),
            k.ctypes.data_as(_ip), alpha.ctypes.data_as(_dp), beta.ctypes.data_as(_dp),
            _vp(A.data_ptr()), _vp(B.data_ptr()), _vp(Cm.data_ptr()), int(bool(tune))),
            "flx_batch_gemm_create")
        return BatchPlan(p, (A, B, Cm), h)


_default = None


def default_handle() -> Handle:
    """Process-wide handle on torch's current stream."""
    global _default
    if _default is None:
        import torch
        if not torch.cuda.is_available():
            raise FlxError("no GPU: flens_amd has no CPU path")
        _default = Handle(torch.cuda.current_stream())
    return _default


def probe_peaks(handle: Handle | None = None, copy_bytes: int = 1 << 30) -> dict:
    """Achievable peaks on this device: dense MFMA TFLOP/s for FP64, FP32,
    BF16 and FP16 operands (flx_probe_mfma) and streaming-copy GB/s
    (flx_probe_copy over two copy_bytes buffers)."""
    import torch
    h = handle or default_handle()
    L = lib()
    out = {}
    for name, kind, iters in (("fp64", 0, 20000), ("fp32", 1, 20000),
                              ("bf16", 2, 80000), ("fp16", 3, 80000)):
        v = C.c_double(0.0)
        _check(L.flx_probe_mfma(h.ptr, kind, iters, C.byref(v)), "flx_probe_mfma")
        out[f"{name}_mfma_tflops"] = v.value
    src = torch.empty(copy_bytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    v = C.c_double(0.0)
    _check(L.flx_probe_copy(h.ptr, _vp(src.data_ptr()), _vp(dst.data_ptr()), copy_bytes,
                            C.byref(v)), "flx_probe_copy")
    out["hbm_copy_gbps"] = v.value
    del src, dst
    return out
