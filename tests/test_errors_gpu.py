"""GPU tests of the bounded-wait error path: a device-side wait that gives
up raises the handle's flag; the synchronous entry point that queued it
reports FLX_SYNC_TIMEOUT and re-arms the flag, so the next call is clean.

The flag is raised from the host (flx_test_raise_sync_flag queues a write of
it on the handle's stream, as a failed wait would) so that no kernel has to
time out.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
SYNC_TIMEOUT = 4


def test_synchronous_entry_points_report_their_own_failed_wait(gpu):
    L = gpu.lib()
    h = gpu.Handle()
    n = 300
    a0 = O.uniform(n, n, 5)
    A = gpu.DeviceMatrix.from_numpy(a0)
    piv, info = gpu.lapack.trf(A, handle=h)
    h.sync()
    rc = C.c_double(0.0)

    def con():
        return L.flx_dgecon(h.ptr, b"O", n, gpu._vp(A.ptr), A.ld, 10.0, C.byref(rc))

    assert con() == 0
    # con: the flag raised on its stream is its error, then re-armed
    assert L.flx_test_raise_sync_flag(h.ptr) == 0
    assert con() == SYNC_TIMEOUT
    assert con() == 0 and rc.value > 0.0
    # flx_sync reports a failed wait of the asynchronous calls queued before it
    assert L.flx_test_raise_sync_flag(h.ptr) == 0
    assert L.flx_sync(h.ptr) == SYNC_TIMEOUT
    assert L.flx_sync(h.ptr) == 0
    # flx_dgesv reads INFO only after checking the flag
    B = gpu.DeviceMatrix.from_numpy(O.uniform(n, 2, 6))
    A2 = gpu.DeviceMatrix.from_numpy(a0)
    ipiv = gpu._int_buffer(n)
    inf = C.c_int(-1)
    assert L.flx_test_raise_sync_flag(h.ptr) == 0
    st = L.flx_dgesv(h.ptr, n, 2, gpu._vp(A2.ptr), A2.ld, gpu._vp(ipiv.data_ptr()),
                     gpu._vp(B.ptr), B.ld, C.byref(inf))
    assert st == SYNC_TIMEOUT
    A2 = gpu.DeviceMatrix.from_numpy(a0)
    st = L.flx_dgesv(h.ptr, n, 2, gpu._vp(A2.ptr), A2.ld, gpu._vp(ipiv.data_ptr()),
                     gpu._vp(B.ptr), B.ld, C.byref(inf))
    assert st == 0 and inf.value == 0
    h.close()


def test_mixed_fallback_reports_a_failed_wait(gpu):
    """The FP64 fallback of the mixed solver (kappa = 1e12 sends the
    automatic ladder to FP64 at once) checks the flag after its LU."""
    from flens_amd import testmat as T
    L = gpu.lib()
    h = gpu.Handle()
    n = 200
    a = T.svd_matrix(n, 1e12, 3)
    b, _ = T.rhs_from_solution(a, 1, 3)
    A = gpu.DeviceMatrix.from_numpy(a)
    B = gpu.DeviceMatrix.from_numpy(b)
    X = gpu.DeviceMatrix.from_numpy(np.zeros_like(b))
    assert L.flx_test_raise_sync_flag(h.ptr) == 0
    with pytest.raises(gpu.FlxError, match="timed out"):
        gpu.lapack.sv_mixed(A, B, X, prec="fp32", refine="auto", fallback=True, handle=h)
    A = gpu.DeviceMatrix.from_numpy(a)
    st = gpu.lapack.sv_mixed(A, B, X, prec="fp32", refine="auto", fallback=True, handle=h)
    assert st["fallback"] == 1
    h.close()
