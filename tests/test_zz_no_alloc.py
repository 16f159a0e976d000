"""GPU test of flx_dgetrf_noalloc, the LU entry point that allocates
nothing (its scratch is a fixed static device array): on two calls in a
row, the second smaller, its factors, pivots and INFO are bit-identical to
flx_dgetrf's on the same input (both without lookahead), and they pass the
oracle's checks (pivots equal the reference's away from near-ties, dget01 <
30).  Shapes beyond the scratch are rejected without a launch."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
SCRATCH_LIMIT = 6


def test_noalloc_lu_matches_dgetrf(gpu):
    import torch
    L = gpu.lib()
    h = gpu.Handle(torch.cuda.current_stream())
    h.set_lu_lookahead(0)
    h.set_lu_blocking(128, 32)
    for (m, n, seed) in [(2100, 2100, 1), (900, 700, 2), (300, 500, 3)]:
        assert L.flx_noalloc_bytes(m, n) <= gpu.NOALLOC_SCRATCH_MIB << 20
        a0 = O.uniform(m, n, 900 + seed)
        A1 = gpu.DeviceMatrix.from_numpy(a0)
        piv1, info1 = gpu.lapack.trf(A1, handle=h)
        A2 = gpu.DeviceMatrix.from_numpy(a0)
        piv2 = gpu._int_buffer(min(m, n))
        info2 = gpu._int_buffer(1)
        st = L.flx_dgetrf_noalloc(h.ptr, m, n, gpu._vp(A2.ptr), A2.ld, gpu._vp(piv2.data_ptr()),
                                  gpu._vp(info2.data_ptr()))
        assert st == 0, st
        h.sync()
        lu1, lu2 = A1.numpy(), A2.numpy()
        p1, p2 = piv1.cpu().numpy(), piv2.cpu().numpy()
        assert int(info1.cpu()[0]) == int(info2.cpu()[0]) == 0, (m, n)
        assert np.array_equal(p1, p2), (m, n)
        assert np.array_equal(lu1, lu2), (m, n)
        rpiv, rinfo, gap = O.dgetrf_gaps(a0.copy(order="F"))
        d = np.nonzero(p2[: min(m, n)] != rpiv)[0]
        assert d.size == 0 or gap[d[0]] < 64 * max(m, n) * 2.0 ** -53, (m, n, d[:1])
        assert O.dget01(a0, lu2, p2[: min(m, n)]) < 30, (m, n)
    # beyond the scratch: refused, nothing launched, the arrays untouched
    big = 4096
    assert L.flx_noalloc_bytes(big, big) > gpu.NOALLOC_SCRATCH_MIB << 20
    A3 = gpu.DeviceMatrix.from_numpy(np.ones((4, 4), order="F"))
    st = L.flx_dgetrf_noalloc(h.ptr, big, big, gpu._vp(A3.ptr), big, gpu._vp(A3.ptr),
                              gpu._vp(A3.ptr))
    assert st == SCRATCH_LIMIT
    h.sync()
    assert np.array_equal(A3.numpy(), np.ones((4, 4)))
    h.close()
