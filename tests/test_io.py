"""flens::save / flens::load file format (GeMatrix and DenseVector, IndexType
int): the package's host writer/reader against files written and read by
the compiled reference (flens/io/fullstorage/save.tcc:49-82,
load.tcc:52-80, flens/io/array/save.tcc:43-66), and -- on the GPU -- the
device save/load entry points against the same bytes."""
import ctypes as C
import os

import numpy as np
import pytest

import flens_amd as F
from oracle import oracle as O

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="compiled reference (oracle/_ref) absent")


def _ref():
    R = O.ref()
    R.ref_save_gematrix.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int]
    R.ref_load_gematrix.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double), C.c_int]
    R.ref_save_ivector.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    return R


def test_host_roundtrip(tmp_path):
    a = np.asfortranarray(np.random.default_rng(1).standard_normal((7, 5)))
    p = str(tmp_path / "a.bin")
    F.io.save(p, a)
    assert os.path.getsize(p) == 16 + 8 * 35
    hdr = np.fromfile(p, dtype=np.int32, count=4)
    assert list(hdr) == [7, 5, 1, 1]
    raw = np.fromfile(p, dtype=np.float64, offset=16)
    assert np.array_equal(raw, a.ravel(order="C"))          # row by row
    assert np.array_equal(F.io.load_numpy(p), a)
    piv = np.array([3, 3, 5, 4, 5], dtype=np.int32)
    q = str(tmp_path / "p.bin")
    F.io.save(q, piv)
    assert np.array_equal(F.io.load_vector_numpy(q, np.int32), piv)


@needs_ref
def test_files_match_reference(tmp_path):
    R = _ref()
    a = np.asfortranarray(np.random.default_rng(2).standard_normal((13, 9)))
    pr, pp = str(tmp_path / "ref.bin"), str(tmp_path / "ours.bin")
    assert R.ref_save_gematrix(pr.encode(), 13, 9, a.ctypes.data_as(C.POINTER(C.c_double)), 13) == 0
    F.io.save(pp, a)
    assert open(pr, "rb").read() == open(pp, "rb").read()
    out = np.zeros(13 * 9)
    m, n = C.c_int(0), C.c_int(0)
    assert R.ref_load_gematrix(pp.encode(), C.byref(m), C.byref(n),
                               out.ctypes.data_as(C.POINTER(C.c_double)), out.size) == 0
    assert (m.value, n.value) == (13, 9)
    assert np.array_equal(out.reshape(9, 13).T, a)
    piv = np.array([2, 3, 3, 4], dtype=np.int32)
    vr, vp = str(tmp_path / "vr.bin"), str(tmp_path / "vp.bin")
    assert R.ref_save_ivector(vr.encode(), 4, piv.ctypes.data_as(C.POINTER(C.c_int))) == 0
    F.io.save(vp, piv)
    assert open(vr, "rb").read() == open(vp, "rb").read()


@pytest.mark.gpu
def test_device_save_load(gpu, tmp_path):
    import torch
    a = np.asfortranarray(np.random.default_rng(3).standard_normal((300, 170)))
    A = gpu.DeviceMatrix.from_numpy(a)
    pd, ph = str(tmp_path / "dev.bin"), str(tmp_path / "host.bin")
    gpu.io.save(pd, A)
    gpu.io.save(ph, a)
    assert open(pd, "rb").read() == open(ph, "rb").read()
    B = gpu.io.load(ph)
    assert (B.m, B.n) == (300, 170) and np.array_equal(B.numpy(), a)
    piv = torch.arange(1, 301, dtype=torch.int32, device="cuda")
    pv = str(tmp_path / "piv.bin")
    gpu.io.save(pv, piv)
    assert np.array_equal(gpu.io.load_vector_numpy(pv, np.int32), np.arange(1, 301))
