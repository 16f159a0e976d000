"""The netlib BLAS-3 GEMM test (dblat3's DCHK1 grid, DMMCH ratio < 16, the
DCHKE error exits) and the LAPACK LU test (dchkge's dget01/02/03/04/06/07
ratios < 30 over DLATMS-style matrix types 1-11, sizes 0..50, 1/2/15
right-hand sides, both transposes), restated in C++ over libflx's Fortran
seams (tests/netlib/netlib_checks.cc, built by build()).  Runs the program
as a child process on the GPU and requires every group to pass."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "netlib", "_build",
                   "netlib_checks")


def test_netlib_suites_through_fortran_seams(gpu):
    assert os.path.exists(EXE), "tests/netlib/_build/netlib_checks missing: run build() first"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
