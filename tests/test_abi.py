"""CPU tests of the C-ABI boundary: libflx.so loads without a GPU and exports
every function include/flx.h declares (no compute calls here)."""
import os
import re
import subprocess

import flens_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "flx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return")))


def test_header_declares_expected_symbols():
    names = declared_functions()
    assert set(F.EXPORTS) == set(names), set(F.EXPORTS) ^ set(names)


def test_library_exports_every_declared_symbol():
    L = F.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in declared_functions():
        assert name in exported, name


def test_no_cpu_fallback_without_gpu():
    import torch
    if torch.cuda.is_available():
        return
    try:
        F.default_handle()
    except F.FlxError:
        return
    raise AssertionError("device path must refuse to run without a GPU")


def test_oracle_not_linked_into_product():
    out = subprocess.run(["nm", "-D", F.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "orc_" not in out and "ref_d" not in out
    src = open(os.path.join(ROOT, "flens_amd", "__init__.py")).read()
    assert "oracle" not in src.split('"""', 2)[2]


DROPIN = os.path.join(ROOT, "oracle", "_ref", "flens_dropin")


def test_flens_program_binds_to_libflx():
    """A FLENS program built with FLENS's external BLAS/LAPACK switches
    (tests/dropin/flens_dropin.cc, built by oracle/Makefile where the
    reference headers exist) resolves its GEMM, triangular-solve and LU calls
    to libflx."""
    import pytest
    if not os.path.exists(DROPIN):
        pytest.skip("drop-in program not built (needs the reference headers)")
    und = subprocess.run(["nm", "-u", DROPIN], capture_output=True, text=True, check=True).stdout
    needed = {line.split()[-1] for line in und.splitlines() if line.strip()}
    assert {"cblas_dgemm", "cblas_dtrsm", "dgetrf_", "dgetrs_", "dgesv_"} <= needed
    ldd = subprocess.run(["ldd", DROPIN], capture_output=True, text=True, check=True).stdout
    line = next(l for l in ldd.splitlines() if "libflx.so" in l)
    assert os.path.realpath(line.split("=>")[1].split("(")[0].strip()) == os.path.realpath(F.LIB_PATH)
