"""The HIP library loads on a CPU-only host and exports every entry point the
public header declares (no compute calls here)."""
import ctypes
import os
import re

import calypso_amd
from calypso_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "calypso_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(calypso_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_symbols()


def test_host_only_entry_points():
    lib = calypso_amd.load()
    assert lib.calypso_legendre_mode_id(b"hip_fp64") == 1
    assert lib.calypso_legendre_mode_id(b"Hip_BF16x3") == 3
    assert lib.calypso_legendre_mode_id(b"symmetric_matmul") == -1
    assert calypso_amd.set_legendre_trans_mode_ctl("hip_fp32") == 2


def test_gauss_points_match_oracle():
    import numpy as np
    import oracle
    for n in (12, 48, 384):
        a = calypso_amd.gauss_points(n)
        b = oracle.gauss_points(n)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_fortran_shim_built():
    lib_dir = _lib.LIB_DIR
    assert os.path.exists(os.path.join(lib_dir, "libcalypso_amd_f.so"))
    assert os.path.exists(os.path.join(lib_dir, "legendre_trans_hip.mod"))
    assert os.path.exists(os.path.join(lib_dir, "leg_driver"))
