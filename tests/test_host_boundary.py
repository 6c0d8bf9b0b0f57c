"""Host-buffer entry points shaped after the reference's Fortran modules
(MHD_FFT_selector, lubksb_357band_mul): argument checks that return an error
before any device work, so they run without a GPU."""
import ctypes
import os

import numpy as np

import calypso_amd
from calypso_amd import _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def last_error():
    return calypso_amd.load().calypso_last_error().decode()


def band_case(nb=3, n=9, mcomp=6):
    lu = np.ones((mcomp, n, 2 * nb - 1))
    piv = np.tile(np.arange(1, n + 1, dtype=np.int32), (mcomp, 1))
    x = np.arange(n * mcomp, dtype=np.float64).reshape(n, mcomp)
    return lu, piv, x


def test_lubksb_rejects_pivot_outside_matrix():
    lib = calypso_amd.load()
    lu, piv, x = band_case()
    piv[2, 4] = 10                        # row 10 of a 9-row system
    stack = np.array([0, 3, 6], dtype=np.int32)
    x0 = x.copy()
    rc = lib.calypso_lubksb_3band_mul(2, ptr(stack), 6, 9, ptr(lu), ptr(piv), ptr(x))
    assert rc != 0 and "i_pivot" in last_error()
    assert np.array_equal(x, x0)


def test_lubksb_rejects_bad_stack():
    lib = calypso_amd.load()
    lu, piv, x = band_case(nb=5)
    for stack in ([0, 4, 2], [0, 7], [-1, 3]):
        st = np.array(stack, dtype=np.int32)
        rc = lib.calypso_lubksb_5band_mul(len(stack) - 1, ptr(st), 6, 9, ptr(lu), ptr(piv), ptr(x))
        assert rc != 0 and "Msmp_stack" in last_error()


def test_lubksb_empty_range_leaves_x():
    lib = calypso_amd.load()
    lu, piv, x = band_case()
    x0 = x.copy()
    st = np.array([4, 4, 4], dtype=np.int32)
    assert lib.calypso_lubksb_3band_mul(2, ptr(st), 6, 9, ptr(lu), ptr(piv), ptr(x)) == 0
    assert np.array_equal(x, x0)


def test_fft_selector_needs_init():
    lib = calypso_amd.load()
    lib.calypso_finalize_MHD_FFT_select()
    wr = np.zeros(16)
    out = np.zeros(16)
    assert lib.calypso_back_MHD_FFT_sel_from_recv(1, 16, ptr(wr), ptr(out)) != 0
    assert "not initialised" in last_error()
    assert lib.calypso_fwd_MHD_FFT_sel_from_recv(1, 16, ptr(wr), ptr(out)) != 0


def test_fortran_modules_built():
    for f in ("mhd_fft_select_hip.mod", "lubksb_band_hip.mod", "cal_nonlinear_sph_mhd_hip.mod",
              "fft_lu_driver"):
        assert os.path.exists(os.path.join(_lib.LIB_DIR, f)), f


def test_products_rejects_bad_arguments():
    lib = calypso_amd.load()
    fld = np.zeros(13)
    assert lib.calypso_nonlinear_products(0, ptr(fld), ptr(fld), ptr(fld)) != 0
    assert "calypso_nonlinear_products" in last_error()


def test_no_fallback_without_the_library(tmp_path):
    """Every entry point fails loudly when the HIP library is absent: there is
    no CPU path to fall back on."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {root!r})\n"
        "import calypso_amd, calypso_amd._lib as L\n"
        f"L.LIB_PATH = {str(tmp_path / 'absent.so')!r}\n"
        "from calypso_amd import mhd\n"
        "fails = 0\n"
        "for make in (lambda: calypso_amd.LegendreTransform(7, 12, np.linspace(0.5, 1.5, 3), 3, 1, 0, 'hip_fp64'),\n"
        "             lambda: mhd.MhdSimulation(*mhd.dynamobench_case1()),\n"
        "             lambda: calypso_amd.load()):\n"
        "    try:\n"
        "        make()\n"
        "    except L.CalypsoError as e:\n"
        "        fails += 'not found' in str(e)\n"
        "print(fails)\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "3"
