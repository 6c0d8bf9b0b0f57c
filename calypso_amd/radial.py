"""Python face of the radial solvers: the reference's direct band solver
(SOLVER_src/direct_solver: ludcmp_3band, ludcmp_band, lubksb_3band_mul,
lubksb_5band_mul) on device tensors, batched over modes."""
from __future__ import annotations

import ctypes

from . import _lib
from .fft import _dev, _stream

_P = ctypes.c_void_p


def ludcmp_band(nb, band_a, band_lu, i_pivot, ierr=None, stream=None):
    """band_a (mcomp, n, nb) float64 -> band_lu (mcomp, n, 2nb-1), i_pivot (mcomp, n) int32."""
    import torch
    mcomp, n, k = band_a.shape
    if k != nb or band_lu.shape != (mcomp, n, 2 * nb - 1) or i_pivot.shape != (mcomp, n):
        raise _lib.CalypsoError("ludcmp_band: shape mismatch")
    if i_pivot.dtype != torch.int32 or (ierr is not None and ierr.dtype != torch.int32):
        raise _lib.CalypsoError("ludcmp_band: pivots and ierr are int32")
    _lib.check(_lib.load().calypso_ludcmp_band_device(
        nb, n, mcomp, _dev(band_a, band_a.numel(), "band_a"), _dev(band_lu, band_lu.numel(), "band_lu"),
        _P(i_pivot.data_ptr()), _P(ierr.data_ptr()) if ierr is not None else None,
        _stream(stream)), "calypso_ludcmp_band_device")


def lubksb_band_mul(nb, band_lu, i_pivot, x, stream=None):
    """Solve in place x (n, mcomp) (Fortran x(mcomp, n))."""
    mcomp, n, _ = band_lu.shape
    if tuple(x.shape) != (n, mcomp):
        raise _lib.CalypsoError("lubksb_band_mul: x must be (n, mcomp)")
    _lib.check(_lib.load().calypso_lubksb_band_mul_device(
        nb, n, mcomp, _dev(band_lu, band_lu.numel(), "band_lu"), _P(i_pivot.data_ptr()),
        _dev(x, x.numel(), "x"), _stream(stream)), "calypso_lubksb_band_mul_device")
    return x
