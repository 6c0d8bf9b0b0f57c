"""Python face of the radial kernels: the reference's direct band solver
(SOLVER_src/direct_solver: ludcmp_3band, ludcmp_band, lubksb_3band_mul,
lubksb_5band_mul) on device tensors, batched over modes; the explicit
Adams-Bashforth/Euler update; the volume mean-square (energy) diagnostics."""
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


def explicit_update(kind, scheme, dt, coef_exp, x, diffuse, force, pre=None, stream=None):
    """In-place explicit update of one d_rj component (see include/calypso_amd.h)."""
    n = x.numel()
    _lib.check(_lib.load().calypso_explicit_update_device(
        kind, scheme, n, dt, coef_exp, _dev(x, n, "x"), _dev(diffuse, n, "diffuse"),
        _dev(force, n, "force"), _dev(pre, n, "pre") if pre is not None else None,
        _stream(stream)), "calypso_explicit_update_device")


def sph_mean_square(kind, radius, l_of_j, kr_st, kr_ed, f, stream=None, noalloc=False):
    """Volume mean square of one field; f is a (ncomp, nr, J) device tensor.
    noalloc=True uses the entry point that allocates no device memory."""
    import numpy as np
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    l_of_j = np.ascontiguousarray(l_of_j, dtype=np.int32)
    out = np.zeros(3)
    name = "calypso_sph_mean_square_noalloc" if noalloc else "calypso_sph_mean_square"
    _lib.check(getattr(_lib.load(), name)(
        kind, radius.size, l_of_j.size, l_of_j.ctypes.data_as(_P), radius.ctypes.data_as(_P),
        kr_st, kr_ed, _dev(f, f.numel(), "f"), out.ctypes.data_as(_P), _stream(stream)), name)
    return out
