"""Python face of the longitude FFT, grid products and rtm<->rtp exchange.

Mirrors the reference module MHD_FFT_selector (MHD_FFT_selector.F90:86-249):
``back_MHD_FFT_sel_from_recv`` -> ``LongitudeFFT.backward``,
``fwd_MHD_FFT_sel_from_recv`` -> ``LongitudeFFT.forward``, and
``s_cal_nonlinear_sph_MHD`` (cal_nonlinear_sph_MHD.f90:35-75) ->
``nonlinear_products``.  Device tensors only (torch on ROCm supplies memory
and streams); the arithmetic is the HIP library's.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_P = ctypes.c_void_p


def _stream(stream):
    if stream is not None:
        return _P(stream)
    import torch
    return _P(torch.cuda.current_stream().cuda_stream)


def _dev(t, n, what):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
            and t.is_contiguous()):
        raise _lib.CalypsoError(f"{what}: contiguous float64 GPU tensor required")
    if t.numel() < n:
        raise _lib.CalypsoError(f"{what}: {t.numel()} elements, needs {n}")
    return _P(t.data_ptr())


class LongitudeFFT:
    def __init__(self, l_truncation: int, nth: int, nr: int, nphi: int):
        self.L, self.nth, self.nr, self.nphi = l_truncation, nth, nr, nphi
        self.M = 2 * l_truncation + 1
        self.nnod_rtp = nr * nth * nphi
        h = _P()
        _lib.check(_lib.load().calypso_fft_create(l_truncation, nth, nr, nphi, ctypes.byref(h)),
                   "calypso_fft_create")
        self._h = h

    def n_rtm(self, ncomp):
        return ncomp * self.nth * self.nr * self.M

    def backward(self, vr_rtm, fld_rtp, ncomp, stream=None):
        _lib.check(_lib.load().calypso_fft_backward_device(
            self._h, ncomp, _dev(vr_rtm, self.n_rtm(ncomp), "vr_rtm"),
            _dev(fld_rtp, ncomp * self.nnod_rtp, "fld_rtp"), _stream(stream)),
            "calypso_fft_backward_device")
        return fld_rtp

    def forward(self, frc_rtp, vr_rtm, ncomp, stream=None):
        _lib.check(_lib.load().calypso_fft_forward_device(
            self._h, ncomp, _dev(frc_rtp, ncomp * self.nnod_rtp, "frc_rtp"),
            _dev(vr_rtm, self.n_rtm(ncomp), "vr_rtm"), _stream(stream)),
            "calypso_fft_forward_device")
        return vr_rtm

    def mhd_fused(self, bwd_rtm, fwd_rtm, coef, stream=None):
        c = np.ascontiguousarray(coef, dtype=np.float64)
        _lib.check(_lib.load().calypso_fft_mhd_fused_device(
            self._h, _dev(bwd_rtm, self.n_rtm(13), "bwd_rtm"),
            _dev(fwd_rtm, self.n_rtm(19), "fwd_rtm"), c.ctypes.data_as(_P), _stream(stream)),
            "calypso_fft_mhd_fused_device")
        return fwd_rtm

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().calypso_fft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def nonlinear_products(fld_rtp, frc_rtp, nnod, coef, stream=None):
    c = np.ascontiguousarray(coef, dtype=np.float64)
    _lib.check(_lib.load().calypso_nonlinear_products_device(
        nnod, _dev(fld_rtp, 13 * nnod, "fld_rtp"), _dev(frc_rtp, 19 * nnod, "frc_rtp"),
        c.ctypes.data_as(_P), _stream(stream)), "calypso_nonlinear_products_device")
    return frc_rtp
