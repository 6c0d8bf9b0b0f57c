"""Python face of the GPU Legendre transforms.

Mirrors the reference's operator boundary, module ``legendre_transform_select``
(SPH_SHELL_src/legendre_transform_select.f90):

=====================================  ==========================================
reference                              here
=====================================  ==========================================
set_legendre_trans_mode_ctl (:181)     ``set_legendre_trans_mode_ctl(name)``
sel_init_legendre_trans (:228)         ``sel_init_legendre_trans(...)``
sel_backward_legendre_trans (:309)     ``sel_backward_legendre_trans(ncomp, nvector,
                                       nscalar, n_WR, n_WS, WR, WS)``
sel_forward_legendre_trans (:375)      ``sel_forward_legendre_trans(...)``
sel_finalize_legendre_trans (:269)     ``sel_finalize_legendre_trans()``
=====================================  ==========================================

``LegendreTransform`` is the handle-based form used with device tensors
(torch on ROCm provides the memory and the stream; the arithmetic is the HIP
library's).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

LEG_MODES = {"hip_fp64": 1, "hip_fp32": 2, "hip_bf16x3": 3, "hip_int8x6": 4}
_P = ctypes.c_void_p


def set_legendre_trans_mode_ctl(name: str) -> int:
    """Mode string -> id; unknown strings raise (the GPU build has no CPU loops)."""
    mid = _lib.load().calypso_legendre_mode_id(name.encode())
    if mid < 0:
        raise _lib.CalypsoError(f"unknown Legendre transform mode {name!r}; "
                                f"expected one of {sorted(LEG_MODES)}")
    return mid


def gauss_points(n: int):
    p, w, c = np.zeros(n), np.zeros(n), np.zeros(n)
    _lib.check(_lib.load().calypso_gauss_points(n, p.ctypes.data, w.ctypes.data, c.ctypes.data),
               "calypso_gauss_points")
    return p, w, c


def _dptr(x):
    return x.ctypes.data_as(_P) if isinstance(x, np.ndarray) else _P(x.data_ptr())


class LegendreTransform:
    """Device-resident transform context for one (L, nth, radii) grid."""

    def __init__(self, l_truncation: int, nth: int, radius, ncomp: int, nvector: int,
                 nscalar: int, mode="hip_fp64", colat=None, weight=None):
        lib = _lib.load()
        self.L, self.nth = int(l_truncation), int(nth)
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.nr = int(self.radius.size)
        self.J, self.M = (self.L + 1) ** 2, 2 * self.L + 1
        self.ncomp, self.nvector, self.nscalar = ncomp, nvector, nscalar
        self.mode = mode if isinstance(mode, int) else set_legendre_trans_mode_ctl(mode)
        c = np.ascontiguousarray(colat, dtype=np.float64) if colat is not None else None
        w = np.ascontiguousarray(weight, dtype=np.float64) if weight is not None else None
        h = _P()
        _lib.check(lib.calypso_leg_create(self.L, self.nth, self.nr, self.radius.ctypes.data,
                                          c.ctypes.data if c is not None else None,
                                          w.ctypes.data if w is not None else None,
                                          ncomp, nvector, nscalar, self.mode, ctypes.byref(h)),
                   "calypso_leg_create")
        self._h = h

    # sizes of the two spaces for a field set
    def n_rlm(self, ncomp=None):
        return (ncomp or self.ncomp) * self.J * self.nr

    def n_rtm(self, ncomp=None):
        return (ncomp or self.ncomp) * self.nth * self.nr * self.M

    def _fs(self, ncomp, nvector, nscalar):
        if ncomp is None:
            return self.ncomp, self.nvector, self.nscalar
        return ncomp, nvector, nscalar

    def backward(self, sp_rlm, vr_rtm, ncomp=None, nvector=None, nscalar=None, stream=None):
        """vr_rtm <- backward(sp_rlm); torch device tensors (async) or numpy (sync)."""
        nc, nv, ns = self._fs(ncomp, nvector, nscalar)
        lib = _lib.load()
        if isinstance(sp_rlm, np.ndarray):
            _lib.check(lib.calypso_leg_backward(self._h, nc, nv, ns, sp_rlm.size, vr_rtm.size,
                                                _dptr(sp_rlm), _dptr(vr_rtm)),
                       "calypso_leg_backward")
        else:
            self._check_dev(sp_rlm, self.n_rlm(nc), vr_rtm, self.n_rtm(nc))
            _lib.check(lib.calypso_leg_backward_device(self._h, nc, nv, ns, _dptr(sp_rlm),
                                                       _dptr(vr_rtm), self._stream(stream)),
                       "calypso_leg_backward_device")
        return vr_rtm

    def forward(self, vr_rtm, sp_rlm, ncomp=None, nvector=None, nscalar=None, stream=None):
        """sp_rlm <- forward(vr_rtm)."""
        nc, nv, ns = self._fs(ncomp, nvector, nscalar)
        lib = _lib.load()
        if isinstance(vr_rtm, np.ndarray):
            _lib.check(lib.calypso_leg_forward(self._h, nc, nv, ns, vr_rtm.size, sp_rlm.size,
                                               _dptr(vr_rtm), _dptr(sp_rlm)),
                       "calypso_leg_forward")
        else:
            self._check_dev(vr_rtm, self.n_rtm(nc), sp_rlm, self.n_rlm(nc))
            _lib.check(lib.calypso_leg_forward_device(self._h, nc, nv, ns, _dptr(vr_rtm),
                                                      _dptr(sp_rlm), self._stream(stream)),
                       "calypso_leg_forward_device")
        return sp_rlm

    @staticmethod
    def _check_dev(x, nx, y, ny):
        import torch
        for t, n in ((x, nx), (y, ny)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64
                    and t.is_contiguous()):
                raise _lib.CalypsoError("device buffers must be contiguous float64 GPU tensors")
            if t.numel() < n:
                raise _lib.CalypsoError(f"device buffer has {t.numel()} elements, needs {n}")

    @staticmethod
    def _stream(stream):
        if stream is not None:
            return _P(stream)
        import torch
        return _P(torch.cuda.current_stream().cuda_stream)

    def stream(self):
        return _lib.load().calypso_leg_stream(self._h)

    def tables(self):
        """Device-computed P_rtm, dPdt_rtm as (J, nth) arrays (rlm order)."""
        P = np.zeros((self.J, self.nth))
        dP = np.zeros((self.J, self.nth))
        _lib.check(_lib.load().calypso_leg_tables(self._h, _dptr(P), _dptr(dP)),
                   "calypso_leg_tables")
        return P, dP

    def info(self):
        a = np.zeros(8, dtype=np.int64)
        _lib.check(_lib.load().calypso_leg_info(self._h, a.ctypes.data), "calypso_leg_info")
        keys = ["L", "nth", "nr", "mode", "table_bytes", "kt_bwd", "kt_fwd", "wg_bwd"]
        return dict(zip(keys, (int(v) for v in a)))

    def plan_info(self, fwd: bool, ncomp=None, nvector=None, nscalar=None):
        nc, nv, ns = self._fs(ncomp, nvector, nscalar)
        a = np.zeros(11, dtype=np.int64)
        _lib.check(_lib.load().calypso_leg_plan_info(self._h, nc, nv, ns, int(fwd), a.ctypes.data),
                   "calypso_leg_plan_info")
        keys = ["KT", "mct1", "mct2", "rows_per_wave", "pair", "tc", "tuned_us", "work_groups",
                "kernel_form", "waves", "prefetch"]
        return dict(zip(keys, (int(v) for v in a)))

    def plan_count(self, fwd: bool, ncomp=None, nvector=None, nscalar=None) -> int:
        nc, nv, ns = self._fs(ncomp, nvector, nscalar)
        n = _lib.load().calypso_leg_plan_count(self._h, nc, nv, ns, int(fwd))
        if n < 0:
            _lib.check(n, "calypso_leg_plan_count")
        return n

    def force_plan(self, fwd: bool, idx: int, ncomp=None, nvector=None, nscalar=None):
        nc, nv, ns = self._fs(ncomp, nvector, nscalar)
        _lib.check(_lib.load().calypso_leg_force_plan(self._h, nc, nv, ns, int(fwd), idx),
                   "calypso_leg_force_plan")

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().calypso_leg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------- process-wide selector --

def sel_init_legendre_trans(mode, l_truncation, nth, radius, colat, weight, ncomp, nvector,
                            nscalar):
    mid = mode if isinstance(mode, int) else set_legendre_trans_mode_ctl(mode)
    r = np.ascontiguousarray(radius, dtype=np.float64)
    c = np.ascontiguousarray(colat, dtype=np.float64)
    w = np.ascontiguousarray(weight, dtype=np.float64)
    _lib.check(_lib.load().calypso_sel_init_legendre_trans(
        mid, l_truncation, nth, r.size, r.ctypes.data, c.ctypes.data, w.ctypes.data, ncomp,
        nvector, nscalar), "sel_init_legendre_trans")


def sel_backward_legendre_trans(ncomp, nvector, nscalar, n_WR, n_WS, WR, WS):
    _lib.check(_lib.load().calypso_sel_backward_legendre_trans(
        ncomp, nvector, nscalar, n_WR, n_WS, _dptr(WR), _dptr(WS)), "sel_backward_legendre_trans")


def sel_forward_legendre_trans(ncomp, nvector, nscalar, n_WR, n_WS, WR, WS):
    _lib.check(_lib.load().calypso_sel_forward_legendre_trans(
        ncomp, nvector, nscalar, n_WR, n_WS, _dptr(WR), _dptr(WS)), "sel_forward_legendre_trans")


def sel_finalize_legendre_trans():
    _lib.check(_lib.load().calypso_sel_finalize_legendre_trans(), "sel_finalize_legendre_trans")
