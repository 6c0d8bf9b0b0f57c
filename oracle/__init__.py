"""Test-only checkers for the spherical-harmonic transform path.

``oracle`` wraps two CPU libraries built by ``oracle/Makefile``:

* ``liboracle.so``            -- the plain-C FP64 restatement (calypso_oracle.c);
* ``_ref/libcalypso_ref.so``  -- the reference's own Fortran routines compiled
  from the reference sources, driven through ``ref_harness.f90``.

Only ``tests/``, ``__graft_entry__.smoke`` and the CPU-baseline leg of
``bench.py`` import this package, always as the checker, never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

_lib = None
_ref = None


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle oracle`")
        _lib = ctypes.CDLL(path)
        _lib.orc_num_modes.restype = _I
    return _lib


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libcalypso_ref.so"))


def ref():
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libcalypso_ref.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle ref`")
        _ref = ctypes.CDLL(path)
    return _ref


# ---------------------------------------------------------------- tables --

def rlm_index(L: int):
    J = (L + 1) ** 2
    jlm = np.zeros((J, 3), dtype=np.int32)
    lstack = np.zeros(2 * L + 2, dtype=np.int32)
    lib().orc_rlm_index(_I(L), _ptr(jlm), _ptr(lstack))
    return jlm, lstack


def gauss_points(n: int, use_ref: bool = False):
    point = np.zeros(n)
    weight = np.zeros(n)
    colat = np.zeros(n)
    f = ref().ref_gauss_points if use_ref else lib().orc_gauss_points
    f(_I(n), _ptr(point), _ptr(weight), _ptr(colat))
    return point, weight, colat


def schmidt_legendres_m(ltr: int, m: int, theta: float, use_ref: bool = False):
    p = np.zeros(ltr + 1)
    dp = np.zeros(ltr + 1)
    f = ref().ref_schmidt_legendres_m if use_ref else lib().orc_schmidt_legendres_m
    f(_I(ltr), _I(m), _D(theta), _ptr(p), _ptr(dp))
    return p, dp


def sph_normalizations(l: int, m: int):
    g = np.zeros(17)
    lib().orc_sph_normalizations(_I(l), _I(m), _ptr(g))
    return g


def g_sph_rlm(L: int):
    g = np.zeros(((L + 1) ** 2, 17))
    lib().orc_g_sph_rlm(_I(L), _ptr(g))
    return g


def legendre_tables(L: int, colat: np.ndarray):
    nth = colat.shape[0]
    J = (L + 1) ** 2
    P = np.zeros((J, nth))
    dP = np.zeros((J, nth))
    c = np.ascontiguousarray(colat, dtype=np.float64)
    lib().orc_legendre_tables(_I(L), _I(nth), _ptr(c), _ptr(P), _ptr(dP))
    return P, dP


# ------------------------------------------------------------ transforms --

class Grid:
    """Single-domain transform geometry: truncation L, nth Gauss points, radii."""

    def __init__(self, L: int, nth: int, radius: np.ndarray):
        self.L = L
        self.nth = nth
        self.nr = int(radius.shape[0])
        self.J = (L + 1) ** 2
        self.M = 2 * L + 1
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.point, self.weight, self.colat = gauss_points(nth)
        self.P, self.dP = legendre_tables(L, self.colat)
        self.jlm, self.lstack = rlm_index(L)
        self.g = g_sph_rlm(L)

    def n_rlm(self, ncomp):
        return ncomp * self.J * self.nr

    def n_rtm(self, ncomp):
        return ncomp * self.nth * self.nr * self.M


def leg_bwd(grid: Grid, ncomp, nvec, nscl, sp: np.ndarray) -> np.ndarray:
    sp = np.ascontiguousarray(sp, dtype=np.float64).ravel()
    assert sp.size == grid.n_rlm(ncomp)
    vr = np.zeros(grid.n_rtm(ncomp))
    lib().orc_leg_bwd(_I(grid.L), _I(grid.nth), _I(grid.nr), _I(ncomp),
                      _I(nvec), _I(nscl), _ptr(grid.radius), _ptr(grid.colat),
                      _ptr(grid.P), _ptr(grid.dP), _ptr(sp), _ptr(vr))
    return vr


def leg_fwd(grid: Grid, ncomp, nvec, nscl, vr: np.ndarray) -> np.ndarray:
    vr = np.ascontiguousarray(vr, dtype=np.float64).ravel()
    assert vr.size == grid.n_rtm(ncomp)
    sp = np.zeros(grid.n_rlm(ncomp))
    lib().orc_leg_fwd(_I(grid.L), _I(grid.nth), _I(grid.nr), _I(ncomp),
                      _I(nvec), _I(nscl), _ptr(grid.radius), _ptr(grid.colat),
                      _ptr(grid.weight), _ptr(grid.P), _ptr(grid.dP), _ptr(vr),
                      _ptr(sp))
    return sp


class RefLegendre:
    """The reference's own org-loop transforms (compiled from its sources)."""

    def __init__(self, grid: Grid, nthreads: int = 1):
        self.grid = grid
        ref().ref_leg_setup(_I(grid.L), _I(grid.nth), _I(grid.nr),
                            _ptr(grid.jlm), _ptr(grid.lstack),
                            _ptr(grid.radius), _ptr(grid.colat),
                            _ptr(grid.weight), _ptr(grid.P), _ptr(grid.dP),
                            _ptr(grid.g), _I(nthreads))

    def bwd(self, ncomp, nvec, nscl, sp):
        sp = np.ascontiguousarray(sp, dtype=np.float64).ravel()
        vr = np.zeros(self.grid.n_rtm(ncomp))
        ref().ref_leg_bwd(_I(ncomp), _I(nvec), _I(nscl), _ptr(sp), _ptr(vr))
        return vr

    def fwd(self, ncomp, nvec, nscl, vr):
        vr = np.ascontiguousarray(vr, dtype=np.float64).ravel()
        sp = np.zeros(self.grid.n_rlm(ncomp))
        ref().ref_leg_fwd(_I(ncomp), _I(nvec), _I(nscl), _ptr(vr), _ptr(sp))
        return sp

    # the SYMMETRIC_MATMUL variant (legendre_trans_sym_matmul.f90), the one
    # the reference's own tests select
    def _sym(self, nvec, nscl):
        if getattr(self, "_sym_set", None) != (nvec, nscl):
            ref().ref_leg_sym_init(_I(nvec), _I(nscl))
            self._sym_set = (nvec, nscl)

    def bwd_sym(self, ncomp, nvec, nscl, sp):
        self._sym(nvec, nscl)
        sp = np.ascontiguousarray(sp, dtype=np.float64).ravel()
        vr = np.zeros(self.grid.n_rtm(ncomp))
        ref().ref_leg_bwd_sym(_I(ncomp), _I(nvec), _I(nscl), _ptr(sp), _ptr(vr))
        return vr

    def fwd_sym(self, ncomp, nvec, nscl, vr):
        self._sym(nvec, nscl)
        vr = np.ascontiguousarray(vr, dtype=np.float64).ravel()
        sp = np.zeros(self.grid.n_rlm(ncomp))
        ref().ref_leg_fwd_sym(_I(ncomp), _I(nvec), _I(nscl), _ptr(vr), _ptr(sp))
        return sp


# ------------------------------------------------------- FFT, products --

def fft_bwd(L, nth, nr, nphi, ncomp, vr_rtm):
    vr = np.ascontiguousarray(vr_rtm, dtype=np.float64).ravel()
    out = np.zeros(ncomp * nr * nth * nphi)
    lib().orc_fft_bwd(_I(L), _I(nth), _I(nr), _I(nphi), _I(ncomp), _ptr(vr), _ptr(out))
    return out


def fft_fwd(L, nth, nr, nphi, ncomp, fld_rtp):
    x = np.ascontiguousarray(fld_rtp, dtype=np.float64).ravel()
    out = np.zeros(ncomp * nth * nr * (2 * L + 1))
    lib().orc_fft_fwd(_I(L), _I(nth), _I(nr), _I(nphi), _I(ncomp), _ptr(x), _ptr(out))
    return out


def products_mhd(nnod, fld, coef_velo, coef_lor, coef_induct, coef_temp):
    f = np.ascontiguousarray(fld, dtype=np.float64).ravel()
    assert f.size == 13 * nnod
    out = np.zeros(19 * nnod)
    lib().orc_products_mhd(ctypes.c_size_t(nnod), _ptr(f), _ptr(out), _D(coef_velo),
                           _D(coef_lor), _D(coef_induct), _D(coef_temp))
    return out


def ref_rfftm(x, forward=True):
    """FFTPACK5 RFFTMF/RFFTMB of the compiled reference on rows of x (lot x N)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    lot, n = x.shape
    r = ref()
    I = lambda v: ctypes.byref(ctypes.c_int(v))
    lensav = 2 * n + 64
    wsave = np.zeros(lensav)
    ier = ctypes.c_int(0)
    r.rfftmi_(I(n), _ptr(wsave), I(lensav), ctypes.byref(ier))
    # FFTPACK's multi-vector layout: element i of sequence s at s + i*lot
    buf = np.ascontiguousarray(x.T).ravel()
    work = np.zeros(lot * n)
    f = r.rfftmf_ if forward else r.rfftmb_
    f(I(lot), I(1), I(n), I(lot), _ptr(buf), I(lot * n), _ptr(wsave), I(lensav), _ptr(work),
      I(lot * n), ctypes.byref(ier))
    if ier.value != 0:
        raise RuntimeError(f"FFTPACK error {ier.value}")
    return buf.reshape(n, lot).T.copy()


# ------------------------------------------------------------ banded LU --

def ludcmp(nb, band_a, use_ref=False):
    """band_a: (mcomp, n, nb) C-order == Fortran band_a(nb, n, mcomp)."""
    band_a = np.ascontiguousarray(band_a, dtype=np.float64)
    mcomp, n, _ = band_a.shape
    w = 5 if nb == 3 else 9
    lu = np.zeros((mcomp, n, w))
    piv = np.zeros((mcomp, n), dtype=np.int32)
    if nb == 3:
        ierr = np.zeros(mcomp, dtype=np.int32)
        f = ref().ref_ludcmp_3band if use_ref else lib().orc_ludcmp_3band
        if use_ref:
            f(_I(n), _I(mcomp), _ptr(band_a), _ptr(lu), _ptr(piv), _ptr(ierr))
        else:
            f(_I(n), _I(mcomp), _ptr(band_a), _ptr(lu), _ptr(piv), _ptr(ierr))
    else:
        f = ref().ref_ludcmp_5band if use_ref else lib().orc_ludcmp_5band
        f(_I(n), _I(mcomp), _ptr(band_a), _ptr(lu), _ptr(piv))
    return lu, piv


def lubksb(nb, lu, piv, x, use_ref=False):
    """x: (n, mcomp) C-order == Fortran x(mcomp, n); returns the solution."""
    lu = np.ascontiguousarray(lu, dtype=np.float64)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    x = np.array(x, dtype=np.float64, order="C")
    mcomp, n, _ = lu.shape
    f = ref().ref_lubksb_mul if use_ref else lib().orc_lubksb_mul
    f(_I(nb), _I(n), _I(mcomp), _ptr(lu), _ptr(piv), _ptr(x))
    return x


# ------------------------------------------------- explicit step, energies --

def explicit_update(kind, scheme, dt, coef_exp, x, diff, force, pre):
    x = np.array(x, dtype=np.float64)
    pre = np.array(pre, dtype=np.float64)
    diff = np.ascontiguousarray(diff, dtype=np.float64)
    force = np.ascontiguousarray(force, dtype=np.float64)
    lib().orc_explicit_update(_I(kind), _I(scheme), ctypes.c_size_t(x.size), _D(dt),
                              _D(coef_exp), _ptr(x), _ptr(diff), _ptr(force), _ptr(pre))
    return x, pre


def sph_mean_square(kind, radius, l_of_j, kr_st, kr_ed, f):
    """f: (ncomp, nr, J) C-order d_rj slice of one field."""
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    l_of_j = np.ascontiguousarray(l_of_j, dtype=np.int32)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_sph_mean_square(_I(kind), _I(radius.size), _I(l_of_j.size), _ptr(l_of_j),
                              _I(int(l_of_j.max())), _ptr(radius), _I(kr_st), _I(kr_ed),
                              _ptr(f), _ptr(out))
    return out


# ------------------------------------------------------- seeded inputs --

def splitmix64(state: int) -> int:
    lib().orc_splitmix64.restype = ctypes.c_uint64
    lib().orc_splitmix64.argtypes = [ctypes.c_uint64]
    return int(lib().orc_splitmix64(state))


def box_muller(u1: float, u2: float) -> float:
    lib().orc_box_muller.restype = _D
    return float(lib().orc_box_muller(_D(u1), _D(u2)))


def seeded_spectrum(seed, field, L, radius, amplitude=1.0, power=2.0, envelope=True, l_min=0):
    """(nr, J) seeded spectrum, the C restatement orc_seeded_spectrum."""
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    out = np.zeros((radius.size, (L + 1) ** 2))
    lib().orc_seeded_spectrum(ctypes.c_uint64(int(seed)), _I(field), _I(L), _I(radius.size),
                              _ptr(radius), _D(amplitude), _D(power), _I(1 if envelope else 0),
                              _I(l_min), _ptr(out))
    return out
