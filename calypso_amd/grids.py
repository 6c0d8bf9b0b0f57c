"""One-domain grid tables of the spherical shell (host only): the local
mode order of the reference's rj / rlm arrays and its grid files
<prefix>.0.{rj,rlm,rtm,rtp} (gen_sph_grids, sph_modes_grids_data_IO.f90)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_P = ctypes.c_void_p


def rj_mode_order(l_truncation: int) -> np.ndarray:
    """Global mode index l(l+1)+m of each local rj mode (orders -L..L,
    degrees |m|..L inside)."""
    out = np.zeros((l_truncation + 1) ** 2, dtype=np.int32)
    _lib.check(_lib.load().calypso_sph_rj_mode_order(l_truncation, out.ctypes.data_as(_P)),
               "calypso_sph_rj_mode_order")
    return out


def write_grid_files(prefix, l_truncation, nth, nphi, radius):
    """Write <prefix>.0.rj, .rlm, .rtm, .rtp as the reference's generator does."""
    r = np.ascontiguousarray(radius, dtype=np.float64)
    _lib.check(_lib.load().calypso_sph_write_grid_files(
        str(prefix).encode(), l_truncation, nth, nphi, r.size, r.ctypes.data_as(_P)),
        "calypso_sph_write_grid_files")
