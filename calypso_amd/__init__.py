"""calypso_amd -- MI355X (gfx950) spherical-harmonic MHD time step of the
Calypso dynamo code, as hand-written HIP kernels behind a C ABI
(include/calypso_amd.h), with Fortran (calypso_amd/fortran) and Python faces.

There is no CPU fallback: the HIP library must be built
(``python -m calypso_amd.build``) and a GPU present for any transform.
"""
from ._lib import CalypsoError, LIB_PATH, load  # noqa: F401
from .fft import LongitudeFFT, nonlinear_products  # noqa: F401
from .mhd import MhdParams, MhdSimulation, dynamobench_case1  # noqa: F401
from .legendre import (  # noqa: F401
    LEG_MODES,
    LegendreTransform,
    gauss_points,
    sel_backward_legendre_trans,
    sel_finalize_legendre_trans,
    sel_forward_legendre_trans,
    sel_init_legendre_trans,
    set_legendre_trans_mode_ctl,
)

__all__ = [
    "CalypsoError", "LIB_PATH", "load", "LEG_MODES", "LegendreTransform", "gauss_points",
    "sel_init_legendre_trans", "sel_backward_legendre_trans", "sel_forward_legendre_trans",
    "sel_finalize_legendre_trans", "set_legendre_trans_mode_ctl", "LongitudeFFT",
    "nonlinear_products", "MhdParams", "MhdSimulation", "dynamobench_case1",
]
