"""ctypes binding of libcalypso_amd.so (include/calypso_amd.h).

The library is loaded from ``calypso_amd/lib`` only.  There is no CPU
fallback: if the HIP library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libcalypso_amd.so")

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double

# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("calypso_last_error", ctypes.c_char_p, []),
    ("calypso_legendre_mode_id", _I, [ctypes.c_char_p]),
    ("calypso_gauss_points", _I, [_I, _P, _P, _P]),
    ("calypso_leg_create", _I, [_I, _I, _I, _P, _P, _P, _I, _I, _I, _I, ctypes.POINTER(_P)]),
    ("calypso_leg_destroy", _I, [_P]),
    ("calypso_leg_backward_device", _I, [_P, _I, _I, _I, _P, _P, _P]),
    ("calypso_leg_forward_device", _I, [_P, _I, _I, _I, _P, _P, _P]),
    ("calypso_leg_backward", _I, [_P, _I, _I, _I, _I, _I, _P, _P]),
    ("calypso_leg_forward", _I, [_P, _I, _I, _I, _I, _I, _P, _P]),
    ("calypso_leg_tables", _I, [_P, _P, _P]),
    ("calypso_leg_info", _I, [_P, _P]),
    ("calypso_leg_stream", _P, [_P]),
    ("calypso_leg_plan_info", _I, [_P, _I, _I, _I, _I, _P]),
    ("calypso_leg_set_autotune", _I, [_P, _I]),
    ("calypso_leg_plan_count", _I, [_P, _I, _I, _I, _I]),
    ("calypso_leg_force_plan", _I, [_P, _I, _I, _I, _I, _I]),
    ("calypso_sel_init_legendre_trans", _I, [_I, _I, _I, _I, _P, _P, _P, _I, _I, _I]),
    ("calypso_sel_backward_legendre_trans", _I, [_I, _I, _I, _I, _I, _P, _P]),
    ("calypso_sel_forward_legendre_trans", _I, [_I, _I, _I, _I, _I, _P, _P]),
    ("calypso_sel_finalize_legendre_trans", _I, []),
    ("calypso_init_MHD_FFT_select", _I, [_I, _I, _I, _I, _I, _I, _I]),
    ("calypso_back_MHD_FFT_sel_from_recv", _I, [_I, _I, _P, _P]),
    ("calypso_fwd_MHD_FFT_sel_from_recv", _I, [_I, _I, _P, _P]),
    ("calypso_finalize_MHD_FFT_select", _I, []),
    ("calypso_nonlinear_products", _I, [ctypes.c_longlong, _P, _P, _P]),
    ("calypso_lubksb_3band_mul", _I, [_I, _P, _I, _I, _P, _P, _P]),
    ("calypso_lubksb_5band_mul", _I, [_I, _P, _I, _I, _P, _P, _P]),
    ("calypso_fft_create", _I, [_I, _I, _I, _I, ctypes.POINTER(_P)]),
    ("calypso_fft_destroy", _I, [_P]),
    ("calypso_fft_backward_device", _I, [_P, _I, _P, _P, _P]),
    ("calypso_fft_forward_device", _I, [_P, _I, _P, _P, _P]),
    ("calypso_nonlinear_products_device", _I, [ctypes.c_longlong, _P, _P, _P, _P]),
    ("calypso_fft_mhd_fused_device", _I, [_P, _P, _P, _P, _P]),
    ("calypso_ludcmp_band_device", _I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    ("calypso_lubksb_band_mul_device", _I, [_I, _I, _I, _P, _P, _P, _P]),
    ("calypso_explicit_update_device", _I, [_I, _I, ctypes.c_longlong, _D, _D, _P, _P, _P, _P, _P]),
    ("calypso_sph_mean_square", _I, [_I, _I, _I, _P, _P, _I, _I, _P, _P, _P]),
    ("calypso_sph_mean_square_noalloc", _I, [_I, _I, _I, _P, _P, _I, _I, _P, _P, _P]),
    ("calypso_mhd_create", _I, [_P, _P, _P, ctypes.POINTER(_P)]),
    ("calypso_mhd_destroy", _I, [_P]),
    ("calypso_mhd_set_field", _I, [_P, _I, _P]),
    ("calypso_mhd_snapshot_rtp", _I, [_P, _I, _P, _I, _P, _P]),
    ("calypso_mhd_get_field", _I, [_P, _I, _P]),
    ("calypso_mhd_set_dynamobench_ic", _I, [_P, _I]),
    ("calypso_mhd_set_dynamobench2_ic", _I, [_P, _I]),
    ("calypso_sph_write_grid_files", _I, [ctypes.c_char_p, _I, _I, _I, _I, _P]),
    ("calypso_sph_rj_mode_order", _I, [_I, _P]),
    ("calypso_mhd_set_rj_order", _I, [_P, _I, _P]),
    ("calypso_mhd_set_seeded_ic", _I, [_P, ctypes.c_ulonglong, _D, _D, _D]),
    ("calypso_seeded_spectrum", _I, [ctypes.c_ulonglong, _I, _I, _I, _P, _D, _D, _I, _I, _P]),
    ("calypso_mhd_initialize", _I, [_P]),
    ("calypso_mhd_step", _I, [_P, _I]),
    ("calypso_mhd_energies", _I, [_P, _P]),
    ("calypso_mhd_pressure", _I, [_P, _P]),
    ("calypso_mhd_energy_fluxes", _I, [_P, _I, _P, _P]),
    ("calypso_mhd_info", _I, [_P, _P]),
    ("calypso_mhd_set_option", _I, [_P, ctypes.c_char_p, ctypes.c_longlong]),
    ("calypso_mhd_time", _I, [_P, _P]),
    ("calypso_mhd_write_volume_ms", _I, [_P, ctypes.c_char_p]),
    ("calypso_mhd_write_restart", _I, [_P, ctypes.c_char_p, _I, ctypes.c_char_p, _I]),
    ("calypso_mhd_read_restart", _I, [_P, ctypes.c_char_p]),
    ("calypso_mhd_coriolis_table", _I, [_I, _P, _P]),
    ("calypso_sph_write_field_file", _I, [ctypes.c_char_p, _I, _I, _D, _D, ctypes.c_longlong, _I,
                                          _P, _P, _P]),
    ("calypso_scalar_shell_create", _I, [_I, _I, _P, _D, _I, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    ("calypso_scalar_shell_destroy", _I, [_P]),
    ("calypso_scalar_shell_set_noise", _I, [_P]),
    ("calypso_scalar_shell_set_field", _I, [_P, _I, _P]),
    ("calypso_scalar_shell_get_field", _I, [_P, _I, _P]),
    ("calypso_scalar_shell_initialize", _I, [_P]),
    ("calypso_scalar_shell_step", _I, [_P, _I]),
    ("calypso_scalar_shell_mean_square", _I, [_P, _P]),
    ("calypso_scalar_shell_write_volume_ms", _I, [_P, ctypes.c_char_p]),
]

_lib = None


class CalypsoError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CalypsoError(
                f"{LIB_PATH} not found: build it with `python -m calypso_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = load().calypso_last_error().decode(errors="replace")
        raise CalypsoError(f"{what}: {msg}" if what else msg)
