"""Python face of the scalar-shell model (libcalypso_amd.so): scalars
diffusing in a shell at rest with per-mode boundary values or fluxes.

Mirrors the reference's run of tests/heterogineous_temp: the boundary data
come from its spectral file (read_boundary_spectr_file,
m_sph_boundary_input_data.f90:147-300), the initial fields from
set_noize_scalar_sph (set_initial_sph_dynamo.f90:390-441), the step from
SPH_analyze_MHD with only the scalars evolved, the output from
sph_pwr_volume_s.dat.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .mhd import chebyshev_radius

_P = ctypes.c_void_p
FIXED, FLUX = 0, 1


class ScalarBC(ctypes.Structure):
    _fields_ = [("coef_d", ctypes.c_double), ("coef_imp", ctypes.c_double),
                ("bc_icb", ctypes.c_int), ("bc_cmb", ctypes.c_int),
                ("name", ctypes.c_char_p)]


def read_bc_spectr(path):
    """The spectral boundary file (read_boundary_spectr_file): a count of
    blocks, then per block the field name, the group (ICB / CMB), the number
    of modes and lines `l m value`; '#' lines are comments.  Returns
    {(field, GROUP): [(l, m, value), ...]}."""
    with open(path) as f:
        lines = [l.strip() for l in f if l.strip() and not l.lstrip().startswith("#")]
    n = int(lines[0].split()[0])
    out, i = {}, 1
    for _ in range(n):
        field, group = lines[i].split()[0], lines[i + 1].split()[0]
        nm = int(lines[i + 2].split()[0])
        rows = []
        for q in range(nm):
            t = lines[i + 3 + q].split()
            rows.append((int(t[0]), int(t[1]), float(t[2])))
        out[(field.lower(), group.upper())] = rows
        i += 3 + nm
    return out


def mode_values(rows, l_truncation):
    """Per-mode array (index l(l+1)+m) of (l, m, value) rows; rows beyond the
    truncation are dropped, as set_fixed_scalar_bc_by_file does."""
    J = (l_truncation + 1) ** 2
    v = np.zeros(J)
    for l, m, x in rows:
        if 0 <= l <= l_truncation and abs(m) <= l:
            v[l * (l + 1) + m] = x
    return v


class ScalarShell:
    """Scalars (each a dict: name, coef_d, coef_imp, icb=(kind, values[J]),
    cmb=(kind, values[J])) on the radial grid `radius`."""

    def __init__(self, l_truncation, radius, dt, scalars, stream=None):
        self.L = int(l_truncation)
        self.J = (self.L + 1) ** 2
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        self.nr = self.radius.size
        self.dt = float(dt)
        self.names = [s.get("name", f"scalar_{q + 1}") for q, s in enumerate(scalars)]
        self.nsc = len(scalars)
        self._names_b = [n.encode() for n in self.names]
        bcs = (ScalarBC * self.nsc)()
        vi = np.zeros((self.nsc, self.J))
        vc = np.zeros((self.nsc, self.J))
        for q, s in enumerate(scalars):
            bcs[q].coef_d = s["coef_d"]
            bcs[q].coef_imp = s["coef_imp"]
            bcs[q].bc_icb = s["icb"][0]
            bcs[q].bc_cmb = s["cmb"][0]
            bcs[q].name = self._names_b[q]
            vi[q] = s["icb"][1]
            vc[q] = s["cmb"][1]
        lib = _lib.load()
        h = _P()
        _lib.check(lib.calypso_scalar_shell_create(
            self.L, self.nr, self.radius.ctypes.data, self.dt, self.nsc,
            ctypes.cast(bcs, _P), vi.ctypes.data, vc.ctypes.data,
            _P(stream) if stream else None, ctypes.byref(h)), "calypso_scalar_shell_create")
        self._h = h
        self._lib = lib

    def close(self):
        if getattr(self, "_h", None):
            self._lib.calypso_scalar_shell_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_noise(self):
        _lib.check(self._lib.calypso_scalar_shell_set_noise(self._h))

    def set_field(self, q, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.nr, self.J)
        _lib.check(self._lib.calypso_scalar_shell_set_field(self._h, int(q), x.ctypes.data))

    def get_field(self, q, which=0):
        """which: 0 the scalar, 1 its diffusion term, 2 advection, 3 Adams history."""
        out = np.empty((self.nr, self.J))
        _lib.check(self._lib.calypso_scalar_shell_get_field(
            self._h, int(which) * self.nsc + int(q), out.ctypes.data))
        return out

    def initialize(self):
        _lib.check(self._lib.calypso_scalar_shell_initialize(self._h))

    def step(self, nsteps=1):
        _lib.check(self._lib.calypso_scalar_shell_step(self._h, int(nsteps)))

    def mean_squares(self):
        out = np.empty(self.nsc)
        _lib.check(self._lib.calypso_scalar_shell_mean_square(self._h, out.ctypes.data))
        return out

    def write_volume_ms(self, path):
        _lib.check(self._lib.calypso_scalar_shell_write_volume_ms(self._h, str(path).encode()))


def heterogeneous_temp_params(bc_path, l_truncation=4, nri=64):
    """tests/heterogineous_temp (control_MHD, control_sph_shell, bc_spectr.btx):
    r_i/r_o = 0.35 with r_o - r_i = 1, nri Chebyshev intervals, dt = 1e-4;
    temperature: Pr = 1 (diffusivity 1), Crank-Nicolson weight 0.6, value at
    the ICB, flux at the CMB; composition: composite Prandtl number 2
    (diffusivity 1/2), weight 0.5 (the control file spells the key
    coef_imp_d_ctl, which the reader does not know, so the default stays),
    flux at the ICB, value at the CMB.  Returns (radius, dt, scalars)."""
    bc = read_bc_spectr(bc_path)
    r_cmb = 1.0 / (1.0 - 0.35)
    r_icb = r_cmb * 0.35
    radius = chebyshev_radius(nri, r_icb, r_cmb)
    L = l_truncation
    scalars = [
        dict(name="temperature", coef_d=1.0 * 1.0 ** -1.0, coef_imp=0.6,
             icb=(FIXED, mode_values(bc[("temperature", "ICB")], L)),
             cmb=(FLUX, mode_values(bc[("heat_flux", "CMB")], L))),
        dict(name="composition", coef_d=1.0 * 2.0 ** -1.0, coef_imp=0.5,
             icb=(FLUX, mode_values(bc[("composite_flux", "ICB")], L)),
             cmb=(FIXED, mode_values(bc[("composition", "CMB")], L))),
    ]
    return radius, 1.0e-4, scalars
