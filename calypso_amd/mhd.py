"""Python face of the device-resident MHD time step (libcalypso_amd.so).

Mirrors module SPH_analyzer_MHD (programs/SPH_MHD/SPH_analyzer_MHD.f90):
``MhdSimulation.initialize`` is SPH_initialize_MHD after the initial fields,
``MhdSimulation.step`` is SPH_analyze_MHD, ``energies`` the numbers of
sph_pwr_volume_s.dat.  The coefficients are those m_physical_property holds;
``dynamobench_case1`` builds them from the reference's
tests/Dynamobench_case1 control files the way construct_coefficient
(construct_MHD_coefficient.f90) does.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .legendre import set_legendre_trans_mode_ctl

_P = ctypes.c_void_p

FIELDS = {"velocity": 0, "vorticity": 3, "magnetic_field": 6, "current_density": 9,
          "temperature": 12, "vorticity_diffusion": 13, "magnetic_diffusion": 16,
          "thermal_diffusion": 19, "forces": 20, "induction": 26, "heat_advect": 32,
          "buoyancy": 34}
ENERGY_COLUMNS = ["K_ene_pol", "K_ene_tor", "K_ene", "temperature",
                  "M_ene_pol", "M_ene_tor", "M_ene"]


class MhdParams(ctypes.Structure):
    _fields_ = [("l_truncation", ctypes.c_int), ("nth", ctypes.c_int), ("nphi", ctypes.c_int),
                ("nr", ctypes.c_int), ("legendre_mode", ctypes.c_int),
                ("dt", ctypes.c_double),
                ("coef_imp_v", ctypes.c_double), ("coef_imp_t", ctypes.c_double),
                ("coef_imp_b", ctypes.c_double),
                ("coef_velo", ctypes.c_double), ("coef_d_velo", ctypes.c_double),
                ("coef_buo", ctypes.c_double), ("coef_cor", ctypes.c_double),
                ("coef_lor", ctypes.c_double),
                ("coef_temp", ctypes.c_double), ("coef_d_temp", ctypes.c_double),
                ("coef_magne", ctypes.c_double), ("coef_d_magne", ctypes.c_double),
                ("coef_induct", ctypes.c_double),
                ("depth_high_t", ctypes.c_double), ("depth_low_t", ctypes.c_double),
                ("high_temp", ctypes.c_double), ("low_temp", ctypes.c_double),
                ("bc_velo_icb", ctypes.c_int), ("bc_velo_cmb", ctypes.c_int),
                ("kr_icb", ctypes.c_int), ("kr_cmb", ctypes.c_int),
                ("bc_magne_icb", ctypes.c_int),
                ("coef_press", ctypes.c_double)]


# velocity walls (bc_velocity: non_slip_sph, free_slip_sph, rot_inner_core)
VELO_RIGID, VELO_FREE_SLIP, VELO_ROT_INNER_CORE = 0, 1, 2
# magnetic condition below the shell (bc_magnetic_field: insulator, sph_to_center)
MAGNE_INSULATOR, MAGNE_TO_CENTER = 0, 1


def chebyshev_radius(nri: int, r_icb: float, r_cmb: float) -> np.ndarray:
    """set_chebyshev_distance_shell (set_radial_grid_sph_shell.f90) for a grid
    with no points outside the shell: nri + 1 points from r_icb to r_cmb."""
    pi = 4.0 * math.atan(1.0)
    shell = r_cmb - r_icb
    return np.array([r_icb + 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))
                     for k in range(nri + 1)])


def chebyshev_radius_ext(nri: int, r_icb: float, r_cmb: float, r_min: float, r_max: float):
    """count_chebyshev_ext_layers then set_chebyshev_distance_shell
    (set_radial_grid_sph_shell.f90:60-109, :171-229), the grid gen_sph_grids
    writes for Min_radius / Max_radius: nri Chebyshev intervals in the shell,
    the mirrored Chebyshev spacing below the ICB down towards r_min (then
    evenly spaced) and above the CMB up to r_max.  Returns (radius, kr_icb,
    kr_cmb), the boundary points 1-based as in the reference's groups."""
    pi = 4.0 * math.atan(1.0)
    shell = r_cmb - r_icb
    half = nri // 2

    def below(k):
        return r_icb - 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))

    def above(k):
        return r_cmb + 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))

    r, k, dr = r_icb, 0, 0.5 * shell * (1.0 - math.cos(pi / float(nri)))
    while r > r_min and k < half:
        k += 1
        r = below(k)
        dr = 0.5 * shell * (math.cos(pi * float(k - 1) / float(nri)) - math.cos(pi * float(k) / float(nri)))
    if k == half:
        k = int((r - r_min) / dr) + half
    n_icore = max(k - 1, 0)
    r, k = r_cmb, 0
    while r < r_max and k < half:
        k += 1
        r = above(k)
        dr = 0.5 * shell * (-math.cos(pi * float(k) / float(nri)) + math.cos(pi * float(k - 1) / float(nri)))
    while r < r_max:
        k += 1
        r = r + dr
    n_ext = 0 if k <= 1 else k
    k_icb = n_icore + 1
    k_cmb = k_icb + nri
    n = k_cmb + n_ext
    out = np.zeros(n)
    for q in range(1, k_icb - half):
        out[q - 1] = (r_icb - 0.5 * shell) * float(q) / float(k_icb - half)
    for q in range(max(1, k_icb - half), k_icb):
        out[q - 1] = below(k_icb - q)
    for q in range(k_icb, k_cmb + 1):
        out[q - 1] = r_icb + 0.5 * shell * (1.0 - math.cos(pi * float(q - k_icb) / float(nri)))
    top = min(n, k_cmb + half)
    for q in range(k_cmb + 1, top + 1):
        out[q - 1] = above(q - k_cmb)
    if top > k_cmb:
        dr = out[top - 1] - out[top - 2]
    for q in range(k_cmb + half + 1, n + 1):
        out[q - 1] = out[q - 2] + dr
    return out, k_icb, k_cmb


def dynamobench_case1(l_truncation=31, nth=48, nphi=96, nri=48, dt=5.0e-5, coef_imp=0.6,
                      legendre_mode="hip_fp64", Ra=1.0e2, bc_velo=("non_slip_sph", "non_slip_sph")):
    """Parameters of tests/Dynamobench_case1 (control_MHD, control_sph_shell):
    E = 1e-3, modified Ra = 100, Pr = 1, Pm = 5, r_i/r_o = 0.35 with
    r_o - r_i = 1; returns (MhdParams, radius).  The grid, time step,
    Crank-Nicolson weight and Rayleigh number can be changed, and the
    velocity walls (ICB, CMB) set to "non_slip_sph" or "free_slip_sph"."""
    E, Pr, Pm = 1.0e-3, 1.0, 5.0
    # set_ctl_gen_shell_grids.f90:179-185 (fluid_core_size 1, ratio 0.35)
    r_icb = 1.0 * 0.35 / (1.0 - 0.35)
    r_cmb = r_icb + 1.0
    p = MhdParams()
    p.l_truncation, p.nth, p.nphi, p.nr = l_truncation, nth, nphi, nri + 1
    p.legendre_mode = set_legendre_trans_mode_ctl(legendre_mode)
    p.dt = dt
    p.coef_imp_v = p.coef_imp_t = p.coef_imp_b = coef_imp
    # construct_coefficient: 'One' multiplies by 1, 'Two' by 2**power,
    # 'Radial_35' by (1 - 0.35), a dimensionless number by value**power
    p.coef_velo = 1.0 * 1.0
    p.coef_d_velo = 1.0 * 1.0
    p.coef_buo = 1.0 * (1.0 - 0.35) * Ra ** 1.0 * E ** -1.0
    p.coef_cor = 1.0 * 2.0 ** 1.0 * E ** -1.0
    p.coef_lor = 1.0 * Pm ** -1.0 * E ** -1.0
    p.coef_temp = 1.0 * 1.0
    p.coef_d_temp = 1.0 * Pr ** -1.0
    p.coef_magne = 1.0 * 1.0
    p.coef_d_magne = 1.0 * Pm ** -1.0
    p.coef_induct = 1.0 * 1.0
    p.coef_press = 1.0 * E ** -1.0
    p.depth_high_t, p.depth_low_t = 0.5384615384615384, 1.5384615384615384
    p.high_temp, p.low_temp = 1.0, 0.0
    kinds = {"non_slip_sph": VELO_RIGID, "free_slip_sph": VELO_FREE_SLIP}
    for name in bc_velo:
        if name not in kinds:
            raise _lib.CalypsoError(f"velocity boundary {name!r} is not supported "
                                    "(non_slip_sph, free_slip_sph)")
    p.bc_velo_icb, p.bc_velo_cmb = kinds[bc_velo[0]], kinds[bc_velo[1]]
    return p, chebyshev_radius(nri, r_icb, r_cmb)


def dynamobench_case2(l_truncation=31, nth=48, nphi=96, nri=48, dt=5.0e-5, coef_imp=0.6,
                      legendre_mode="hip_fp64", Ra=1.1e2, r_min=0.0, r_max=2.0):
    """Parameters of tests/Dynamobench_case2: case 1's shell and numbers with
    modified Ra = 110, the grid extended to Min_radius 0 and Max_radius 2
    (24 points in the inner core, 49 in the shell, 23 above the CMB for the
    default grid), a rotatable inner core (bc_velocity rot_inner_core) and the
    magnetic field solved to the centre (sph_to_center); returns
    (MhdParams, radius)."""
    p, _ = dynamobench_case1(l_truncation, nth, nphi, nri, dt, coef_imp, legendre_mode, Ra)
    r_icb = 1.0 * 0.35 / (1.0 - 0.35)
    r_cmb = r_icb + 1.0
    radius, kr_icb, kr_cmb = chebyshev_radius_ext(nri, r_icb, r_cmb, r_min, r_max)
    p.nr = radius.size
    p.kr_icb, p.kr_cmb = kr_icb, kr_cmb
    p.bc_velo_icb = VELO_ROT_INNER_CORE
    p.bc_magne_icb = MAGNE_TO_CENTER
    return p, radius


class MhdSimulation:
    """One device-resident MHD state (d_rj) and its step."""

    def __init__(self, params: MhdParams, radius, stream=None):
        lib = _lib.load()
        self.params = params
        self.radius = np.ascontiguousarray(radius, dtype=np.float64)
        if self.radius.size != params.nr:
            raise _lib.CalypsoError("radius must have params.nr points")
        self.L = params.l_truncation
        self.J = (self.L + 1) ** 2
        self.nr = params.nr
        h = _P()
        _lib.check(lib.calypso_mhd_create(ctypes.byref(params), self.radius.ctypes.data_as(_P),
                                          _P(stream) if stream else None, ctypes.byref(h)),
                   "calypso_mhd_create")
        self._h = h

    def _comp(self, field):
        if isinstance(field, str):
            base, _, c = field.partition(":")
            return FIELDS[base] + (int(c) if c else 0)
        return int(field)

    def set_field(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.nr, self.J)
        _lib.check(_lib.load().calypso_mhd_set_field(self._h, self._comp(field),
                                                     v.ctypes.data_as(_P)),
                   "calypso_mhd_set_field")

    def get_field(self, field) -> np.ndarray:
        out = np.zeros((self.nr, self.J))
        _lib.check(_lib.load().calypso_mhd_get_field(self._h, self._comp(field),
                                                     out.ctypes.data_as(_P)),
                   "calypso_mhd_get_field")
        return out

    def snapshot_rtp(self, vectors=("velocity", "magnetic_field"), scalars=("temperature",)):
        """Fields on the physical grid (sph_back_trans_snapshot_MHD): returns
        (3 len(vectors) + len(scalars), nphi, nth, nr) with the vectors as
        (r, theta, phi) components, scalars after."""
        vec = np.array([self._comp(v) for v in vectors], dtype=np.int32)
        scl = np.array([self._comp(v) for v in scalars], dtype=np.int32)
        nc = 3 * vec.size + scl.size
        p = self.params
        out = np.zeros((nc, p.nphi, p.nth, p.nr))
        _lib.check(_lib.load().calypso_mhd_snapshot_rtp(
            self._h, vec.size, vec.ctypes.data_as(_P), scl.size, scl.ctypes.data_as(_P),
            out.ctypes.data_as(_P)), "calypso_mhd_snapshot_rtp")
        return out

    def set_dynamobench_ic(self, isig=400):
        _lib.check(_lib.load().calypso_mhd_set_dynamobench_ic(self._h, isig),
                   "calypso_mhd_set_dynamobench_ic")

    def set_dynamobench2_ic(self, isig=400):
        """dynamo_benchmark_2: initial_b_dynamobench_2's field to the centre."""
        _lib.check(_lib.load().calypso_mhd_set_dynamobench2_ic(self._h, isig),
                   "calypso_mhd_set_dynamobench2_ic")

    def set_rj_order(self, j_global):
        """Take and give d_rj (set_field, get_field, restart files) in the
        caller's local mode order: j_global[j] = l(l+1)+m of local mode j."""
        jg = np.ascontiguousarray(j_global, dtype=np.int32)
        _lib.check(_lib.load().calypso_mhd_set_rj_order(self._h, jg.size, jg.ctypes.data_as(_P)),
                   "calypso_mhd_set_rj_order")

    def set_seeded_ic(self, seed, amp_u=1.0, amp_b=1.0, amp_t=0.1):
        """Seeded random u, B (l >= 1) and temperature perturbation with
        amplitude (l+1)^-2 and the wall envelope (see include/calypso_amd.h)."""
        _lib.check(_lib.load().calypso_mhd_set_seeded_ic(self._h, int(seed), float(amp_u),
                                                         float(amp_b), float(amp_t)),
                   "calypso_mhd_set_seeded_ic")

    def initialize(self):
        _lib.check(_lib.load().calypso_mhd_initialize(self._h), "calypso_mhd_initialize")

    def step(self, nsteps=1):
        _lib.check(_lib.load().calypso_mhd_step(self._h, nsteps), "calypso_mhd_step")

    def set_option(self, name: str, value: int):
        """"fused_fft" (0/1), "legendre_autotune" (0/1) or "graph" (0/1: replay
        the captured Adams-Bashforth step as a HIP graph; on by default)."""
        _lib.check(_lib.load().calypso_mhd_set_option(self._h, name.encode(), int(value)),
                   "calypso_mhd_set_option")

    def pressure(self) -> np.ndarray:
        """The pressure of the current state as the reference computes it at
        an output step (pressure_4_sph_mhd, lead_fields_4_sph_mhd.f90:70-95),
        (nr, J) in the caller's rj order, zero outside the shell."""
        out = np.empty((self.nr, self.J))
        _lib.check(_lib.load().calypso_mhd_pressure(self._h, out.ctypes.data_as(_P)),
                   "calypso_mhd_pressure")
        return out

    ENERGY_FLUXES = ("Lorentz_work", "work_against_Lorentz", "magnetic_ene_generation",
                     "buoyancy_flux")

    def energy_fluxes(self, names=ENERGY_FLUXES) -> np.ndarray:
        """Spectra of the output-step energy fluxes named (enegy_fluxes_4_sph_mhd,
        lead_fields_4_sph_mhd.f90:98-121; the reference's field labels,
        m_phys_labels.f90:425-441), (len(names), nr, J) in the caller's rj order."""
        ids = np.array([self.ENERGY_FLUXES.index(n) for n in names], dtype=np.int32)
        out = np.empty((len(ids), self.nr, self.J))
        _lib.check(_lib.load().calypso_mhd_energy_fluxes(
            self._h, len(ids), ids.ctypes.data_as(_P), out.ctypes.data_as(_P)),
            "calypso_mhd_energy_fluxes")
        return out

    def energies(self) -> np.ndarray:
        out = np.zeros(7)
        _lib.check(_lib.load().calypso_mhd_energies(self._h, out.ctypes.data_as(_P)),
                   "calypso_mhd_energies")
        return out

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        _lib.check(_lib.load().calypso_mhd_time(self._h, ctypes.byref(t)), "calypso_mhd_time")
        return t.value

    def write_volume_ms(self, path):
        """Append this step's energies to sph_pwr_volume_s.dat (created with
        the reference's header if absent)."""
        _lib.check(_lib.load().calypso_mhd_write_volume_ms(self._h, str(path).encode()),
                   "calypso_mhd_write_volume_ms")

    def write_restart(self, prefix, istep_fld) -> str:
        """Write <prefix>.<istep_fld>.0.fst in the reference's ASCII restart
        format; returns the file name."""
        buf = ctypes.create_string_buffer(4096)
        _lib.check(_lib.load().calypso_mhd_write_restart(self._h, str(prefix).encode(),
                                                         int(istep_fld), buf, 4096),
                   "calypso_mhd_write_restart")
        return buf.value.decode()

    def read_restart(self, path):
        """Read a restart file; initialize() then resumes at its step and time."""
        _lib.check(_lib.load().calypso_mhd_read_restart(self._h, str(path).encode()),
                   "calypso_mhd_read_restart")

    @property
    def i_step(self) -> int:
        a = np.zeros(4, dtype=np.int64)
        _lib.check(_lib.load().calypso_mhd_info(self._h, a.ctypes.data_as(_P)), "calypso_mhd_info")
        return int(a[0])

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().calypso_mhd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def seeded_spectrum(seed, field, l_truncation, radius, amplitude=1.0, power=2.0,
                    envelope=True, l_min=0):
    """Counter-based seeded spectrum (nr, J) of calypso_seeded_spectrum."""
    radius = np.ascontiguousarray(radius, dtype=np.float64)
    J = (l_truncation + 1) ** 2
    out = np.zeros((radius.size, J))
    _lib.check(_lib.load().calypso_seeded_spectrum(
        int(seed), int(field), int(l_truncation), radius.size, radius.ctypes.data_as(_P),
        float(amplitude), float(power), 1 if envelope else 0, int(l_min),
        out.ctypes.data_as(_P)), "calypso_seeded_spectrum")
    return out


def coriolis_table(l_truncation: int):
    """Host-side coupling table of the Coriolis curl: (cor (12, J), idx (3, J))."""
    J = (l_truncation + 1) ** 2
    cor = np.zeros((12, J))
    idx = np.zeros((3, J), dtype=np.int32)
    _lib.check(_lib.load().calypso_mhd_coriolis_table(l_truncation, cor.ctypes.data_as(_P),
                                                      idx.ctypes.data_as(_P)),
               "calypso_mhd_coriolis_table")
    return cor, idx
