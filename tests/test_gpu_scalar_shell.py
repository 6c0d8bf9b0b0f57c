"""The device-resident scalar-shell step (scalars diffusing in a shell at
rest, per-mode boundary values and fluxes) against the oracle's restatement
(oracle/scalar_shell.py: every field after the initialisation and after a
few steps, for every combination of wall conditions) and against the
reference's own tests/heterogineous_temp energies (sph_pwr_volume_s.dat,
steps 50 and 100, 1e-12 relative; the reference's check,
compare_sph_mean_square.f90:96-122, allows 1e-9)."""
import os

import numpy as np
import pytest

from calypso_amd import _lib
from calypso_amd import scalar_shell as product
from oracle import scalar_shell as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "heterogeneous_temp")
BC_FILE = os.path.join(GOLD, "bc_spectr.btx")
FIELD_TOL = 1e-13


def golden_rows():
    rows = {}
    for line in open(os.path.join(GOLD, "sph_pwr_volume_s.dat")):
        t = line.split()
        if len(t) == 4 and t[0].isdigit():
            rows[int(t[0])] = np.array([float(x) for x in t[1:]])
    return rows


def rel_err(got, want):
    return np.abs(got - want).max() / max(np.abs(want).max(), 1e-300)


def compare_fields(sim, ref, label):
    errs = []
    for q in range(sim.nsc):
        for which, arr in ((0, ref.x[q]), (1, ref.diff[q]), (2, ref.adv[q])):
            errs.append(rel_err(sim.get_field(q, which), arr))
    print(label, " ".join(f"{e:.1e}" for e in errs), flush=True)
    assert max(errs) < FIELD_TOL, f"{label}: field error {max(errs):.2e}"


def test_heterogeneous_temp_golden(torch_cuda, tmp_path):
    print("[scalar_shell] heterogineous_temp, 100 steps", flush=True)
    radius, dt, scalars = product.heterogeneous_temp_params(BC_FILE)
    sim = product.ScalarShell(4, radius, dt, scalars)
    ref = orc.heterogeneous_temp_case(BC_FILE)
    np.testing.assert_array_equal(radius, ref.radius)
    sim.set_noise()
    ref.set_noise()
    sim.initialize()
    ref.initialize()
    compare_fields(sim, ref, "init")
    gold = golden_rows()
    out = tmp_path / "sph_pwr_volume_s.dat"
    for n in range(1, 101):
        sim.step(1)
        ref.step()
        if n in (1, 2, 3):
            compare_fields(sim, ref, f"step {n}")
        if n in gold:
            got = sim.mean_squares()
            print(f"step {n}: {got} golden {gold[n][1:]}", flush=True)
            np.testing.assert_allclose(got, gold[n][1:], rtol=1e-12, atol=0)
            np.testing.assert_allclose(got, ref.mean_squares(), rtol=1e-13, atol=0)
            sim.write_volume_ms(out)
    compare_fields(sim, ref, "step 100")
    # the file: the reference's header, then one row per output step
    mine = out.read_text().splitlines()
    theirs = open(os.path.join(GOLD, "sph_pwr_volume_s.dat")).read().splitlines()
    assert len(mine) == len(theirs) == 10
    assert [l.rstrip() for l in mine[:8]] == [l.rstrip() for l in theirs[:8]]
    for a, b in zip(mine[8:], theirs[8:]):
        ta, tb = a.split(), b.split()
        assert ta[0] == tb[0]
        np.testing.assert_allclose([float(x) for x in ta[1:]], [float(x) for x in tb[1:]],
                                   rtol=1e-12, atol=0)
    sim.close()


@pytest.mark.parametrize("kinds", [(orc.FIXED, orc.FIXED), (orc.FLUX, orc.FLUX),
                                   (orc.FLUX, orc.FIXED), (orc.FIXED, orc.FLUX)])
def test_wall_conditions_vs_oracle(torch_cuda, kinds):
    """Seeded fields and seeded per-mode wall data (L = 7, 17 points), three
    scalars with different diffusivities and weights, for each pair of wall
    conditions."""
    print(f"[scalar_shell] walls icb/cmb = {kinds}", flush=True)
    L, nri = 7, 16
    J = (L + 1) ** 2
    rng = np.random.default_rng(11 + 3 * kinds[0] + kinds[1])
    r_cmb = 1.0 / (1.0 - 0.35)
    radius = orc.chebyshev_radius(nri, 0.35 * r_cmb, r_cmb)
    specs = []
    for q, (cd, ci) in enumerate([(1.0, 0.5), (0.5, 0.6), (2.0, 1.0)]):
        vi, vc = rng.standard_normal(J) * 0.1, rng.standard_normal(J) * 0.1
        l_of_j = np.floor(np.sqrt(np.arange(J))).astype(int)
        m_of_j = np.arange(J) - l_of_j * (l_of_j + 1)
        rows_i = [(int(l), int(m), float(v)) for l, m, v in zip(l_of_j, m_of_j, vi)]
        rows_c = [(int(l), int(m), float(v)) for l, m, v in zip(l_of_j, m_of_j, vc)]
        specs.append(dict(name=f"s{q}", coef_d=cd, coef_imp=ci,
                          icb=(kinds[0], rows_i), cmb=(kinds[1], rows_c),
                          vi=vi, vc=vc))
    ref = orc.ScalarShell(L, nri, 0.35 * r_cmb, r_cmb, 2.0e-3,
                          [dict(coef_d=s["coef_d"], coef_imp=s["coef_imp"], icb=s["icb"],
                                cmb=s["cmb"]) for s in specs])
    sim = product.ScalarShell(L, radius, 2.0e-3,
                              [dict(name=s["name"], coef_d=s["coef_d"], coef_imp=s["coef_imp"],
                                    icb=(kinds[0], s["vi"]), cmb=(kinds[1], s["vc"]))
                               for s in specs])
    np.testing.assert_array_equal(ref.radius, radius)
    for q in range(3):
        x0 = rng.standard_normal((radius.size, J)) * 0.01
        ref.x[q][:] = x0
        sim.set_field(q, x0)
    ref.initialize()
    sim.initialize()
    compare_fields(sim, ref, "init")
    for n in range(1, 6):
        sim.step(1)
        ref.step()
        compare_fields(sim, ref, f"step {n}")
    np.testing.assert_allclose(sim.mean_squares(), ref.mean_squares(), rtol=1e-13, atol=0)
    sim.close()


def test_rejects_bad_arguments(torch_cuda):
    radius, dt, scalars = product.heterogeneous_temp_params(BC_FILE)
    bad = [dict(s) for s in scalars]
    bad[0]["icb"] = (7, bad[0]["icb"][1])
    with pytest.raises(_lib.CalypsoError, match="boundary kind"):
        product.ScalarShell(4, radius, dt, bad)
    with pytest.raises(_lib.CalypsoError, match="increasing"):
        product.ScalarShell(4, radius[::-1].copy(), dt, scalars)
    with pytest.raises(_lib.CalypsoError, match="1..8 scalars"):
        product.ScalarShell(4, radius, dt, scalars * 5)
    sim = product.ScalarShell(4, radius, dt, scalars)
    with pytest.raises(_lib.CalypsoError, match="out of range"):
        sim.set_field(2, np.zeros((radius.size, 25)))
    sim.close()


def test_fortran_driver_heterogeneous_temp(torch_cuda, tmp_path):
    """The Fortran face (bind(C) interfaces of calypso_amd_capi) runs the
    same 100 steps and writes the same energy file as the reference's test."""
    import subprocess
    exe = os.path.join(_lib.LIB_DIR, "scalar_shell_driver")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no amdflang)")
    print("[scalar_shell] Fortran driver, 100 steps", flush=True)
    radius, dt, scalars = product.heterogeneous_temp_params(BC_FILE)
    setup = tmp_path / "setup.txt"
    with open(setup, "w") as f:
        f.write(f"4 {radius.size - 1} {dt!r} {len(scalars)}\n")
        for s in scalars:
            f.write(f"{s['coef_d']!r} {s['coef_imp']!r} {s['icb'][0]} {s['cmb'][0]}\n")
            f.write(" ".join(repr(float(v)) for v in s["icb"][1]) + "\n")
            f.write(" ".join(repr(float(v)) for v in s["cmb"][1]) + "\n")
    efile = tmp_path / "sph_pwr_volume_s.dat"
    out = subprocess.run([exe, str(setup), str(efile), "100", "50"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    mine = efile.read_text().splitlines()
    theirs = open(os.path.join(GOLD, "sph_pwr_volume_s.dat")).read().splitlines()
    assert len(mine) == len(theirs) == 10
    assert [l.rstrip() for l in mine[:8]] == [l.rstrip() for l in theirs[:8]]
    for a, b in zip(mine[8:], theirs[8:]):
        ta, tb = a.split(), b.split()
        assert ta[0] == tb[0]
        np.testing.assert_allclose([float(x) for x in ta[1:]], [float(x) for x in tb[1:]],
                                   rtol=1e-12, atol=0)
