"""Writes tests/golden/energy_fluxes_case1.json: spectral coefficients of the
output-step energy fluxes (enegy_fluxes_4_sph_mhd, lead_fields_4_sph_mhd.f90:
98-121) at chosen (l, m) modes on every radial point after steps 10 and 20 of
the reference's tests/Dynamobench_case1, as the reference's own MHD program
writes them through its pickup monitor (pickup_sph_spectr.f90).  The program is
the reference's grid generator and MHD program built from its sources
(`make -C oracle grids mhd`).  The control files are the test's own, with one
subdomain and one thread, an output row every 10 steps, 20 steps, and a field
list that adds the magnetic induction (the backward-transformed input of the
magnetic energy generation) and four monitored fluxes: Lorentz_work,
work_against_Lorentz, magnetic_ene_generation and buoyancy_flux
(m_phys_labels.f90:425-441).

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_energy_fluxes.py
"""
import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_picked_modes import CASES, GEN, MHD, MODES, monitor_block, one_domain  # noqa: E402

FIELDS = [("velocity", "Monitor_On"), ("temperature", "Monitor_On"),
          ("pressure", "Monitor_Off"), ("vorticity", "Monitor_Off"),
          ("magnetic_field", "Monitor_On"), ("current_density", "Monitor_Off"),
          ("magnetic_induction", "Monitor_Off"),
          ("Lorentz_work", "Monitor_On"), ("work_against_Lorentz", "Monitor_On"),
          ("magnetic_ene_generation", "Monitor_On"), ("buoyancy_flux", "Monitor_On")]
FLUXES = ["Lorentz_work", "work_against_Lorentz", "magnetic_ene_generation", "buoyancy_flux"]


def field_block():
    lines = [f"      array nod_value_ctl      {len(FIELDS)}"]
    lines += [f"        nod_value_ctl  {name:24s}Viz_Off  {mon}" for name, mon in FIELDS]
    lines += ["      end array nod_value_ctl"]
    return "\n".join(lines)


def run_case(case):
    test_dir, grid = CASES[case]
    shell = one_domain(open(os.path.join(test_dir, "control_sph_shell")).read(), grid)
    mhd = one_domain(open(os.path.join(test_dir, "control_MHD")).read(), grid)
    mhd = re.sub(r"'rst_4/rst'", "'rst/rst'", mhd)
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    10", mhd)
    mhd = re.sub(r"i_step_finish_ctl\s+\d+", "i_step_finish_ctl    20", mhd)
    mhd, n = re.subn(r"(?s)[ \t]*array nod_value_ctl\s+\d+.*?end array nod_value_ctl",
                     field_block(), mhd)
    assert n == 1, "field list not found in the control file"
    mhd = re.sub(r"(?s)\s*begin sph_monitor_ctl.*?end sph_monitor_ctl\s*", "\n", mhd)
    mhd = mhd.replace("end MHD_control", monitor_block() + "end MHD_control")
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("grid", "rst", "field", "sph_spectr"):
            os.makedirs(os.path.join(tmp, d))
        open(os.path.join(tmp, "control_sph_shell"), "w").write(shell)
        subprocess.run([GEN], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        open(os.path.join(tmp, "control_MHD"), "w").write(mhd)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MHD], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        rows = open(os.path.join(tmp, "sph_spectr", "picked_mode.dat")).read().splitlines()
    head = next(i for i, l in enumerate(rows) if l.startswith("t_step"))
    names = rows[head].split()[6:]
    keep = [names.index(f) for f in FLUXES]
    out = {}
    for line in rows[head + 1:]:
        v = line.split()
        step, k, l, m = v[0], int(v[2]), int(v[4]), int(v[5])
        vals = [float(x) for x in v[6:]]
        rec = out.setdefault(step, {}).setdefault(f"{l},{m}", {})
        rec[k] = [vals[i] for i in keep]
    data = {step: {lm: [vals[k] for k in sorted(vals)] for lm, vals in modes.items()}
            for step, modes in out.items()}
    return data


def main():
    data = run_case(1)
    path = os.path.join(HERE, "energy_fluxes_case1.json")
    json.dump({"source": "the reference's sph_mhd built from its sources, run on "
                         "tests/Dynamobench_case1's control files with one subdomain and the "
                         "energy fluxes monitored; pickup monitor on every radial point",
               "columns": FLUXES, "modes": [f"{l},{m}" for l, m in MODES],
               "steps": data}, open(path, "w"), separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
