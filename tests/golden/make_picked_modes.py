"""Writes tests/golden/picked_modes_case{1,2}.json: spectral coefficients of
velocity, temperature, magnetic field and pressure at chosen (l, m) modes on
every radial point after steps 10 and 20 of the reference's tests/Dynamobench_case1
and tests/Dynamobench_case2, as the reference's own MHD program writes them
through its pickup monitor (picked_sph_prefix, pickup_sph_spectr.f90).  The
programs are the reference's grid generator and MHD program built from its
sources (`make -C oracle grids mhd`).  The reference's control files are used
as they are except for one subdomain and one thread, the grid and restart
paths, an output row every 10 steps, 20 steps, the pressure monitored
(Monitor_On: it is solved at every output step either way) and the pickup
block.  (1, 1)
stays at round-off in these runs and pins the inner core's x/y turns of case 2
only through their absence.

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_picked_modes.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(ROOT, "oracle", "_ref", "gen_sph_grids")
MHD = os.path.join(ROOT, "oracle", "_ref", "sph_mhd")
# modes the benchmark runs excite (m = 0, +-4, +-8, ...: the initial (4, 4)
# temperature and the axisymmetric field), down to 1e-7 of the largest
MODES = [(0, 0), (1, 0), (1, 1), (2, 0), (3, 0), (4, 4), (4, -4), (5, 4), (6, -4), (8, 8), (9, 0),
         (10, -8), (12, 12)]
CASES = {1: ("/root/reference/tests/Dynamobench_case1", "'sph_lm31r48c_4/in'"),
         2: ("/root/reference/tests/Dynamobench_case2", "'sph_lm31r48c_ic_4/in'")}


def one_domain(text, grid):
    text = re.sub(r"num_subdomain_ctl\s+\d+", "num_subdomain_ctl    1", text)
    text = re.sub(r"num_smp_ctl\s+\d+", "num_smp_ctl    1", text)
    text = re.sub(r"num_radial_domain_ctl\s+\d+", "num_radial_domain_ctl    1", text)
    text = re.sub(r"num_horizontal_domain_ctl\s+\d+", "num_horizontal_domain_ctl    1", text)
    return text.replace(grid, "'grid/in'")


def monitor_block():
    lines = ["  begin sph_monitor_ctl",
             "    volume_pwr_spectr_prefix     'sph_pwr_volume'",
             "    picked_sph_prefix            'sph_spectr/picked_mode'",
             "    array pick_layer_ctl  0", "    end array",
             f"    array pick_sph_spectr_ctl  {len(MODES)}"]
    lines += [f"      pick_sph_spectr_ctl   {l}   {m}" for l, m in MODES]
    lines += ["    end array pick_sph_spectr_ctl", "  end sph_monitor_ctl"]
    return "\n".join(lines) + "\n"


def run_case(case):
    test_dir, grid = CASES[case]
    shell = one_domain(open(os.path.join(test_dir, "control_sph_shell")).read(), grid)
    mhd = one_domain(open(os.path.join(test_dir, "control_MHD")).read(), grid)
    mhd = re.sub(r"'rst_4/rst'", "'rst/rst'", mhd)
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    10", mhd)
    mhd = re.sub(r"i_step_finish_ctl\s+\d+", "i_step_finish_ctl    20", mhd)
    mhd = re.sub(r"nod_value_ctl\s+pressure\s+\w+\s+Monitor_Off",
                 "nod_value_ctl  pressure                Viz_Off  Monitor_On", mhd)
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
    out = {}
    for line in rows[head + 1:]:
        v = line.split()
        step, k, l, m = v[0], int(v[2]), int(v[4]), int(v[5])
        if k == 0:
            continue      # the centre node of case 2 (degree 0 at r = 0), not kept here
        rec = out.setdefault(step, {}).setdefault(f"{l},{m}", {})
        rec[k] = [float(x) for x in v[6:]]
    data = {}
    for step, modes in out.items():
        data[step] = {lm: [vals[k] for k in sorted(vals)] for lm, vals in modes.items()}
    return names, data


def main():
    for case in CASES:
        names, data = run_case(case)
        path = os.path.join(HERE, f"picked_modes_case{case}.json")
        json.dump({"source": f"the reference's sph_mhd built from its sources, run on "
                             f"tests/Dynamobench_case{case}'s control files with one subdomain; "
                             "pickup monitor (pickup_sph_spectr.f90) on every radial point",
                   "columns": names, "modes": [f"{l},{m}" for l, m in MODES],
                   "steps": data}, open(path, "w"), separators=(",", ":"))
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
