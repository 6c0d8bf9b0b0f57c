"""Writes tests/golden/dynamobench_free_slip_steps.json: the energies
(sph_pwr_volume_s.dat columns) after every one of 20 steps of the reference's
tests/Dynamobench_case1 with its velocity walls changed to free slip, at both
walls ("free_free") and at the ICB only ("free_rigid"), from the reference's
own programs built from its sources (`make -C oracle grids mhd`:
oracle/_ref/gen_sph_grids, oracle/_ref/sph_mhd).

The control files are the reference's, except: one subdomain and one thread,
the grid and restart paths, an energy row every step, 20 steps, and the
bc_velocity kinds.

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_free_slip_steps.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_TEST = "/root/reference/tests/Dynamobench_case1"
GEN = os.path.join(ROOT, "oracle", "_ref", "gen_sph_grids")
MHD = os.path.join(ROOT, "oracle", "_ref", "sph_mhd")
OUT = os.path.join(HERE, "dynamobench_free_slip_steps.json")
NSTEP = 20
WALLS = {"free_free": ("free_slip_sph", "free_slip_sph"),
         "free_rigid": ("free_slip_sph", "non_slip_sph")}


def one_domain(text):
    text = re.sub(r"num_subdomain_ctl\s+\d+", "num_subdomain_ctl    1", text)
    text = re.sub(r"num_smp_ctl\s+\d+", "num_smp_ctl    1", text)
    text = re.sub(r"num_radial_domain_ctl\s+\d+", "num_radial_domain_ctl    1", text)
    text = re.sub(r"num_horizontal_domain_ctl\s+\d+", "num_horizontal_domain_ctl    1", text)
    return text.replace("'sph_lm31r48c_4/in'", "'grid/in'")


def run(icb, cmb):
    shell = one_domain(open(os.path.join(REF_TEST, "control_sph_shell")).read())
    mhd = one_domain(open(os.path.join(REF_TEST, "control_MHD")).read())
    mhd = mhd.replace("'rst_4/rst'", "'rst/rst'")
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    1", mhd)
    mhd = re.sub(r"i_step_finish_ctl\s+\d+", f"i_step_finish_ctl    {NSTEP}", mhd)
    mhd = re.sub(r"bc_velocity\s+\w+\s+ICB", f"bc_velocity  {icb}  ICB", mhd)
    mhd = re.sub(r"bc_velocity\s+\w+\s+CMB", f"bc_velocity  {cmb}  CMB", mhd)
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("grid", "rst", "field"):
            os.makedirs(os.path.join(tmp, d))
        open(os.path.join(tmp, "control_sph_shell"), "w").write(shell)
        subprocess.run([GEN], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        open(os.path.join(tmp, "control_MHD"), "w").write(mhd)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MHD], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        rows = open(os.path.join(tmp, "sph_pwr_volume_s.dat")).read().splitlines()
    steps = {}
    for line in rows[8:]:
        v = line.split()
        steps[v[0]] = [float(x) for x in v[2:]]
    return steps


def main():
    out = {"source": "the reference's sph_mhd built from its sources, run on "
                     "tests/Dynamobench_case1's control files with one subdomain and the "
                     "velocity walls changed; sph_pwr_volume_s.dat after every step",
           "columns": ["K_ene_pol", "K_ene_tor", "K_ene", "temperature", "M_ene_pol",
                       "M_ene_tor", "M_ene"],
           "runs": {}}
    for name, (icb, cmb) in WALLS.items():
        out["runs"][name] = {"bc_velocity": [icb, cmb], "steps": run(icb, cmb)}
        print(name, len(out["runs"][name]["steps"]), "rows")
    json.dump(out, open(OUT, "w"), indent=0)


if __name__ == "__main__":
    main()
