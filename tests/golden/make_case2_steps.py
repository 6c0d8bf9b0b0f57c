"""Writes tests/golden/dynamobench_case2_steps.json: the energies
(sph_pwr_volume_s.dat columns) after every one of the 100 steps of the
reference's tests/Dynamobench_case2, and the radial grid it runs on, from the
reference's own programs built from its sources: the grid generator
(`make -C oracle grids` -> oracle/_ref/gen_sph_grids) and the MHD program
(`make -C oracle mhd` -> oracle/_ref/sph_mhd).

The reference's control files are used as they are, except: one subdomain and
one thread instead of four (its published file comes from four), the grid
and restart paths, and an energy row every step instead of every 50.  Rows 50
and 100 agree with the reference's published
tests/Dynamobench_case2/reference/sph_pwr_volume_s.dat to ~1e-14.

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_case2_steps.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_TEST = "/root/reference/tests/Dynamobench_case2"
GEN = os.path.join(ROOT, "oracle", "_ref", "gen_sph_grids")
MHD = os.path.join(ROOT, "oracle", "_ref", "sph_mhd")
OUT = os.path.join(HERE, "dynamobench_case2_steps.json")


def one_domain(text):
    text = re.sub(r"num_subdomain_ctl\s+\d+", "num_subdomain_ctl    1", text)
    text = re.sub(r"num_smp_ctl\s+\d+", "num_smp_ctl    1", text)
    text = re.sub(r"num_radial_domain_ctl\s+\d+", "num_radial_domain_ctl    1", text)
    text = re.sub(r"num_horizontal_domain_ctl\s+\d+", "num_horizontal_domain_ctl    1", text)
    return text.replace("'sph_lm31r48c_ic_4/in'", "'grid/in'")


def main():
    shell = one_domain(open(os.path.join(REF_TEST, "control_sph_shell")).read())
    mhd = one_domain(open(os.path.join(REF_TEST, "control_MHD")).read())
    mhd = mhd.replace("'rst_4/rst'", "'rst/rst'")
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    1", mhd)
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("grid", "rst", "field"):
            os.makedirs(os.path.join(tmp, d))
        open(os.path.join(tmp, "control_sph_shell"), "w").write(shell)
        subprocess.run([GEN], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        open(os.path.join(tmp, "control_MHD"), "w").write(mhd)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MHD], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        rows = open(os.path.join(tmp, "sph_pwr_volume_s.dat")).read().splitlines()
        rj = open(os.path.join(tmp, "grid", "in.0.rj")).read().splitlines()
    steps, times = {}, {}
    for line in rows[8:]:
        v = line.split()
        steps[v[0]] = [float(x) for x in v[2:]]
        times[v[0]] = float(v[1])
    # the radial points: "n 1 n" then n lines "k r"
    i = next(k for k, l in enumerate(rj) if l.strip().startswith("! r-direction"))
    i += 2
    n = int(rj[i].split()[0])
    radius = [float(rj[i + 1 + k].split()[1]) for k in range(n)]
    json.dump({"source": "the reference's sph_mhd and gen_sph_grids built from its sources "
                         "(oracle/build_ref_mhd.py, oracle/build_ref_grids.py), run on "
                         "tests/Dynamobench_case2's control files with one subdomain",
               "columns": ["K_ene_pol", "K_ene_tor", "K_ene", "temperature",
                           "M_ene_pol", "M_ene_tor", "M_ene"],
               "icb_id": 25, "cmb_id": 73, "radius": radius,
               "steps": steps, "time": times, "rtol": 1e-12},
              open(OUT, "w"), indent=0)
    print("wrote", OUT, len(steps), "steps")


if __name__ == "__main__":
    main()
