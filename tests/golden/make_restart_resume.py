"""Writes tests/golden/restart_resume_l7/: a restart file written by the
reference's own MHD program (built from its sources by `make -C oracle grids
mhd`) after step 10 of tests/Dynamobench_case1's physics on a small grid
(L = 7, 12 x 24, 12 Chebyshev intervals) and the one after step 20, gzipped,
and the energies
(sph_pwr_volume_s.dat columns) after each of the run's 20 steps.  A run of
the library read from the first file must continue the reference's run and
write the second.

The control files are the reference's, except: one subdomain and one thread,
the grid (truncation, points, intervals) and restart paths, an energy row
every step, a restart every 10 steps, 20 steps.

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_restart_resume.py
"""
import gzip
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
OUT = os.path.join(HERE, "restart_resume_l7")
L, NTH, NPHI, NRI = 7, 12, 24, 12


def one_domain(text):
    text = re.sub(r"num_subdomain_ctl\s+\d+", "num_subdomain_ctl    1", text)
    text = re.sub(r"num_smp_ctl\s+\d+", "num_smp_ctl    1", text)
    text = re.sub(r"num_radial_domain_ctl\s+\d+", "num_radial_domain_ctl    1", text)
    text = re.sub(r"num_horizontal_domain_ctl\s+\d+", "num_horizontal_domain_ctl    1", text)
    return text.replace("'sph_lm31r48c_4/in'", "'grid/in'")


def main():
    shell = one_domain(open(os.path.join(REF_TEST, "control_sph_shell")).read())
    shell = re.sub(r"truncation_level_ctl\s+\d+", f"truncation_level_ctl  {L}", shell)
    shell = re.sub(r"ngrid_meridonal_ctl\s+\d+", f"ngrid_meridonal_ctl  {NTH}", shell)
    shell = re.sub(r"ngrid_zonal_ctl\s+\d+", f"ngrid_zonal_ctl  {NPHI}", shell)
    shell = re.sub(r"num_fluid_grid_ctl\s+\d+", f"num_fluid_grid_ctl  {NRI}", shell)
    mhd = one_domain(open(os.path.join(REF_TEST, "control_MHD")).read())
    mhd = mhd.replace("'rst_4/rst'", "'rst/rst'")
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    1", mhd)
    mhd = re.sub(r"i_step_rst_ctl\s+\d+", "i_step_rst_ctl    10", mhd)
    mhd = re.sub(r"i_step_finish_ctl\s+\d+", "i_step_finish_ctl    20", mhd)
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("grid", "rst", "field"):
            os.makedirs(os.path.join(tmp, d))
        open(os.path.join(tmp, "control_sph_shell"), "w").write(shell)
        subprocess.run([GEN], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        open(os.path.join(tmp, "control_MHD"), "w").write(mhd)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MHD], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        rows = open(os.path.join(tmp, "sph_pwr_volume_s.dat")).read().splitlines()
        print(sorted(os.listdir(os.path.join(tmp, "rst"))))
        rst = open(os.path.join(tmp, "rst", "rst.1.0.fst"), "rb").read()
        rst2 = open(os.path.join(tmp, "rst", "rst.2.0.fst"), "rb").read()
    os.makedirs(OUT, exist_ok=True)
    for name, data in (("rst.1.0.fst.gz", rst), ("rst.2.0.fst.gz", rst2)):
        with gzip.GzipFile(os.path.join(OUT, name), "wb", mtime=0) as g:
            g.write(data)
    steps = {}
    for line in rows[8:]:
        v = line.split()
        steps[v[0]] = [float(x) for x in v[2:]]
    json.dump({"source": "the reference's sph_mhd built from its sources on "
                         "tests/Dynamobench_case1's physics at L=7, 12 x 24, 12 intervals; "
                         "restart after step 10, sph_pwr_volume_s.dat after every step",
               "grid": {"l_truncation": L, "nth": NTH, "nphi": NPHI, "nri": NRI},
               "restart": "rst.1.0.fst.gz", "restart_step": 10,
               "restart_end": "rst.2.0.fst.gz", "steps": steps},
              open(os.path.join(OUT, "energies.json"), "w"), indent=0)
    print("wrote", OUT, len(steps), "rows", len(rst), "bytes of restart")


if __name__ == "__main__":
    main()
