"""Writes tests/golden/seeded_c1_reference_steps.json: the energies
(sph_pwr_volume_s.dat columns) after every one of 100 steps of the
reference's own MHD program (built from its sources by `make -C oracle grids
mhd`: oracle/_ref/gen_sph_grids, oracle/_ref/sph_mhd) started from the seeded
C1 fields (seed 2024, amplitudes 0.3 / 0.03 / 0.1 of u, B and the
temperature perturbation, (l+1)^-2 spectra; the oracle's seeded generator)
on tests/Dynamobench_case1's shell and physics with dt = 1e-4 and
Crank-Nicolson weight 0.5.

The seeded state goes to the reference program as its own restart file
(rst/rst.0.0.fst, step 0, time 0: velocity, full temperature, magnetic field,
zero previous terms; vectors as (pol, tor, dpol/dr), modes in the generator's
rj order), written by the
library's host writer, which is byte-identical to the reference's
write_step_data / write_field_data.  The control file is the reference's
with one subdomain and one thread, the grid and restart paths, rst_ctl
start_from_rst_file, dt, the implicit weights and an energy row every step.

Run in the build container:  make -C oracle grids mhd && python tests/golden/make_seeded_reference.py
"""
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from calypso_amd import _lib, grids       # noqa: E402
from oracle import mhd_step               # noqa: E402

REF_TEST = "/root/reference/tests/Dynamobench_case1"
GEN = os.path.join(ROOT, "oracle", "_ref", "gen_sph_grids")
MHD = os.path.join(ROOT, "oracle", "_ref", "sph_mhd")
OUT = os.path.join(HERE, "seeded_c1_reference_steps.json")
SEEDED = (2024, 0.3, 0.03, 0.1)
DT, COEF_IMP, NSTEP = 1.0e-4, 0.5, 100


def one_domain(text):
    text = re.sub(r"num_subdomain_ctl\s+\d+", "num_subdomain_ctl    1", text)
    text = re.sub(r"num_smp_ctl\s+\d+", "num_smp_ctl    1", text)
    text = re.sub(r"num_radial_domain_ctl\s+\d+", "num_radial_domain_ctl    1", text)
    text = re.sub(r"num_horizontal_domain_ctl\s+\d+", "num_horizontal_domain_ctl    1", text)
    return text.replace("'sph_lm31r48c_4/in'", "'grid/in'")


def seeded_restart(path):
    """The seeded start (before SPH_initialize_MHD's derived fields) as the
    reference's restart file."""
    s = mhd_step.MhdStep(mhd_step.shell_case(dt=DT, coef_imp=COEF_IMP))
    start = {}
    orig = s._start
    s._start = lambda: start.update({k: np.array(v, copy=True) for k, v in s.f.items()})
    s.initialize(seeded=SEEDED)
    s._start = orig
    L, nr = s.L, s.nr
    J = (L + 1) ** 2
    jg = grids.rj_mode_order(L)
    nnod = nr * J

    def local(a):            # (nr, J) in l(l+1)+m order -> the generator's rj order
        return np.ascontiguousarray(a[:, jg]).reshape(nnod)

    zero = np.zeros(nnod)
    # a vector in the file is (pol, tor, dpol/dr): copy_each_sph_vector_to_IO
    cols = [local(start["velo"][0]), local(start["velo"][2]), local(start["velo"][1]),
            local(start["temp"]),
            local(start["magne"][0]), local(start["magne"][2]), local(start["magne"][1]),
            zero, zero, zero, zero, zero, zero, zero]
    data = np.ascontiguousarray(np.stack(cols, 1).T).reshape(-1)   # component c at c*nnod
    names = [b"velocity", b"temperature", b"magnetic_field", b"previous_momentum",
             b"previous_induction", b"previous_heat"]
    ncomp = np.array([3, 1, 3, 3, 3, 1], dtype=np.int32)
    arr = (ctypes.c_char_p * len(names))(*names)
    _lib.check(_lib.load().calypso_sph_write_field_file(
        path.encode(), 0, 0, 0.0, DT, nnod, len(names), ncomp.ctypes.data_as(ctypes.c_void_p),
        arr, data.ctypes.data_as(ctypes.c_void_p)), "calypso_sph_write_field_file")


def main():
    shell = one_domain(open(os.path.join(REF_TEST, "control_sph_shell")).read())
    mhd = one_domain(open(os.path.join(REF_TEST, "control_MHD")).read())
    mhd = mhd.replace("'rst_4/rst'", "'rst/rst'")
    mhd = re.sub(r"i_step_check_ctl\s+\d+", "i_step_check_ctl    1", mhd)
    mhd = re.sub(r"i_step_finish_ctl\s+\d+", f"i_step_finish_ctl    {NSTEP}", mhd)
    mhd = re.sub(r"dt_ctl\s+\S+", f"dt_ctl    {DT:.1e}", mhd)
    mhd = re.sub(r"(coef_imp_[vtbd]_ctl)\s+\S+", rf"\1    {COEF_IMP:.1e}", mhd)
    mhd = re.sub(r"rst_ctl\s+dynamo_benchmark_1", "rst_ctl    start_from_rst_file", mhd)
    with tempfile.TemporaryDirectory() as tmp:
        for d in ("grid", "rst", "field"):
            os.makedirs(os.path.join(tmp, d))
        open(os.path.join(tmp, "control_sph_shell"), "w").write(shell)
        subprocess.run([GEN], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
        seeded_restart(os.path.join(tmp, "rst", "rst.0.0.fst"))
        open(os.path.join(tmp, "control_MHD"), "w").write(mhd)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([MHD], cwd=tmp, check=True, stdout=subprocess.DEVNULL, env=env)
        rows = open(os.path.join(tmp, "sph_pwr_volume_s.dat")).read().splitlines()
    steps = {}
    for line in rows[8:]:
        v = line.split()
        steps[v[0]] = [float(x) for x in v[2:]]
    json.dump({"source": "the reference's sph_mhd built from its sources, started from the "
                         "seeded C1 fields written as its restart file; sph_pwr_volume_s.dat "
                         "after every step",
               "seeded": {"seed": SEEDED[0], "amp_u": SEEDED[1], "amp_b": SEEDED[2],
                          "amp_t": SEEDED[3]},
               "dt": DT, "coef_imp": COEF_IMP, "steps": steps}, open(OUT, "w"), indent=0)
    print("wrote", OUT, len(steps), "rows")


if __name__ == "__main__":
    main()
