"""The device-resident step's fields against the reference's own program:
velocity, temperature and magnetic field on every radial point at 13 excited
modes after steps 10 and 20 of Dynamobench case 1 and case 2, as the
reference's MHD program (built from its sources) writes them through its
pickup monitor (tests/golden/picked_modes_case*.json, written by
tests/golden/make_picked_modes.py), within 1e-12 of each column's largest
value."""
import json
import os

import numpy as np
import pytest

from calypso_amd import mhd
from oracle import mhd_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [1, 2])
def test_fields_match_reference_pickup(torch_cuda, case):
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      f"picked_modes_case{case}.json")))
    print(f"[reference fields] Dynamobench case {case}", flush=True)
    if case == 1:
        p, r = mhd.dynamobench_case1()
        reft0 = mhd_step.MhdStep(mhd_step.DynamoBenchmarkCase1).reft0
    else:
        p, r = mhd.dynamobench_case2()
        reft0 = mhd_step.MhdStep(mhd_step.DynamoBenchmarkCase2).reft0
    sim = mhd.MhdSimulation(p, r)
    if case == 1:
        sim.set_dynamobench_ic(400)
    else:
        sim.set_dynamobench2_ic(400)
    sim.initialize()
    js = [int(lm.split(",")[0]) * (int(lm.split(",")[0]) + 1) + int(lm.split(",")[1])
          for lm in ref["modes"]]
    for step in (10, 20):
        sim.step(step - sim.i_step)
        t = sim.get_field(12)
        t[:, 0] = t[:, 0] + reft0
        cols = [sim.get_field(0), sim.get_field(2), sim.get_field(1), t,
                sim.get_field(6), sim.get_field(8), sim.get_field(7)]
        got = np.stack([np.stack([c[:, j] for c in cols], 1) for j in js])
        want = np.array([ref["steps"][str(step)][lm] for lm in ref["modes"]])
        errs = []
        for c, name in enumerate(ref["columns"]):
            scale = np.abs(want[:, :, c]).max()
            errs.append(np.abs(got[:, :, c] - want[:, :, c]).max() / scale)
        print(f"[reference fields] case {case} step {step}: "
              + " ".join(f"{n}={e:.1e}" for n, e in zip(ref["columns"], errs)), flush=True)
        assert max(errs) < 1e-12, dict(zip(ref["columns"], errs))
    sim.close()
