"""The device-resident MHD step with free-slip velocity walls (bc_velocity
free_slip_sph) against the oracle's restatement, for each choice of walls
(every stage field after the initialisation and three steps from a seeded
start), and the invariant that pins the walls physically: between two
free-slip walls a solid-body rotation (toroidal l = 1, m = 0 velocity
proportional to r^2) keeps its amplitude to rounding, while a rigid wall
brakes it; and Dynamobench case 1 with free-slip walls against the energies
the reference's own MHD program writes after every one of 20 steps."""
import json
import os

import numpy as np
import pytest

from calypso_amd import mhd
from oracle import mhd_step

from test_gpu_mhd import _compare

pytestmark = pytest.mark.gpu

SEEDED = (77, 0.3, 0.03, 0.1)
GRID = dict(l_truncation=15, nth=24, nphi=48, nri=24, dt=1.0e-4, coef_imp=0.5)
NAMES = {True: "free_slip_sph", False: "non_slip_sph"}


@pytest.mark.parametrize("walls", [(True, True), (True, False), (False, True)])
def test_free_slip_steps_match_oracle(torch_cuda, walls):
    print(f"[free slip] walls icb/cmb free = {walls}: init + 3 steps vs oracle", flush=True)
    p, r = mhd.dynamobench_case1(bc_velo=(NAMES[walls[0]], NAMES[walls[1]]), **GRID)
    sim = mhd.MhdSimulation(p, r)
    sim.set_seeded_ic(*SEEDED)
    sim.initialize()
    case = mhd_step.shell_case(L=15, nth=24, nphi=48, nri=24, dt=1.0e-4, coef_imp=0.5,
                               free_icb=walls[0], free_cmb=walls[1])
    ref = mhd_step.MhdStep(case)
    np.testing.assert_array_equal(ref.radius, r)
    ref.initialize(seeded=SEEDED)
    _compare(sim, ref, 1e-11, "init")
    for n in (1, 2, 3):
        sim.step(1)
        ref.step()
        _compare(sim, ref, 1e-10, f"step {n}")
    # the toroidal velocity at a free-slip wall is free, at a rigid one zero
    ut = sim.get_field(2)
    for wall, k in ((walls[0], 0), (walls[1], -1)):
        assert (np.abs(ut[k]).max() > 0.0) == wall
    np.testing.assert_allclose(sim.energies(), ref.energies(), rtol=1e-11)
    sim.close()


def _solid_body(walls, steps=20):
    p, r = mhd.dynamobench_case1(bc_velo=(NAMES[walls[0]], NAMES[walls[1]]), **GRID)
    sim = mhd.MhdSimulation(p, r)
    sim.set_seeded_ic(1, 0.0, 0.0, 0.0)
    ut = np.zeros((r.size, (p.l_truncation + 1) ** 2))
    ut[:, 2] = 0.1 * r * r                       # (l, m) = (1, 0)
    sim.set_field(2, ut)
    sim.initialize()
    sim.step(steps)
    got = sim.get_field(2)
    others = max(np.abs(sim.get_field(c)).max() for c in (0, 1))
    sim.close()
    return np.abs(got[:, 2] - ut[:, 2]).max() / np.abs(ut[:, 2]).max(), others


def test_free_slip_keeps_solid_body_rotation(torch_cuda):
    print("[free slip] solid-body rotation, 20 steps", flush=True)
    change, others = _solid_body((True, True))
    print(f"   free/free: change {change:.1e}, poloidal {others:.1e}", flush=True)
    assert change < 1e-12 and others < 1e-12
    change, _ = _solid_body((False, False))
    print(f"   rigid/rigid: change {change:.1e}", flush=True)
    assert change > 1e-2


@pytest.mark.parametrize("run,walls", [("free_free", (True, True)), ("free_rigid", (True, False))])
def test_free_slip_energies_match_reference_program(torch_cuda, run, walls):
    """Dynamobench case 1 with its velocity walls changed to free slip, on the
    GPU, against the energies of the reference's own MHD program (built from
    its sources) after every one of 20 steps
    (tests/golden/dynamobench_free_slip_steps.json): within 1e-12."""
    print(f"[free slip] Dynamobench case 1, {run}: 20 steps vs the reference program", flush=True)
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "dynamobench_free_slip_steps.json")))["runs"][run]
    p, r = mhd.dynamobench_case1(bc_velo=(NAMES[walls[0]], NAMES[walls[1]]))
    sim = mhd.MhdSimulation(p, r)
    sim.set_dynamobench_ic(400)
    sim.initialize()
    worst = 0.0
    for n in range(1, 21):
        sim.step(1)
        want = np.array(ref["steps"][str(n)])
        worst = max(worst, float(np.max(np.abs(sim.energies() / want - 1.0))))
    print(f"[free slip] {run}: worst relative energy difference {worst:.1e}", flush=True)
    assert worst < 1e-12, worst
    sim.close()
