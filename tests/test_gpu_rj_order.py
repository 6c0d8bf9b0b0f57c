"""The performance boundary in the reference's own rj mode order: a caller
that hands over d_rj in the one-domain order of the reference's grid files
(orders grouped, calypso_sph_rj_mode_order, pinned against gen_sph_grids by
tests/test_sph_grids.py) gets every mode back where it put it, through the
field copies and the restart files."""
import numpy as np
import pytest

from calypso_amd import grids, mhd

pytestmark = pytest.mark.gpu
SEEDED = (2024, 0.3, 0.03, 0.1)


def test_fields_in_reference_mode_order(torch_cuda, tmp_path):
    print("[rj order] L=31, seeded start, 5 steps, default vs reference order", flush=True)
    p, r = mhd.dynamobench_case1(dt=1.0e-4, coef_imp=0.5)
    order = grids.rj_mode_order(p.l_truncation)
    assert not np.array_equal(order, np.arange(order.size))
    a = mhd.MhdSimulation(p, r)
    b = mhd.MhdSimulation(p, r)
    b.set_rj_order(order)
    for sim in (a, b):
        sim.set_seeded_ic(*SEEDED)
        sim.initialize()
        sim.step(5)
    for comp in (0, 1, 2, 6, 8, 12, 20, 26):
        fa, fb = a.get_field(comp), b.get_field(comp)
        assert np.array_equal(fb, fa[:, order]), comp
    np.testing.assert_array_equal(a.energies(), b.energies())
    # a field given in the reference order lands on the same modes
    u = a.get_field(0) * 1.5
    a.set_field(0, u)
    b.set_field(0, u[:, order])
    a.step(2)
    b.step(2)
    assert np.array_equal(b.get_field(0), a.get_field(0)[:, order])
    # restart files follow the caller's order and resume the caller's state
    name_b = b.write_restart(str(tmp_path / "rb"), 1)
    c = mhd.MhdSimulation(p, r)
    c.set_rj_order(order)
    c.read_restart(name_b)
    c.initialize()
    b.step(3)
    c.step(3)
    for comp in (0, 2, 6, 8, 12):
        x, y = b.get_field(comp), c.get_field(comp)
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(x).max(), 1e-300), comp
    for sim in (a, b, c):
        sim.close()


def test_rejects_an_order_that_is_not_a_permutation(torch_cuda):
    p, r = mhd.dynamobench_case1()
    sim = mhd.MhdSimulation(p, r)
    bad = grids.rj_mode_order(p.l_truncation)
    bad[3] = bad[4]
    with pytest.raises(Exception, match="permutation"):
        sim.set_rj_order(bad)
    with pytest.raises(Exception, match="modes given"):
        sim.set_rj_order(np.arange(10))
    sim.close()
