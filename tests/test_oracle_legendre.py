"""Pins the oracle's Legendre transforms against the reference's compiled
org loops (golden fixtures and, where built, the live library) and against
the quadrature round trip fwd(bwd(x)) = x on band-limited data."""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "legendre_ref.npz")


def grid_from_gold(g, name):
    L, nth, nr, nv, ns = (int(x) for x in g[f"{name}_meta"])
    return oracle.Grid(L, nth, g[f"{name}_radius"]), nv, ns


@pytest.mark.parametrize("name", ["l7", "l10_odd"])
def test_oracle_bitwise_vs_reference_golden(name):
    g = np.load(GOLD)
    grid, nv, ns = grid_from_gold(g, name)
    nc = 3 * nv + ns
    assert np.array_equal(oracle.leg_bwd(grid, nc, nv, ns, g[f"{name}_sp_in"]), g[f"{name}_vr_out"])
    assert np.array_equal(oracle.leg_fwd(grid, nc, nv, ns, g[f"{name}_vr_in"]), g[f"{name}_sp_out"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("L,nth,nr,nv,ns", [(5, 8, 2, 1, 0), (9, 14, 3, 0, 2), (15, 24, 2, 4, 1)])
def test_oracle_bitwise_vs_compiled_reference(L, nth, nr, nv, ns):
    rng = np.random.default_rng(L)
    grid = oracle.Grid(L, nth, np.linspace(0.5, 1.5, nr))
    ref = oracle.RefLegendre(grid)
    nc = 3 * nv + ns
    sp = rng.standard_normal(grid.n_rlm(nc))
    vr = rng.standard_normal(grid.n_rtm(nc))
    assert np.array_equal(oracle.leg_bwd(grid, nc, nv, ns, sp), ref.bwd(nc, nv, ns, sp))
    assert np.array_equal(oracle.leg_fwd(grid, nc, nv, ns, vr), ref.fwd(nc, nv, ns, vr))


def band_limited_spectrum(grid, nc, nv, rng):
    """Random spectrum with the l=0 vector parts zeroed (they carry no field)."""
    sp = rng.standard_normal((grid.nr, grid.J, nc))
    l0 = grid.jlm[:, 1] == 0
    sp[:, l0, :3 * nv] = 0.0
    return sp.ravel()


@pytest.mark.parametrize("L,nth,nv,ns", [(7, 12, 2, 1), (12, 19, 1, 1), (16, 26, 0, 3)])
def test_roundtrip_identity(L, nth, nv, ns):
    # Gauss quadrature with nth >= L+1 points integrates P_l P_l' exactly,
    # so fwd(bwd(x)) = x; a dropped term, a wrong sign or a wrong
    # normalisation breaks it.  The reference's Gauss weights carry ~3e-13
    # error (Newton stops at 3e-12), which bounds the residual.
    rng = np.random.default_rng(7)
    grid = oracle.Grid(L, nth, np.array([0.6, 1.3]))
    nc = 3 * nv + ns
    sp = band_limited_spectrum(grid, nc, nv, rng)
    back = oracle.leg_fwd(grid, nc, nv, ns, oracle.leg_bwd(grid, nc, nv, ns, sp))
    assert np.max(np.abs(back - sp)) < 1e-10 * np.max(np.abs(sp))


def test_backward_single_mode_closed_form():
    # scalar mode (l, m) = (3, -2) with unit coefficient: the grid value at
    # the -2 slot is P_3^2(theta); every other slot is zero
    L, nth = 4, 8
    grid = oracle.Grid(L, nth, np.array([1.0]))
    sp = np.zeros(grid.n_rlm(1))
    j = int(np.where((grid.jlm[:, 1] == 3) & (grid.jlm[:, 2] == -2))[0][0])
    sp[j] = 1.0
    vr = oracle.leg_bwd(grid, 1, 0, 1, sp).reshape(grid.M, 1, nth)
    p32 = np.array([oracle.schmidt_legendres_m(L, 2, t)[0][3] for t in grid.colat])
    np.testing.assert_allclose(vr[L - 2, 0], p32, rtol=0, atol=0)
    assert np.count_nonzero(np.delete(vr, L - 2, axis=0)) == 0


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("L,nth,nr,nv,ns", [(7, 12, 3, 3, 1), (31, 48, 4, 4, 1), (20, 33, 2, 1, 2)])
def test_reference_sym_matmul_agrees_with_org_loops(L, nth, nr, nv, ns):
    """The reference's SYMMETRIC_MATMUL transforms (the CPU baseline, the
    variant its own tests select) against its org loops and the oracle: the
    same sums in another order, so agreement to rounding (1e-13 normwise)."""
    rng = np.random.default_rng(L + nr)
    grid = oracle.Grid(L, nth, np.linspace(0.5, 1.5, nr))
    ref = oracle.RefLegendre(grid, nthreads=2)
    nc = 3 * nv + ns
    sp = rng.standard_normal(grid.n_rlm(nc))
    vr = rng.standard_normal(grid.n_rtm(nc))
    for a, b in ((ref.bwd_sym(nc, nv, ns, sp), oracle.leg_bwd(grid, nc, nv, ns, sp)),
                 (ref.fwd_sym(nc, nv, ns, vr), oracle.leg_fwd(grid, nc, nv, ns, vr))):
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
