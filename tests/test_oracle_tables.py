"""Pins the oracle's grid tables: Gauss points, Schmidt P_l^m and dP/dtheta,
normalisation constants and the rlm ordering."""
import math
import os

import numpy as np
import pytest
from scipy.special import lpmv

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "legendre_ref.npz")


@pytest.mark.parametrize("n", [1, 2, 12, 17, 48, 96])
def test_gauss_matches_numpy_and_quadrature(n):
    p, w, c = oracle.gauss_points(n)
    if n == 1:
        assert p[0] == 0.0 and w[0] == 2.0
        return
    x, wx = np.polynomial.legendre.leggauss(n)
    assert np.all(np.diff(p) > 0)                       # ascending from -1
    np.testing.assert_allclose(p, x, atol=2e-15)
    np.testing.assert_allclose(w, wx, rtol=0, atol=2e-12)  # Newton stops at 3e-12
    assert abs(w.sum() - 2.0) < 5e-12
    np.testing.assert_allclose(c, np.arccos(p), atol=0)
    # exact for polynomials up to degree 2n-1
    for k in range(0, 2 * n, 3):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.sum(w * p ** k) - exact) < 1e-11


def test_gauss_bitwise_vs_golden():
    g = np.load(GOLD)
    for name in ("l7", "l10_odd"):
        nth = int(g[f"{name}_meta"][1])
        p, w, c = oracle.gauss_points(nth)
        assert np.array_equal(p, g[f"{name}_point"])
        assert np.array_equal(w, g[f"{name}_weight"])
        assert np.array_equal(c, g[f"{name}_colat"])


def schmidt_closed_form(L, m, theta):
    """P_l^m Schmidt semi-normalised, no Condon-Shortley phase, from scipy."""
    x = math.cos(theta)
    out = np.zeros(L + 1)
    for l in range(m, L + 1):
        v = lpmv(m, l, x) * (-1) ** m          # remove the CS phase scipy applies
        if m > 0:
            v *= math.sqrt(2.0 * math.factorial(l - m) / math.factorial(l + m))
        out[l] = v
    return out


@pytest.mark.parametrize("m", [0, 1, 2, 5, 12])
@pytest.mark.parametrize("theta", [0.05, 0.7, 1.5707963267948966, 2.9])
def test_schmidt_vs_closed_form(m, theta):
    L = 12
    p, dp = oracle.schmidt_legendres_m(L, m, theta)
    ref = schmidt_closed_form(L, m, theta)
    np.testing.assert_allclose(p[m:], ref[m:], rtol=1e-12, atol=1e-14)
    # dP/dtheta by central differences of the closed form
    h = 1e-6
    num = (schmidt_closed_form(L, m, theta + h) - schmidt_closed_form(L, m, theta - h)) / (2 * h)
    np.testing.assert_allclose(dp[m:], num[m:], rtol=1e-6, atol=1e-8)


def test_schmidt_sectoral_closed_form():
    # P_l^l = sqrt(2 (2l)!)/(2^l l!) sin^l  for l > 0
    th = 0.9
    for l in range(1, 20):
        p, _ = oracle.schmidt_legendres_m(l, l, th)
        ref = math.sqrt(2.0 * math.factorial(2 * l)) / (2 ** l * math.factorial(l)) * math.sin(th) ** l
        assert abs(p[l] - ref) < 1e-13 * max(1.0, abs(ref))


def test_schmidt_tables_bitwise_vs_golden():
    g = np.load(GOLD)
    for name in ("l7", "l10_odd"):
        L, nth = int(g[f"{name}_meta"][0]), int(g[f"{name}_meta"][1])
        _, _, c = oracle.gauss_points(nth)
        P, dP = oracle.legendre_tables(L, c)
        assert np.array_equal(P, g[f"{name}_P"])
        assert np.array_equal(dP, g[f"{name}_dP"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("m", [0, 1, 3, 31])
def test_schmidt_bitwise_vs_compiled_reference(m):
    for th in (0.01, 1.0, 2.5):
        a, da = oracle.schmidt_legendres_m(31, m, th)
        b, db = oracle.schmidt_legendres_m(31, m, th, use_ref=True)
        assert np.array_equal(a[m:], b[m:]) and np.array_equal(da[m:], db[m:])


def test_normalizations_closed_forms():
    for l in range(0, 8):
        for m in range(0, l + 1):
            g = oracle.sph_normalizations(l, m)
            assert g[2] == l * (l + 1)
            assert g[15] == (2 * l + 1) / 4.0
            if l == 0:
                assert g[5] == 0.5 and g[6] == 1.0 and g[16] == 0.0
            elif m == 0:
                assert g[5] == (2 * l + 1) / 2.0
                assert g[6] == (2 * l + 1) / (2.0 * l * (l + 1))
            else:
                assert g[5] == (2 * l + 1) / 4.0
                assert g[6] == (2 * l + 1) / (4.0 * l * (l + 1))
    gr = oracle.g_sph_rlm(3)
    jlm, _ = oracle.rlm_index(3)
    j00 = int(np.where((jlm[:, 1] == 0) & (jlm[:, 2] == 0))[0][0])
    assert gr[j00, 2] == 0.5   # the l=m=0 override of g(3)


def test_rlm_index_order():
    L = 4
    jlm, lstack = oracle.rlm_index(L)
    assert lstack[-1] == (L + 1) ** 2
    ms = jlm[:, 2]
    assert np.all(np.diff(ms) >= 0) and ms[0] == -L and ms[-1] == L
    for j, (jj, l, m) in enumerate(jlm):
        assert jj == l * (l + 1) + m and abs(m) <= l <= L
