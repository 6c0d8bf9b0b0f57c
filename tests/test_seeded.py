"""The seeded synthetic inputs: the oracle's generator pinned by published
values and closed forms, and the library's host generator bitwise equal to
it (no GPU needed: calypso_seeded_spectrum runs on the host)."""
import math

import numpy as np
import pytest

import oracle
from calypso_amd import mhd

GOLDEN = 0x9E3779B97F4A7C15


def test_splitmix64_published_sequence():
    """SplitMix64 from state 1234567: the first five outputs of the published
    generator (Steele, Lea & Flood 2014; Vigna's splitmix64.c)."""
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    x = 1234567
    got = []
    for _ in want:
        got.append(oracle.splitmix64(x))
        x = (x + GOLDEN) % 2 ** 64
    assert got == want


def test_box_muller_closed_form():
    r = math.sqrt(-2.0 * math.log(0.25))          # 2 sqrt(ln 2)
    assert oracle.box_muller(0.25, 0.0) == pytest.approx(r, rel=1e-15)
    assert oracle.box_muller(0.25, 0.5) == pytest.approx(-r, rel=1e-15)
    assert abs(oracle.box_muller(0.25, 0.25)) < 1e-15
    assert oracle.box_muller(math.exp(-0.5), 0.0) == pytest.approx(1.0, rel=1e-15)


def test_gaussian_moments():
    """10^6 draws (L = 999, one radius, no amplitude law): mean 0, variance
    1, skewness 0, kurtosis 3 within a few standard errors; KS against N(0,1)."""
    from scipy import stats
    x = oracle.seeded_spectrum(7, 3, 999, np.array([1.0]), 1.0, 0.0, False, 0).ravel()
    n = x.size
    assert n == 10 ** 6
    assert abs(x.mean()) < 5.0 / math.sqrt(n)
    assert abs(x.var() - 1.0) < 5.0 * math.sqrt(2.0 / n)
    assert abs(stats.skew(x)) < 5.0 * math.sqrt(6.0 / n)
    assert abs(stats.kurtosis(x)) < 5.0 * math.sqrt(24.0 / n)
    assert stats.kstest(x, "norm").pvalue > 1e-4


def test_amplitude_envelope_and_counter_property():
    r = np.linspace(0.5, 1.5, 9)
    L = 15
    base = oracle.seeded_spectrum(11, 6, L, r, 1.0, 0.0, False, 0)
    amp = oracle.seeded_spectrum(11, 6, L, r, 2.5, 2.0, True, 1)
    l_of_j = np.floor(np.sqrt(np.arange((L + 1) ** 2))).astype(int)
    env = ((r - 0.5) * (1.5 - r)) ** 2 / 0.5 ** 4
    want = 2.5 * base * ((l_of_j + 1.0) ** -2.0)[None, :] * env[:, None]
    want[:, l_of_j < 1] = 0.0
    np.testing.assert_allclose(amp, want, rtol=1e-14, atol=0)
    assert np.all(amp[0] == 0.0) and np.all(amp[-1] == 0.0)
    assert env[4] == 1.0
    # a value depends on (seed, field, k, j) only, not on the truncation
    small = oracle.seeded_spectrum(11, 6, 7, r, 1.0, 0.0, False, 0)
    assert np.array_equal(small, base[:, :64])
    # different fields and seeds give different numbers
    assert not np.array_equal(base, oracle.seeded_spectrum(11, 7, L, r, 1.0, 0.0, False, 0))
    assert not np.array_equal(base, oracle.seeded_spectrum(12, 6, L, r, 1.0, 0.0, False, 0))


@pytest.mark.parametrize("field,power,env,lmin", [(0, 2.0, True, 1), (12, 1.5, False, 0),
                                                  (8, 2.0, True, 1)])
def test_library_generator_matches_oracle(field, power, env, lmin):
    r = mhd.chebyshev_radius(48, 7.0 / 13.0, 20.0 / 13.0)
    got = mhd.seeded_spectrum(2024, field, 31, r, 0.3, power, env, lmin)
    want = oracle.seeded_spectrum(2024, field, 31, r, 0.3, power, env, lmin)
    assert np.array_equal(got, want)
