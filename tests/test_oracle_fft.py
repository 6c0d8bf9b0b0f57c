"""Pins the oracle's longitude transforms and products: FFTPACK5 of the
compiled reference (RFFTMF/RFFTMB), numpy's FFT, round trips, and the
product formulas."""
import numpy as np
import pytest

import oracle


def seqs_from_grid(x, nc, nphi, nth, nr):
    X = x.reshape(nc, nphi, nth * nr)          # [nd][i][k + nr t]
    return X.transpose(0, 2, 1).reshape(-1, nphi)


@pytest.mark.parametrize("L,nth,nr,nphi,nc", [(5, 6, 2, 16, 3), (7, 4, 1, 24, 2), (11, 2, 2, 36, 1)])
def test_fft_vs_numpy_and_roundtrip(L, nth, nr, nphi, nc):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(nc * nr * nth * nphi)
    v = oracle.fft_fwd(L, nth, nr, nphi, nc, x)
    S = np.fft.rfft(seqs_from_grid(x, nc, nphi, nth, nr), axis=1)
    V = v.reshape(2 * L + 1, nr, nth, nc)
    for nd in range(nc):
        for t in range(nth):
            for k in range(nr):
                s = S[nd * nth * nr + k + nr * t]
                assert abs(V[L, k, t, nd] - s[0].real / nphi) < 1e-13
                for m in range(1, L + 1):
                    assert abs(V[L + m, k, t, nd] - 2 * s[m].real / nphi) < 1e-13
                    assert abs(V[L - m, k, t, nd] + 2 * s[m].imag / nphi) < 1e-13
    # backward of a band-limited spectrum returns it
    back = oracle.fft_fwd(L, nth, nr, nphi, nc, oracle.fft_bwd(L, nth, nr, nphi, nc, v))
    assert np.max(np.abs(back - v)) < 1e-13


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_fft_vs_compiled_fftpack():
    L, nth, nr, nphi, nc = 9, 3, 2, 32, 2
    rng = np.random.default_rng(3)
    x = rng.standard_normal(nc * nr * nth * nphi)
    v = oracle.fft_fwd(L, nth, nr, nphi, nc, x).reshape(2 * L + 1, nr, nth, nc)
    seqs = seqs_from_grid(x, nc, nphi, nth, nr)
    F = oracle.ref_rfftm(seqs, True)
    for nd in range(nc):
        for t in range(nth):
            for k in range(nr):
                f = F[nd * nth * nr + k + nr * t]
                assert abs(f[0] - v[L, k, t, nd]) < 1e-14
                for m in range(1, L + 1):
                    assert abs(f[2 * m - 1] - v[L + m, k, t, nd]) < 1e-14
                    assert abs(f[2 * m] - v[L - m, k, t, nd]) < 1e-14
    Fz = F.copy()
    Fz[:, 2 * L + 1:] = 0.0
    xb = oracle.fft_bwd(L, nth, nr, nphi, nc, v.ravel())
    assert np.max(np.abs(seqs_from_grid(xb, nc, nphi, nth, nr) - oracle.ref_rfftm(Fz, False))) < 1e-13


def test_products_formulas():
    rng = np.random.default_rng(1)
    n = 50
    fld = rng.standard_normal((13, n))
    c = (1.0, 2.5, -0.5, 3.0)
    frc = oracle.products_mhd(n, fld, *c).reshape(19, n)
    u, w, B, J, T = fld[0:3], fld[3:6], fld[6:9], fld[9:12], fld[12]
    np.testing.assert_allclose(frc[0:3], np.cross(w.T, u.T).T * c[0], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(frc[9:12], np.cross(J.T, B.T).T * c[1], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(frc[12:15], np.cross(u.T, B.T).T * c[2], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(frc[15:18], u * T * c[3], rtol=1e-14, atol=1e-14)
    assert not frc[3:9].any() and not frc[18].any()
