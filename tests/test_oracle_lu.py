"""Pins the oracle's banded LU: bitwise against the reference's own compiled
ludcmp_3band / ludcmp_band / lubksb_*band_mul, and by the residual ||Ax-b||
on diagonally dominant systems (where no pivoting happens)."""
import numpy as np
import pytest

import oracle


def dense(A, nb):
    mcomp, n, _ = A.shape
    nb1 = (nb - 1) // 2
    out = np.zeros((mcomp, n, n))
    for j in range(n):
        for k in range(nb):
            i = k + j - nb1            # band_a(i-j+nb1+1, j) = a(i, j)
            if 0 <= i < n:
                out[:, i, j] = A[:, j, k]
    return out


@pytest.mark.parametrize("nb", [3, 5])
def test_lu_residual_dominant(nb):
    rng = np.random.default_rng(nb)
    mcomp, n = 30, 17
    A = rng.standard_normal((mcomp, n, nb))
    A[:, :, (nb - 1) // 2] += 2.0 * nb
    lu, piv = oracle.ludcmp(nb, A)
    assert np.array_equal(piv, np.tile(np.arange(1, n + 1), (mcomp, 1)))
    b = rng.standard_normal((n, mcomp))
    x = oracle.lubksb(nb, lu, piv, b)
    D = dense(A, nb)
    for m in range(mcomp):
        np.testing.assert_allclose(D[m] @ x[:, m], b[:, m], atol=1e-13)


@pytest.mark.parametrize("nb", [3, 5])
def test_lu_3band_pivoting_residual(nb):
    if nb == 5:
        pytest.skip("the reference pairs ludcmp_band with lubksb_5band_mul, which is exact "
                    "only without row exchanges; its CN matrices never pivot")
    rng = np.random.default_rng(11)
    mcomp, n = 40, 12
    A = rng.standard_normal((mcomp, n, 3))
    lu, piv = oracle.ludcmp(3, A)
    assert (piv != np.arange(1, n + 1)).any()
    b = rng.standard_normal((n, mcomp))
    x = oracle.lubksb(3, lu, piv, b)
    D = dense(A, 3)
    for m in range(mcomp):
        np.testing.assert_allclose(D[m] @ x[:, m], b[:, m], atol=1e-10)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("nb", [3, 5])
def test_lu_bitwise_vs_compiled_reference(nb):
    rng = np.random.default_rng(100 + nb)
    mcomp, n = 64, 33
    A = rng.standard_normal((mcomp, n, nb))       # pivots happen
    A[::2, :, (nb - 1) // 2] += 3.0               # half the modes without pivots
    lu, piv = oracle.ludcmp(nb, A)
    lu2, piv2 = oracle.ludcmp(nb, A, use_ref=True)
    assert np.array_equal(lu, lu2) and np.array_equal(piv, piv2)
    b = rng.standard_normal((n, mcomp))
    assert np.array_equal(oracle.lubksb(nb, lu, piv, b), oracle.lubksb(nb, lu2, piv2, b, use_ref=True))
