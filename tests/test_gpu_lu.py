"""GPU banded LU factor/solve, bitwise against the oracle (itself bitwise the
reference's compiled routines)."""
import numpy as np
import pytest

import oracle
from calypso_amd import radial

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nb", [3, 5])
@pytest.mark.parametrize("n,mcomp", [(49, 1024), (65, 3000), (9, 7)])
def test_lu_bitwise(torch_cuda, nb, n, mcomp):
    torch = torch_cuda
    print(f"[lu] nb={nb} n={n} mcomp={mcomp}", flush=True)
    rng = np.random.default_rng(n * nb)
    A = rng.standard_normal((mcomp, n, nb))
    A[::3, :, (nb - 1) // 2] += 4.0
    lu_ref, piv_ref = oracle.ludcmp(nb, A)
    b = rng.standard_normal((n, mcomp))
    x_ref = oracle.lubksb(nb, lu_ref, piv_ref, b)
    dA = torch.from_numpy(A).cuda()
    dlu = torch.full((mcomp, n, 2 * nb - 1), float("nan"), dtype=torch.float64, device="cuda")
    dpiv = torch.zeros((mcomp, n), dtype=torch.int32, device="cuda")
    derr = torch.full((mcomp,), -1, dtype=torch.int32, device="cuda")
    radial.ludcmp_band(nb, dA, dlu, dpiv, derr)
    dx = torch.from_numpy(b.copy()).cuda()
    radial.lubksb_band_mul(nb, dlu, dpiv, dx)
    torch.cuda.synchronize()
    assert np.array_equal(dlu.cpu().numpy(), lu_ref)
    assert np.array_equal(dpiv.cpu().numpy(), piv_ref)
    assert np.array_equal(dx.cpu().numpy(), x_ref)
    assert not derr.cpu().numpy().any()


def test_singular_3band_reported(torch_cuda):
    torch = torch_cuda
    n, mcomp = 8, 4
    A = np.ones((mcomp, n, 3))
    A[1, 3, :] = 0.0                      # row 4 of matrix 1 is zero (a(4,3..5) = 0)
    A[1, 2, 2] = 0.0
    A[1, 4, 0] = 0.0
    _, ref_piv = oracle.ludcmp(3, A)
    dA = torch.from_numpy(A).cuda()
    dlu = torch.zeros((mcomp, n, 5), dtype=torch.float64, device="cuda")
    dpiv = torch.zeros((mcomp, n), dtype=torch.int32, device="cuda")
    derr = torch.full((mcomp,), -1, dtype=torch.int32, device="cuda")
    radial.ludcmp_band(3, dA, dlu, dpiv, derr)
    torch.cuda.synchronize()
    err = derr.cpu().numpy()
    assert err[1] == 100 and err[0] == 0


@pytest.mark.parametrize("nb", [3, 5])
def test_lubksb_host_arrays_mode_range(torch_cuda, nb):
    """lubksb_3band_mul / lubksb_5band_mul on host arrays with the reference's
    (Msmp, Msmp_stack) ranges: the range solved bitwise as the oracle, the
    other modes untouched, and a second call reusing the device copies."""
    import ctypes
    import calypso_amd
    lib = calypso_amd.load()
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    f = lib.calypso_lubksb_3band_mul if nb == 3 else lib.calypso_lubksb_5band_mul
    for n, mcomp, stack in ((49, 700, [0, 100, 350, 700]), (9, 30, [4, 11, 19])):
        print(f"[lu host] nb={nb} n={n} mcomp={mcomp} stack={stack}", flush=True)
        rng = np.random.default_rng(n + mcomp + nb)
        A = rng.standard_normal((mcomp, n, nb))
        A[::2, :, (nb - 1) // 2] += 4.0
        lu, piv = oracle.ludcmp(nb, A)
        x = rng.standard_normal((n, mcomp))
        xo = x.copy()
        st = np.array(stack, dtype=np.int32)
        assert f(st.size - 1, p(st), mcomp, n, p(lu), p(piv), p(xo)) == 0
        m0, m1 = stack[0], stack[-1]
        ref = oracle.lubksb(nb, lu[m0:m1], piv[m0:m1], x[:, m0:m1])
        assert np.array_equal(xo[:, m0:m1], ref)
        assert np.array_equal(xo[:, :m0], x[:, :m0]) and np.array_equal(xo[:, m1:], x[:, m1:])
