"""The allocation-free volume mean square (calypso_sph_mean_square_noalloc)
against the allocating entry point calypso_sph_mean_square, through the C ABI:
bitwise equal on two calls in a row, the second on a smaller input (the
library's device scratch array keeps the first call's contents), and sizes
beyond the array's limit rejected with an error before any GPU work."""
import ctypes

import numpy as np
import pytest

import oracle
from calypso_amd import _lib, radial

pytestmark = pytest.mark.gpu


def _case(L, nr, kind, seed):
    J = (L + 1) ** 2
    l_of_j = oracle.rlm_index(L)[0][:, 1].astype(np.int32).copy()
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    ncp = 1 if kind == 2 else 3
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((ncp, nr, J)) / (l_of_j + 1.0)
    return J, l_of_j, radius, ncp, f


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_noalloc_matches_allocating_entry(torch_cuda, kind):
    torch = torch_cuda
    # first call large (L=255, all 96 levels), second smaller (L=31, part of the shell)
    for L, nr, kr in ((255, 96, (1, 96)), (31, 49, (3, 40))):
        print(f"[no-alloc mean square] kind={kind} L={L} nr={nr} kr={kr}", flush=True)
        J, l_of_j, radius, ncp, f = _case(L, nr, kind, 100 * L + kind)
        d_f = torch.from_numpy(f).cuda()
        want = radial.sph_mean_square(kind, radius, l_of_j, kr[0], kr[1], d_f)
        got = radial.sph_mean_square(kind, radius, l_of_j, kr[0], kr[1], d_f, noalloc=True)
        assert np.array_equal(got[:ncp], want[:ncp]), (got, want)
        assert np.all(got[ncp:] == 0.0)
        ref = oracle.sph_mean_square(kind, radius, l_of_j, kr[0], kr[1], f)
        assert np.array_equal(got[:ncp], ref[:ncp])


def test_noalloc_repeated_calls_on_one_stream(torch_cuda):
    """Back-to-back calls on a non-default stream, shrinking: each result is
    the allocating entry point's, whatever the previous call left behind."""
    torch = torch_cuda
    s = torch.cuda.Stream()
    for L, nr in ((127, 65), (63, 33), (15, 9)):
        J, l_of_j, radius, ncp, f = _case(L, nr, 0, L)
        d_f = torch.from_numpy(f).cuda()
        torch.cuda.synchronize()
        got = radial.sph_mean_square(0, radius, l_of_j, 1, nr, d_f, stream=s.cuda_stream, noalloc=True)
        want = radial.sph_mean_square(0, radius, l_of_j, 1, nr, d_f, stream=s.cuda_stream)
        assert np.array_equal(got, want)


def test_noalloc_rejects_sizes_beyond_its_array(torch_cuda):
    torch = torch_cuda
    lib = _lib.load()
    d_f = torch.zeros(8, dtype=torch.float64, device="cuda")
    out = np.zeros(3)
    P = ctypes.c_void_p
    # J above (511+1)^2: rejected on the host before any copy or launch
    J = 512 * 512 + 1
    l_of_j = np.zeros(J, dtype=np.int32)
    radius = np.linspace(0.5, 1.5, 4)
    rc = lib.calypso_sph_mean_square_noalloc(
        0, 4, J, l_of_j.ctypes.data_as(P), radius.ctypes.data_as(P), 1, 4,
        P(d_f.data_ptr()), out.ctypes.data_as(P), None)
    assert rc != 0
    assert b"J <= 262144" in lib.calypso_last_error()
    # too many radial points
    radius = np.linspace(0.5, 1.5, 5000)
    l_of_j = np.zeros(4, dtype=np.int32)
    rc = lib.calypso_sph_mean_square_noalloc(
        0, 5000, 4, l_of_j.ctypes.data_as(P), radius.ctypes.data_as(P), 1, 5000,
        P(d_f.data_ptr()), out.ctypes.data_as(P), None)
    assert rc != 0
    assert b"nr <= 4096" in lib.calypso_last_error()
    # a degree outside the reduction's table
    l_of_j = np.array([0, 1, 1, 4000], dtype=np.int32)
    radius = np.linspace(0.5, 1.5, 4)
    rc = lib.calypso_sph_mean_square_noalloc(
        0, 4, 4, l_of_j.ctypes.data_as(P), radius.ctypes.data_as(P), 1, 4,
        P(d_f.data_ptr()), out.ctypes.data_as(P), None)
    assert rc != 0
    assert b"degrees" in lib.calypso_last_error()
