"""GPU parity of the longitude FFTs (with the rtm<->rtp exchange folded in),
the grid products and the fused pencil kernel against the oracle."""
import numpy as np
import pytest

import calypso_amd
import oracle

pytestmark = pytest.mark.gpu


def to_dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def nan_dev(torch, n):
    return torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")


FFT_CASES = [
    # L, nth, nr, nphi, ncomp
    (31, 48, 3, 96, 13),
    (31, 48, 2, 96, 19),
    (10, 6, 2, 24, 3),
    (47, 8, 1, 144, 4),
    (127, 24, 1, 384, 19),
]


@pytest.mark.parametrize("L,nth,nr,nphi,nc", FFT_CASES)
def test_fft_backward_vs_oracle(torch_cuda, L, nth, nr, nphi, nc):
    print(f"[fft bwd] L={L} nth={nth} nr={nr} nphi={nphi} ncomp={nc}", flush=True)
    rng = np.random.default_rng(L + nc)
    vr = rng.standard_normal(nc * nth * nr * (2 * L + 1))
    ref = oracle.fft_bwd(L, nth, nr, nphi, nc, vr)
    f = calypso_amd.LongitudeFFT(L, nth, nr, nphi)
    out = nan_dev(torch_cuda, ref.size)
    f.backward(to_dev(torch_cuda, vr), out, nc)
    torch_cuda.cuda.synchronize()
    err = np.max(np.abs(out.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"   err={err:.2e}", flush=True)
    assert err < 1e-13
    f.close()


@pytest.mark.parametrize("L,nth,nr,nphi,nc", FFT_CASES)
def test_fft_forward_vs_oracle(torch_cuda, L, nth, nr, nphi, nc):
    print(f"[fft fwd] L={L} nth={nth} nr={nr} nphi={nphi} ncomp={nc}", flush=True)
    rng = np.random.default_rng(2 * L + nc)
    x = rng.standard_normal(nc * nr * nth * nphi)
    ref = oracle.fft_fwd(L, nth, nr, nphi, nc, x)
    f = calypso_amd.LongitudeFFT(L, nth, nr, nphi)
    out = nan_dev(torch_cuda, ref.size)
    f.forward(to_dev(torch_cuda, x), out, nc)
    torch_cuda.cuda.synchronize()
    err = np.max(np.abs(out.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print(f"   err={err:.2e}", flush=True)
    assert err < 1e-13
    f.close()


def test_products_bitwise(torch_cuda):
    rng = np.random.default_rng(9)
    n = 10007
    fld = rng.standard_normal(13 * n)
    coef = np.array([1.0, 0.2, -1.0, 3.0])
    ref = oracle.products_mhd(n, fld, *coef)
    out = nan_dev(torch_cuda, 19 * n)
    calypso_amd.nonlinear_products(to_dev(torch_cuda, fld), out, n, coef)
    torch_cuda.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


# L=255 / nphi=768: the 512-thread work-group with 151 KB of LDS, one per CU
@pytest.mark.parametrize("L,nth,nr,nphi", [(31, 48, 3, 96), (15, 24, 2, 48), (255, 6, 2, 768)])
def test_fused_mhd_pencil_vs_oracle(torch_cuda, L, nth, nr, nphi):
    print(f"[fused] L={L} nth={nth} nr={nr} nphi={nphi}", flush=True)
    rng = np.random.default_rng(L)
    M = 2 * L + 1
    vr = rng.standard_normal(13 * nth * nr * M)
    coef = np.array([-1.0, 0.7, 1.3, -0.4])
    nnod = nr * nth * nphi
    grid = oracle.fft_bwd(L, nth, nr, nphi, 13, vr)
    frc = oracle.products_mhd(nnod, grid, *coef)
    ref = oracle.fft_fwd(L, nth, nr, nphi, 19, frc)
    f = calypso_amd.LongitudeFFT(L, nth, nr, nphi)
    out = nan_dev(torch_cuda, ref.size)
    f.mhd_fused(to_dev(torch_cuda, vr), out, coef)
    torch_cuda.cuda.synchronize()
    got = out.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"   err={err:.2e}", flush=True)
    assert err < 1e-12
    f.close()
