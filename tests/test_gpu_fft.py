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
    # powers of two only (radix 4 stages, then one radix 2)
    (31, 8, 1, 64, 5),
    (15, 4, 2, 32, 13),
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
@pytest.mark.parametrize("L,nth,nr,nphi", [(31, 48, 3, 96), (15, 24, 2, 48), (255, 6, 2, 768),
                                            (31, 8, 2, 64)])
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


def _p(a):
    import ctypes
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("L,nth,nr,nphi,ncb,ncf", [(31, 48, 3, 96, 13, 12), (10, 6, 2, 24, 3, 5)])
def test_fft_selector_host_buffers(torch_cuda, L, nth, nr, nphi, ncb, ncf):
    """init / back / fwd / finalize_MHD_FFT_select on host arrays vs the oracle."""
    print(f"[fft sel] L={L} nth={nth} nr={nr} nphi={nphi} bwd={ncb} fwd={ncf}", flush=True)
    lib = calypso_amd.load()
    rng = np.random.default_rng(7 * L + ncb)
    nrtm, nrtp = nth * nr * (2 * L + 1), nth * nr * nphi
    wr = rng.standard_normal(ncb * nrtm)
    frc = rng.standard_normal(nrtp * ncf)
    fld = np.full(nrtp * ncb, np.nan)
    ws = np.full(ncf * nrtm, np.nan)
    calypso_amd._lib.check(lib.calypso_init_MHD_FFT_select(L, nth, nr, nphi, max(ncb, ncf), ncf, ncb))
    calypso_amd._lib.check(lib.calypso_back_MHD_FFT_sel_from_recv(ncb, wr.size, _p(wr), _p(fld)))
    calypso_amd._lib.check(lib.calypso_fwd_MHD_FFT_sel_from_recv(ncf, ws.size, _p(frc), _p(ws)))
    # a receive buffer shorter than ncomp * nnod_rtm is refused
    assert lib.calypso_back_MHD_FFT_sel_from_recv(ncb, wr.size - 1, _p(wr), _p(fld)) != 0
    calypso_amd._lib.check(lib.calypso_finalize_MHD_FFT_select())
    ref_b = oracle.fft_bwd(L, nth, nr, nphi, ncb, wr)
    ref_f = oracle.fft_fwd(L, nth, nr, nphi, ncf, frc)
    eb = np.max(np.abs(fld - ref_b)) / np.max(np.abs(ref_b))
    ef = np.max(np.abs(ws - ref_f)) / np.max(np.abs(ref_f))
    print(f"   err bwd={eb:.2e} fwd={ef:.2e}", flush=True)
    assert eb < 1e-13 and ef < 1e-13


@pytest.mark.parametrize("nb", [3, 5])
def test_fortran_fft_lu_driver(torch_cuda, tmp_path, nb):
    """The Fortran modules mhd_fft_select_hip, cal_nonlinear_sph_mhd_hip and
    lubksb_band_hip, called from a Fortran program with the reference's
    argument lists, against the oracle."""
    import os
    import subprocess
    L, nth, nr, nphi, ncb, ncf = 15, 24, 2, 48, 13, 12
    n, mcomp, stack = 11, 40, np.array([3, 10, 22, 37], dtype=np.int32)
    rng = np.random.default_rng(40 + nb)
    nrtm, nrtp = nth * nr * (2 * L + 1), nth * nr * nphi
    wr = rng.standard_normal(ncb * nrtm)
    frc = rng.standard_normal(nrtp * ncf)
    A = rng.standard_normal((mcomp, n, nb))
    A[:, :, (nb - 1) // 2] += 3.0
    lu, piv = oracle.ludcmp(nb, A)
    x = rng.standard_normal((n, mcomp))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([L, nth, nr, nphi, ncb, ncf], dtype=np.int32).tofile(f)
        wr.tofile(f)
        frc.tofile(f)
        np.array([nb, n, mcomp, stack.size - 1], dtype=np.int32).tofile(f)
        stack.tofile(f)
        np.ascontiguousarray(lu).tofile(f)
        np.ascontiguousarray(piv, dtype=np.int32).tofile(f)
        x.tofile(f)
    drv = os.path.join(calypso_amd._lib.LIB_DIR, "fft_lu_driver")
    print(f"[fortran fft/lu] {drv} nb={nb}", flush=True)
    r = subprocess.run([drv, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    fld, ws, xo = out[:nrtp * ncb], out[nrtp * ncb:nrtp * ncb + ncf * nrtm], out[nrtp * ncb + ncf * nrtm:]
    ref_b = oracle.fft_bwd(L, nth, nr, nphi, ncb, wr)
    ref_f = oracle.fft_fwd(L, nth, nr, nphi, ncf, frc)
    assert np.max(np.abs(fld - ref_b)) / np.max(np.abs(ref_b)) < 1e-13
    assert np.max(np.abs(ws - ref_f)) / np.max(np.abs(ref_f)) < 1e-13
    # modes stack[0]+1 .. stack[-1] solved bitwise as the reference; the rest untouched
    xo, frc2 = xo[:n * mcomp], xo[n * mcomp:]
    # grid products of the backward result, bitwise as the reference formulas
    assert np.array_equal(frc2, oracle.products_mhd(nrtp, fld, -1.0, 0.7, 1.3, -0.4))
    xo = xo.reshape(n, mcomp)
    m0, m1 = stack[0], stack[-1]
    ref_x = oracle.lubksb(nb, lu[m0:m1], piv[m0:m1], x[:, m0:m1])
    assert np.array_equal(xo[:, m0:m1], ref_x)
    assert np.array_equal(xo[:, :m0], x[:, :m0]) and np.array_equal(xo[:, m1:], x[:, m1:])
