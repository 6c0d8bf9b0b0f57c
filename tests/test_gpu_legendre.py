"""GPU parity of the HIP Legendre transforms against the oracle and the
reference golden vectors (FP64: 1e-12 normwise per field; the mixed modes are
held to their own documented bounds)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import calypso_amd
import oracle

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "legendre_ref.npz")

FP64_TOL = 1e-12
# about 4x the largest error each mode reaches over CASES (FP32 4.4e-7,
# split-BF16x3 3.8e-7); the int8-slice mode is held to the 1e-11 bound.
# FP32 and split-BF16x3 do not meet 1e-11.
MIXED_TOL = {"hip_fp32": 2e-6, "hip_bf16x3": 1.5e-6, "hip_int8x6": 1e-11}


def field_err(got, ref, ncomp):
    """max over components of max|got-ref| / max|ref| (normwise per field)."""
    g = np.asarray(got).reshape(-1, ncomp)
    r = np.asarray(ref).reshape(-1, ncomp)
    errs = []
    for nd in range(ncomp):
        s = np.max(np.abs(r[:, nd]))
        errs.append(np.max(np.abs(g[:, nd] - r[:, nd])) / (s if s > 0 else 1.0))
    return max(errs)


def run_dev(torch, lt, sp=None, vr=None):
    dev = torch.device("cuda:0")
    if sp is not None:
        d_sp = torch.from_numpy(np.ascontiguousarray(sp)).to(dev)
        d_vr = torch.full((lt.n_rtm(),), float("nan"), dtype=torch.float64, device=dev)
        lt.backward(d_sp, d_vr)
        torch.cuda.synchronize()
        return d_vr.cpu().numpy()
    d_vr = torch.from_numpy(np.ascontiguousarray(vr)).to(dev)
    d_sp = torch.full((lt.n_rlm(),), float("nan"), dtype=torch.float64, device=dev)
    lt.forward(d_vr, d_sp)
    torch.cuda.synchronize()
    return d_sp.cpu().numpy()


@pytest.mark.parametrize("name", ["l7", "l10_odd"])
def test_tables_fp64_vs_reference_golden(torch_cuda, name):
    g = np.load(GOLD)
    L, nth, nr, nv, ns = (int(x) for x in g[f"{name}_meta"])
    print(f"[tables] case={name} L={L} nth={nth}", flush=True)
    lt = calypso_amd.LegendreTransform(L, nth, g[f"{name}_radius"], 3 * nv + ns, nv, ns, "hip_fp64")
    P, dP = lt.tables()
    assert np.max(np.abs(P - g[f"{name}_P"])) < 1e-13
    assert np.max(np.abs(dP - g[f"{name}_dP"])) < 1e-12
    lt.close()


@pytest.mark.parametrize("name", ["l7", "l10_odd"])
def test_fp64_vs_reference_golden(torch_cuda, name):
    g = np.load(GOLD)
    L, nth, nr, nv, ns = (int(x) for x in g[f"{name}_meta"])
    nc = 3 * nv + ns
    print(f"[golden] case={name} L={L} nth={nth} nr={nr} nv={nv} ns={ns}", flush=True)
    lt = calypso_amd.LegendreTransform(L, nth, g[f"{name}_radius"], nc, nv, ns, "hip_fp64")
    vr = run_dev(torch_cuda, lt, sp=g[f"{name}_sp_in"])
    assert field_err(vr, g[f"{name}_vr_out"], nc) < FP64_TOL
    sp = run_dev(torch_cuda, lt, vr=g[f"{name}_vr_in"])
    assert field_err(sp, g[f"{name}_sp_out"], nc) < FP64_TOL
    lt.close()


CASES = [
    # L, nth, nr, nv, ns : the MHD field sets (13 bwd / 19 fwd) at L=31,
    # the 12-component benchmark set at a reduced grid, scalars only, odd nth
    (31, 48, 5, 4, 1),
    (31, 48, 4, 6, 1),
    (63, 96, 3, 4, 0),
    (20, 33, 2, 0, 3),
    (47, 72, 2, 2, 2),
    # several LDS chunks along degree (bwd) and colatitude (fwd)
    (127, 192, 2, 1, 1),
    # more than one group of row tiles per order (fwd, m small)
    (191, 288, 1, 1, 0),
    # tiny truncations (one partial row tile, one LDS chunk), and a grid with
    # many more colatitudes than degrees
    (1, 2, 1, 1, 1),
    (2, 3, 2, 1, 0),
    (5, 16, 1, 0, 2),
]


@pytest.mark.parametrize("mode", ["hip_fp64", "hip_fp32", "hip_bf16x3", "hip_int8x6"])
@pytest.mark.parametrize("L,nth,nr,nv,ns", CASES)
def test_vs_oracle(torch_cuda, mode, L, nth, nr, nv, ns):
    nc = 3 * nv + ns
    print(f"[oracle] mode={mode} L={L} nth={nth} nr={nr} nv={nv} ns={ns}", flush=True)
    rng = np.random.default_rng(L * 7 + nr)
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    grid = oracle.Grid(L, nth, radius)
    l = grid.jlm[:, 1].astype(float)
    sp = (rng.standard_normal((nr, grid.J, nc)) * ((l + 1.0) ** -1.5)[None, :, None]).ravel()
    vr = rng.standard_normal(grid.n_rtm(nc))
    lt = calypso_amd.LegendreTransform(L, nth, radius, nc, nv, ns, mode)
    tol = FP64_TOL if mode == "hip_fp64" else MIXED_TOL[mode]
    e_b = field_err(run_dev(torch_cuda, lt, sp=sp), oracle.leg_bwd(grid, nc, nv, ns, sp), nc)
    e_f = field_err(run_dev(torch_cuda, lt, vr=vr), oracle.leg_fwd(grid, nc, nv, ns, vr), nc)
    print(f"   err bwd={e_b:.3e} fwd={e_f:.3e}", flush=True)
    assert e_b < tol and e_f < tol
    lt.close()


def test_host_selector_api(torch_cuda):
    g = np.load(GOLD)
    name = "l7"
    L, nth, nr, nv, ns = (int(x) for x in g[f"{name}_meta"])
    nc = 3 * nv + ns
    calypso_amd.sel_init_legendre_trans("hip_fp64", L, nth, g[f"{name}_radius"],
                                        g[f"{name}_colat"], g[f"{name}_weight"], nc, nv, ns)
    WR = np.array(g[f"{name}_sp_in"])
    WS = np.zeros(g[f"{name}_vr_out"].size)
    calypso_amd.sel_backward_legendre_trans(nc, nv, ns, WR.size, WS.size, WR, WS)
    assert field_err(WS, g[f"{name}_vr_out"], nc) < FP64_TOL
    WR2 = np.array(g[f"{name}_vr_in"])
    WS2 = np.zeros(g[f"{name}_sp_out"].size)
    calypso_amd.sel_forward_legendre_trans(nc, nv, ns, WR2.size, WS2.size, WR2, WS2)
    assert field_err(WS2, g[f"{name}_sp_out"], nc) < FP64_TOL
    calypso_amd.sel_finalize_legendre_trans()


def test_fortran_module_driver(torch_cuda, tmp_path):
    g = np.load(GOLD)
    name = "l7"
    L, nth, nr, nv, ns = (int(x) for x in g[f"{name}_meta"])
    nc = 3 * nv + ns
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([L, nth, nr, nv, ns], dtype=np.int32).tofile(f)
        for key in ("radius", "colat", "weight", "sp_in"):
            np.asarray(g[f"{name}_{key}"], dtype=np.float64).tofile(f)
    drv = os.path.join(calypso_amd._lib.LIB_DIR, "leg_driver")
    print(f"[fortran] {drv} hip_fp64", flush=True)
    r = subprocess.run([drv, "hip_fp64", str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    out = np.fromfile(fout, dtype=np.float64)
    n_rtm = g[f"{name}_vr_out"].size
    vr, sp2 = out[:n_rtm], out[n_rtm:]
    assert field_err(vr, g[f"{name}_vr_out"], nc) < FP64_TOL
    ref2 = oracle.leg_fwd(oracle.Grid(L, nth, g[f"{name}_radius"]), nc, nv, ns, vr)
    assert field_err(sp2, ref2, nc) < FP64_TOL


def test_full_size_roundtrip_fp64(torch_cuda):
    """Benchmark shape (L=255, nth=384, nr=96, 4 vectors): fwd(bwd(x)) = x.

    Spectrum amplitude (l+1)^-1.5 as in the benchmark.  Gauss quadrature
    with nth >= L+1 points makes the pair an identity on
    band-limited data up to the error of the Gauss weights, which the
    reference's Newton iteration leaves at ~3e-13 (it stops at 3e-12; the
    weights sum to 2 - 8e-13).  Modes l >= 1 see that error directly; the
    l = 0 poloidal output collects it times l(l+1) <= 6.6e4 from every m = 0
    mode, so it gets its own bound.
    """
    torch = torch_cuda
    L, nth, nr, nv, ns = 255, 384, 96, 4, 0
    nc = 12
    print(f"[roundtrip] L={L} nth={nth} nr={nr} comps={nc}", flush=True)
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    lt = calypso_amd.LegendreTransform(L, nth, radius, nc, nv, ns, "hip_fp64")
    print("   plans:", lt.plan_info(False), lt.plan_info(True), flush=True)
    gen = torch.Generator(device="cuda").manual_seed(5)
    J = (L + 1) ** 2
    degs = np.concatenate([np.arange(abs(m), L + 1) for m in range(-L, L + 1)])
    amp = torch.from_numpy((degs + 1.0) ** -1.5).cuda()
    sp = torch.randn((nr, J, nc), generator=gen, device="cuda", dtype=torch.float64)
    sp = sp * amp[None, :, None]
    j00 = L * (L + 1) // 2                  # (l, m) = (0, 0) in rlm order
    sp[:, j00, :] = 0.0                     # l = 0 vector parts carry no field
    sp = sp.reshape(-1).contiguous()
    vr = torch.empty(lt.n_rtm(), dtype=torch.float64, device="cuda")
    sp2 = torch.empty_like(sp)
    lt.backward(sp, vr)
    lt.forward(vr, sp2)
    torch.cuda.synchronize()
    scale = sp.abs().max().item()
    d = (sp2 - sp).abs().reshape(nr, J, nc)
    err_l0 = d[:, j00, :].max().item() / scale
    d[:, j00, :] = 0.0
    err = d.max().item() / scale
    k, j, c = np.unravel_index(int(d.argmax().item()), (nr, J, nc))
    ords = np.concatenate([np.full(L - abs(m) + 1, m) for m in range(-L, L + 1)])
    print(f"   roundtrip err(l>=1)={err:.3e} at k={k} l={degs[j]} m={ords[j]} comp={c}; "
          f"err(l=0)={err_l0:.3e}", flush=True)
    assert err < 1e-9
    assert err_l0 < 1e-6
    lt.close()


PLAN_CASES = [(31, 48, 9, 4, 1), (31, 48, 5, 6, 1), (47, 72, 7, 4, 0), (64, 96, 6, 1, 1),
              # the golden-fixture shapes: tiny grid, odd number of colatitudes
              (7, 12, 3, 3, 1), (10, 17, 2, 1, 2)]


@pytest.mark.parametrize("mode", ["hip_fp64", "hip_fp32", "hip_bf16x3", "hip_int8x6"])
@pytest.mark.parametrize("L,nth,nr,nv,ns", PLAN_CASES)
def test_every_tile_plan_vs_oracle(torch_cuda, mode, L, nth, nr, nv, ns):
    """Every candidate kernel plan the autotuner may pick, forced one by one."""
    nc = 3 * nv + ns
    rng = np.random.default_rng(L + 31 * nr + nv)
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    grid = oracle.Grid(L, nth, radius)
    l = grid.jlm[:, 1].astype(float)
    sp = (rng.standard_normal((nr, grid.J, nc)) * ((l + 1.0) ** -1.5)[None, :, None]).ravel()
    vr = rng.standard_normal(grid.n_rtm(nc))
    ref_b = oracle.leg_bwd(grid, nc, nv, ns, sp)
    ref_f = oracle.leg_fwd(grid, nc, nv, ns, vr)
    lt = calypso_amd.LegendreTransform(L, nth, radius, nc, nv, ns, mode)
    tol = FP64_TOL if mode == "hip_fp64" else MIXED_TOL[mode]
    for fwd in (False, True):
        n = lt.plan_count(fwd)
        assert n >= 1
        for idx in range(n):
            lt.force_plan(fwd, idx)
            info = lt.plan_info(fwd)
            print(f"[plans] mode={mode} L={L} nr={nr} nv={nv} ns={ns} fwd={fwd} "
                  f"plan {idx}/{n} {info}", flush=True)
            e = field_err(run_dev(torch_cuda, lt, vr=vr), ref_f, nc) if fwd else \
                field_err(run_dev(torch_cuda, lt, sp=sp), ref_b, nc)
            assert e < tol, (fwd, idx, info, e)
    lt.close()


def test_same_shape_contexts_bitwise(torch_cuda):
    """Two contexts of one shape take the same (cached) autotuned plans, so
    their transforms agree bitwise."""
    L, nth, nr, nv, ns = 47, 72, 6, 4, 1
    nc = 3 * nv + ns
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    rng = np.random.default_rng(5)
    grid = oracle.Grid(L, nth, radius)
    sp = rng.standard_normal(grid.n_rlm(nc))
    outs = []
    for _ in range(2):
        lt = calypso_amd.LegendreTransform(L, nth, radius, nc, nv, ns, "hip_fp64")
        vr = run_dev(torch_cuda, lt, sp=sp)
        outs.append((vr, run_dev(torch_cuda, lt, vr=vr), lt.plan_info(False), lt.plan_info(True)))
        lt.close()
    assert outs[0][2] == outs[1][2] and outs[0][3] == outs[1][3]
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_int8x6_extreme_component_scales(torch_cuda):
    """int8-slice mode on components that are zero, ~1e-200 and ~1e+100 (the
    per-column exponents and the exponent floor): each field within 1e-11 of
    the oracle, the zero field exactly zero."""
    L, nth, nr, nv, ns = 31, 48, 3, 1, 3
    nc = 3 * nv + ns
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    rng = np.random.default_rng(11)
    grid = oracle.Grid(L, nth, radius)
    l = grid.jlm[:, 1].astype(float)
    scale = np.array([1.0, 1e-200, 1e100, 1.0, 0.0, 3e-7])
    sp = (rng.standard_normal((nr, grid.J, nc)) * ((l + 1.0) ** -1.5)[None, :, None]
          * scale[None, None, :]).ravel()
    vr = (rng.standard_normal((grid.n_rtm(nc) // nc, nc)) * scale[None, :]).ravel()
    lt = calypso_amd.LegendreTransform(L, nth, radius, nc, nv, ns, "hip_int8x6")
    got_b = run_dev(torch_cuda, lt, sp=sp)
    got_f = run_dev(torch_cuda, lt, vr=vr)
    lt.close()
    ref_b = oracle.leg_bwd(grid, nc, nv, ns, sp)
    ref_f = oracle.leg_fwd(grid, nc, nv, ns, vr)
    for got, ref in ((got_b, ref_b), (got_f, ref_f)):
        g = got.reshape(-1, nc)
        r = ref.reshape(-1, nc)
        assert np.all(g[:, 4] == 0.0)
        for c in (0, 1, 2, 3, 5):
            e = np.max(np.abs(g[:, c] - r[:, c])) / np.max(np.abs(r[:, c]))
            print(f"   comp {c} scale {scale[c]:.0e}: {e:.2e}", flush=True)
            assert e < 1e-11, (c, e)
