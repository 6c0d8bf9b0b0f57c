"""The benchmark configurations on the GPU, each against the oracle or the
FP64 GPU path:

* C1  L=31, 48 x 96 grid, 49 radial points, E=1e-3, Ra=100, Pr=1, Pm=5,
      seeded random start, dt=1e-4, Crank-Nicolson weight 0.5: every stage
      field against the oracle after the initialisation and three steps, the
      energies after 50 and 100 steps against the oracle's
      (tests/golden/seeded_c1_energies.json), and after every one of 100
      steps against the reference's own program started from the same
      fields (tests/golden/seeded_c1_reference_steps.json), also through a
      restart file;
* C3  L=127, 192 x 384 grid, 65 radial points, dt=5e-5: the FP64 step
      against the oracle after two steps, then 200 steps with every Legendre
      precision against the FP64 step (energy drift and fields);
* C4  64 shells of uneven size (L in {31, 47, 63, 95}, Nr in {32, 48, 64,
      96}, seeded Ra in [50, 500]) stepped together: bitwise the state of
      each shell stepped alone;
* C5  L=511, 768 colatitudes, 128 radial points, 2 vectors + 2 scalars:
      every precision's transforms against the oracle on sampled radial
      levels; the L=511, 768 x 1536, 129-point MHD step for 20 steps in FP64
      and split-BF16x3, finite energies, per-field difference and the HBM
      footprint.
"""
import json
import os

import numpy as np
import pytest

import calypso_amd
import oracle
from calypso_amd import mhd
from oracle import mhd_step

from test_gpu_mhd import COMPARED, _compare

pytestmark = pytest.mark.gpu

# seed, amplitudes of u, B and the temperature perturbation: a magnetic field
# ten times weaker than the velocity keeps the Lorentz-driven start-up
# transient within the explicit step's stability limit on the finer grids
# (with B at 0.3 the L=127 step at dt=5e-5 leaves its stable range after
# about 30 steps)
SEEDED = (2024, 0.3, 0.03, 0.1)
GOLD_C1 = os.path.join(os.path.dirname(__file__), "golden", "seeded_c1_energies.json")
MODES = ("hip_fp64", "hip_fp32", "hip_bf16x3", "hip_int8x6")


def _seeded_sim(p, r, stream=None, seeded=SEEDED):
    sim = mhd.MhdSimulation(p, r, stream=stream)
    sim.set_seeded_ic(*seeded)
    sim.initialize()
    return sim


def _field_diff(a, b, comps=None):
    worst = {}
    for _, _, comp in (COMPARED if comps is None else comps):
        x, y = a.get_field(comp), b.get_field(comp)
        worst[comp] = float(np.abs(x - y).max() / max(np.abs(x).max(), 1e-300))
    return worst


# --------------------------------------------------------------------- C1 --

def test_c1_seeded_steps_match_oracle(torch_cuda):
    print("[C1] seeded start, L=31, dt=1e-4, CN 0.5: init + 3 steps vs oracle", flush=True)
    p, r = mhd.dynamobench_case1(dt=1.0e-4, coef_imp=0.5)
    sim = _seeded_sim(p, r)
    ref = mhd_step.MhdStep(mhd_step.shell_case(dt=1.0e-4, coef_imp=0.5))
    ref.initialize(seeded=SEEDED)
    _compare(sim, ref, 1e-11, "C1 init")
    for n in (1, 2, 3):
        sim.step(1)
        ref.step()
        _compare(sim, ref, 1e-10, f"C1 step {n}")
    np.testing.assert_allclose(sim.energies(), ref.energies(), rtol=1e-11)
    sim.close()


def test_c1_seeded_100_steps_energies(torch_cuda):
    gold = json.load(open(GOLD_C1))
    assert tuple(gold["seeded"].values()) == SEEDED
    print("[C1] seeded start, 100 steps vs the oracle's energies", flush=True)
    p, r = mhd.dynamobench_case1(dt=1.0e-4, coef_imp=0.5)
    sim = _seeded_sim(p, r)
    got = {"0": sim.energies()}
    for _ in range(2):
        sim.step(50)
        got[str(sim.i_step)] = sim.energies()
    sim.close()
    for step, want in gold["steps"].items():
        want = np.array(want)
        rel = np.max(np.abs(got[step] - want) / np.abs(want))
        print(f"   step {step}: max rel diff {rel:.2e}", flush=True)
        assert rel < 1e-10, (step, got[step], want)


REF_C1 = os.path.join(os.path.dirname(__file__), "golden", "seeded_c1_reference_steps.json")


def test_c1_seeded_matches_reference_program(torch_cuda):
    """C1 on the GPU against the reference's own MHD program started from the
    same seeded fields (its restart file): energies after every one of 100
    steps within 1e-12."""
    ref = json.load(open(REF_C1))
    assert tuple(ref["seeded"].values()) == SEEDED
    print("[C1] seeded start, 100 steps vs the reference program", flush=True)
    p, r = mhd.dynamobench_case1(dt=ref["dt"], coef_imp=ref["coef_imp"])
    sim = _seeded_sim(p, r)
    worst = 0.0
    for n in range(1, 101):
        sim.step(1)
        want = np.array(ref["steps"][str(n)])
        worst = max(worst, float(np.max(np.abs(sim.energies() / want - 1.0))))
    sim.close()
    print(f"   worst relative energy difference {worst:.1e}", flush=True)
    assert worst < 1e-12, worst


def test_c1_restart_file_layout_and_resume(torch_cuda, tmp_path):
    """The restart file holds each vector as (pol, tor, dpol/dr), the layout
    of copy_each_sph_vector_to_IO (copy_rj_phys_data_4_IO.f90:219-236) that
    the reference program read in the C1 reference run; a state read back
    from it resumes that run: energies of steps 1-10 within 1e-12 of the
    reference program's."""
    ref = json.load(open(REF_C1))
    p, r = mhd.dynamobench_case1(dt=ref["dt"], coef_imp=ref["coef_imp"])
    sim = mhd.MhdSimulation(p, r)
    sim.set_seeded_ic(*SEEDED)
    u = [sim.get_field(c) for c in (0, 1, 2)]
    path = sim.write_restart(str(tmp_path / "rst"), 0)
    sim.close()
    lines = open(path).read().splitlines()
    i = lines.index("velocity")
    nnod = u[0].size
    block = np.array([[float(x) for x in lines[i + 1 + k].split()] for k in range(nnod)])
    # 16 significant digits in the file ('1pE25.15')
    np.testing.assert_allclose(block[:, 0], u[0].reshape(-1), rtol=2e-15, atol=0)
    np.testing.assert_allclose(block[:, 1], u[2].reshape(-1), rtol=2e-15, atol=0)  # toroidal
    np.testing.assert_allclose(block[:, 2], u[1].reshape(-1), rtol=2e-15, atol=0)
    print("[C1] resume from the restart file, 10 steps vs the reference program", flush=True)
    sim = mhd.MhdSimulation(p, r)
    sim.read_restart(path)
    sim.initialize()
    for n in range(1, 11):
        sim.step(1)
        want = np.array(ref["steps"][str(n)])
        np.testing.assert_allclose(sim.energies(), want, rtol=1e-12, atol=0,
                                   err_msg=f"step {n}")
    sim.close()


# --------------------------------------------------------------------- C3 --

C3 = dict(l_truncation=127, nth=192, nphi=384, nri=64, dt=5.0e-5, coef_imp=0.5)


def test_c3_fp64_step_matches_oracle(torch_cuda):
    print("[C3] L=127 192x384 Nr=65, FP64: init + 2 steps vs oracle", flush=True)
    p, r = mhd.dynamobench_case1(**C3)
    sim = _seeded_sim(p, r)
    ref = mhd_step.MhdStep(mhd_step.shell_case(L=127, nth=192, nphi=384, nri=64, dt=5.0e-5,
                                               coef_imp=0.5))
    ref.initialize(seeded=SEEDED)
    _compare(sim, ref, 1e-11, "C3 init")
    for n in (1, 2):
        sim.step(1)
        ref.step()
        _compare(sim, ref, 1e-10, f"C3 step {n}")
    sim.close()


# energy drift after 200 steps against the FP64 step, per precision: about
# 5x what each mode reaches (FP32 6.4e-10, split-BF16x3 2.9e-9, int8 slices
# 9.3e-15)
C3_ENERGY_TOL = {"hip_fp32": 3e-9, "hip_bf16x3": 1.5e-8, "hip_int8x6": 5e-14}


@pytest.mark.timeout(300)
def test_c3_precisions_200_steps(torch_cuda):
    """The graph-captured C3 step for 200 steps with each Legendre precision
    (FP64 solves, recursion and diagnostics in all of them)."""
    sims = {}
    for mode in MODES:
        print(f"[C3] 200 steps, {mode}", flush=True)
        p, r = mhd.dynamobench_case1(legendre_mode=mode, **C3)
        sim = _seeded_sim(p, r)
        sim.step(200)
        sims[mode] = sim
    e64 = sims["hip_fp64"].energies()
    assert np.all(np.isfinite(e64))
    for mode in MODES[1:]:
        e = sims[mode].energies()
        de = float(np.max(np.abs(e - e64) / np.abs(e64)))
        fd = _field_diff(sims["hip_fp64"], sims[mode])
        print(f"   {mode}: energy drift {de:.2e}, worst field diff {max(fd.values()):.2e}",
              flush=True)
        assert de < C3_ENERGY_TOL[mode], (mode, de)
    for sim in sims.values():
        sim.close()


# --------------------------------------------------------------------- C4 --

def _c4_specs(nshell=64):
    rng = np.random.default_rng(64)
    Ls, Nrs = (31, 47, 63, 95), (32, 48, 64, 96)
    specs = []
    for q in range(nshell):
        L, nr = Ls[q % 4], Nrs[(q // 4) % 4]
        specs.append((L, nr, 3 * (L + 1) // 2, float(rng.uniform(50.0, 500.0))))
    return specs


@pytest.mark.timeout(300)
def test_c4_ensemble_matches_shells_alone(torch_cuda):
    torch = torch_cuda
    specs = _c4_specs()
    nsteps = 4

    def make(q, stream=None):
        L, nr, nth, ra = specs[q]
        p, r = mhd.dynamobench_case1(l_truncation=L, nth=nth, nphi=2 * nth, nri=nr - 1, Ra=ra)
        return _seeded_sim(p, r, stream, (1000 + q,) + SEEDED[1:])

    alone = []
    for q in range(len(specs)):
        if q % 16 == 0:
            print(f"[C4] shells alone {q}..{q + 15}", flush=True)
        sim = make(q)
        sim.step(nsteps)
        alone.append((sim.energies(), sim.get_field(0), sim.get_field(12)))
        sim.close()
    streams = [torch.cuda.Stream() for _ in range(4)]
    print("[C4] 64 shells together on 4 streams", flush=True)
    sims = [make(q, streams[q % 4].cuda_stream) for q in range(len(specs))]
    for _ in range(nsteps):
        for sim in sims:
            sim.step(1)
    torch.cuda.synchronize()
    for q, sim in enumerate(sims):
        e, u, t = alone[q]
        assert np.all(np.isfinite(e)), q
        assert np.array_equal(sim.energies(), e), q
        assert np.array_equal(sim.get_field(0), u), q
        assert np.array_equal(sim.get_field(12), t), q
        sim.close()


# --------------------------------------------------------------------- C5 --

# at L=511 the GPU's FP64 and int8-slice sums and the oracle's ordered
# 512-term loops differ by up to 3.4e-12 near the poles (m/sin(theta)
# terms): both are held to the 1e-11 bound; the FP32 and split-BF16x3
# bounds are about 4x what they reach (1.4e-6 / 8.7e-7)
C5_TOL = {"hip_fp64": 1e-11, "hip_fp32": 6e-6, "hip_bf16x3": 4e-6, "hip_int8x6": 1e-11}
C5_LEVELS = [0, 77]


@pytest.fixture(scope="module")
def c5_case():
    """L=511, nth=768, Nr=128, 2 vectors + 2 scalars: seeded inputs on two
    radial levels and the oracle's transforms of them (computed once)."""
    L, nth, nr, nv, ns = 511, 768, 128, 2, 2
    nc = 3 * nv + ns
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
    print(f"[C5] oracle transforms L={L} nth={nth} on levels {C5_LEVELS}", flush=True)
    grid = oracle.Grid(L, nth, radius[C5_LEVELS])
    J, M = (L + 1) ** 2, 2 * L + 1
    rng = np.random.default_rng(511)
    l = grid.jlm[:, 1].astype(float)
    sp_s = rng.standard_normal((len(C5_LEVELS), J, nc)) * ((l + 1.0) ** -1.5)[None, :, None]
    vr_s = rng.standard_normal((M, len(C5_LEVELS), nth, nc))
    want_b = oracle.leg_bwd(grid, nc, nv, ns, sp_s.ravel()).reshape(M, len(C5_LEVELS), nth, nc)
    want_f = oracle.leg_fwd(grid, nc, nv, ns, vr_s.ravel()).reshape(len(C5_LEVELS), J, nc)
    del grid
    return dict(L=L, nth=nth, nr=nr, nv=nv, ns=ns, nc=nc, radius=radius, J=J, M=M,
                sp_s=sp_s, vr_s=vr_s, want_b=want_b, want_f=want_f)


@pytest.mark.timeout(400)
@pytest.mark.parametrize("mode", MODES)
def test_c5_transforms_vs_oracle(torch_cuda, c5_case, mode):
    """The GPU transforms of the whole L=511 shell, checked against the
    oracle on radial levels 0 and 77 (the transforms act level by level)."""
    torch = torch_cuda
    c = c5_case
    ks, nc, nr, M, J = C5_LEVELS, c["nc"], c["nr"], c["M"], c["J"]
    print(f"[C5] transforms L={c['L']} nth={c['nth']} nr={nr} mode={mode}", flush=True)
    dev = torch.device("cuda:0")
    lt = calypso_amd.LegendreTransform(c["L"], c["nth"], c["radius"], nc, c["nv"], c["ns"], mode)
    sp = torch.zeros((nr, J, nc), dtype=torch.float64, device=dev)
    sp[ks] = torch.from_numpy(c["sp_s"]).to(dev)
    vr = torch.full((M, nr, c["nth"], nc), float("nan"), dtype=torch.float64, device=dev)
    lt.backward(sp.view(-1), vr.view(-1))
    got_b = vr[:, ks].cpu().numpy()
    assert not torch.isnan(vr).any().item()
    vr.zero_()
    vr[:, ks] = torch.from_numpy(c["vr_s"]).to(dev)
    lt.forward(vr.view(-1), sp.view(-1))
    got_f = sp[ks].cpu().numpy()
    lt.close()
    del sp, vr
    torch.cuda.synchronize()
    want_b, want_f = c["want_b"], c["want_f"]
    eb = max(np.abs(got_b[..., q] - want_b[..., q]).max() / np.abs(want_b[..., q]).max()
             for q in range(nc))
    ef = max(np.abs(got_f[..., q] - want_f[..., q]).max() / np.abs(want_f[..., q]).max()
             for q in range(nc))
    print(f"   {mode}: bwd err {eb:.2e}, fwd err {ef:.2e}", flush=True)
    assert eb < C5_TOL[mode] and ef < C5_TOL[mode], (eb, ef)


# per-field difference of the split-BF16x3 step from the FP64 step after 20
# steps: about 5x what it reaches (3.6e-7)
C5_STEP_FIELD_TOL = 2e-6


@pytest.mark.timeout(600)
def test_c5_step_20_steps(torch_cuda):
    """The L=511 shell (768 x 1536 grid, 129 radial points) stepped 20 times
    with FP64 and with split-BF16x3 Legendre transforms (FP64 LU solves and
    P_l^m recursion in both): finite energies, the per-field difference, and
    the HBM footprint of each state."""
    torch = torch_cuda
    sims, foot = {}, {}
    for mode in ("hip_fp64", "hip_bf16x3"):
        print(f"[C5] MHD step L=511 768x1536 Nr=129, {mode}", flush=True)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        p, r = mhd.dynamobench_case1(l_truncation=511, nth=768, nphi=1536, nri=128, dt=1.0e-5,
                                     coef_imp=0.5, legendre_mode=mode)
        sim = _seeded_sim(p, r)
        torch.cuda.synchronize()
        foot[mode] = free0 - torch.cuda.mem_get_info()[0]
        sim.step(20)
        e = sim.energies()
        print(f"   {mode}: footprint {foot[mode] / 2**30:.1f} GiB, energies {e}", flush=True)
        assert np.all(np.isfinite(e)) and sim.i_step == 20
        sims[mode] = sim
    fd = _field_diff(sims["hip_fp64"], sims["hip_bf16x3"],
                     [c for c in COMPARED if c[2] in (0, 2, 6, 8, 12, 20, 22, 26, 28)])
    de = np.max(np.abs(sims["hip_fp64"].energies() - sims["hip_bf16x3"].energies())
                / np.abs(sims["hip_fp64"].energies()))
    print(f"   bf16x3 vs fp64: worst field diff {max(fd.values()):.2e}, energy diff {de:.2e}",
          flush=True)
    assert max(fd.values()) < C5_STEP_FIELD_TOL, fd
    for mode in sims:
        assert foot[mode] < 200 * 2**30
        sims[mode].close()
