"""Benchmark: standalone Legendre transform pair on one MI355X.

Workload (the headline configuration): Lmax=255, nth=384 Gauss points,
Nr=96 radial points, 12 field components (4 vectors), seeded Gaussian
spectra with amplitude (l+1)^-1.5.  One step = one backward + one forward
transform, inputs resident in HBM.  Every precision mode runs; each mode's
output is compared with the FP64 GPU path, and the FP64 path with the
reference's own org-loop transforms (oracle/_ref, compiled from the reference
sources) on a bounded sample of radial levels, timed on the host cores as the
CPU baseline.  The full MHD step (Dynamobench case 1 physics) is timed as
well at three sizes ("mhd_step"), at L=127 with each precision of the
Legendre transforms and the energy difference to the FP64 step
("mhd_step_L127_precisions"), an ensemble of 64 uneven shells is stepped
together ("ensemble"), and the large shell Lmax=511, nth=768, Nr=128 with 8
components is transformed in each precision ("legendre_L511").

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--modes ...]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

L, NTH, NR, NV, NS = 255, 384, 96, 4, 0
NCOMP = 3 * NV + NS
ERR_BOUND = 1e-11
PEAK = {"hip_fp64": 78.6, "hip_fp32": 157.3, "hip_bf16x3": 2500.0,
        "hip_int8x6": 5000.0}   # TFLOP/s (TOP/s for int8), dense
# MFMA operations per algorithmic FP64 operation: six BF16 slice products;
# 28 int8 digit-pair products (seven digits per operand, pairs i+j <= 6)
OPS_PER_FLOP = {"hip_fp64": 1, "hip_fp32": 1, "hip_bf16x3": 6, "hip_int8x6": 28}
HBM_PEAK = 8000.0   # GB/s


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "legendre_L255_nth384_nr96_c12_traffic.md")


def kernel_name(direction, mode, plan):
    """Template instance that runs a plan (as rocprofv3 names it)."""
    if mode == "hip_int8x6":
        return (f"k_leg_{direction}z<6, {plan['rows_per_wave']}, {plan['mct1']}, "
                f"{plan['mct2']}>")
    t = {"hip_fp64": "TF64", "hip_fp32": "TF32", "hip_bf16x3": "TBF3"}[mode]
    if plan["kernel_form"] == 3:   # third form: waves per work-group last
        return (f"k_leg_{direction}3<{t}, {plan['rows_per_wave']}, {plan['mct1']}, "
                f"{plan['mct2']}, {plan['prefetch']}, {plan.get('waves', 4)}>")
    return (f"k_leg_{direction}<{t}, {plan['rows_per_wave']}, {plan['mct1']}, "
            f"{plan['mct2']}, {plan['prefetch']}>")


def measured_traffic(name):
    """HBM bytes per launch of `name` from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by a 1 GiB copy), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic: ") and "=" in line:
                k, v = line[len("traffic: "):].rsplit("=", 1)
                if k.strip() == name:
                    return float(v)
    except OSError:
        pass
    return None


def legendre_flops(L, nth, nr, nvec, nscl):
    """Symmetric-form work of one direction: 2 th_h Nr (L+1)^2 (5 nvec + nscl)."""
    th_h = (nth + 1) // 2
    return 2.0 * th_h * nr * (L + 1) ** 2 * (5 * nvec + nscl)


def rlm_degrees(L):
    """Degree l of each mode in rlm order (m = -L..L, then l = |m|..L)."""
    return np.concatenate([np.arange(abs(m), L + 1) for m in range(-L, L + 1)]).astype(np.float64)


def make_spectrum(torch, J, seed=1234):
    l = torch.from_numpy(rlm_degrees(L)).cuda()
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn((NR, J, NCOMP), generator=gen, dtype=torch.float64, device="cuda")
    return (x * ((l + 1.0) ** -1.5)[None, :, None]).reshape(-1).contiguous()


def field_err(got, ref, ncomp):
    g = got.reshape(-1, ncomp)
    r = ref.reshape(-1, ncomp)
    s = r.abs().amax(dim=0).clamp_min(1e-300)
    return ((g - r).abs().amax(dim=0) / s).max().item()


def time_mode(torch, calypso_amd, mode, radius, sp, steps, warmup):
    lt = calypso_amd.LegendreTransform(L, NTH, radius, NCOMP, NV, NS, mode)
    vr = torch.empty(lt.n_rtm(), dtype=torch.float64, device="cuda")
    sp2 = torch.empty_like(sp)
    for _ in range(warmup):
        lt.backward(sp, vr)
        lt.forward(vr, sp2)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1, e2 in evs:
        e0.record()
        lt.backward(sp, vr)
        e1.record()
        lt.forward(vr, sp2)
        e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    bwd_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in evs]))
    fwd_ms = float(np.mean([b.elapsed_time(c) for _, b, c in evs]))
    info = lt.info()
    info["plan_bwd"] = lt.plan_info(False)
    info["plan_fwd"] = lt.plan_info(True)
    lt.close()
    return {"wall_s": wall, "bwd_ms": bwd_ms, "fwd_ms": fwd_ms, "vr": vr, "sp2": sp2,
            "info": info}


def cpu_baseline(calypso_amd, radius, sp_host, vr_gpu_host, nsample, threads):
    """Reference org loops (oracle/_ref) on `nsample` radial levels."""
    import oracle
    rs = radius[:nsample]
    grid = oracle.Grid(L, NTH, rs)
    J = (L + 1) ** 2
    sp_s = sp_host.reshape(NR, J, NCOMP)[:nsample].ravel()
    kind = "reference" if oracle.ref_available() else "port"
    variants = {}
    if kind == "reference":
        # the reference's org loops and its SYMMETRIC_MATMUL transforms (the
        # variant its own tests select); the faster pair is the baseline, as
        # the reference's own init-time selection keeps the fastest variant
        # (select_legendre_transform, sph_transforms_4_MHD.f90:309-426)
        ref = oracle.RefLegendre(grid, nthreads=threads)
        for name, fb, ff in (("org", ref.bwd, ref.fwd), ("sym_matmul", ref.bwd_sym, ref.fwd_sym)):
            t0 = time.perf_counter()
            vr_v = fb(NCOMP, NV, NS, sp_s)
            t1 = time.perf_counter()
            ff(NCOMP, NV, NS, vr_v)
            t2 = time.perf_counter()
            variants[name] = (t1 - t0, t2 - t1, vr_v)
    else:
        threads = 1
        t0 = time.perf_counter()
        vr_v = oracle.leg_bwd(grid, NCOMP, NV, NS, sp_s)
        t1 = time.perf_counter()
        oracle.leg_fwd(grid, NCOMP, NV, NS, vr_v)
        t2 = time.perf_counter()
        variants["port"] = (t1 - t0, t2 - t1, vr_v)
    best = min(variants, key=lambda k: variants[k][0] + variants[k][1])
    tb, tf, vr_c = variants[best]
    sec = (tb + tf) * (NR / nsample)
    # GPU FP64 output on the same levels vs the reference
    M = 2 * L + 1
    g = vr_gpu_host.reshape(M, NR, NTH, NCOMP)[:, :nsample]
    c = vr_c.reshape(M, nsample, NTH, NCOMP)
    errs = [np.max(np.abs(g[..., nd] - c[..., nd])) / np.max(np.abs(c[..., nd]))
            for nd in range(NCOMP)]
    timings = ", ".join(f"{k} bwd {v[0]:.2f}s + fwd {v[1]:.2f}s" for k, v in variants.items())
    return {"value": 1.0 / sec, "unit": "steps/s", "cores": threads, "kind": kind,
            "variant": best,
            "sample": f"{nsample} of {NR} radial levels, scaled x{NR / nsample:.0f} ({timings}; "
                      f"np_smp={threads}); the faster variant is the value",
            "gpu_fp64_err_vs_cpu": float(max(errs))}


MHD_CONFIGS = [   # (name, Lmax, nth, nphi, radial intervals, timed steps)
    ("dynamobench_case1_L31", 31, 48, 96, 48, 200),
    ("shell_L127_Nr65", 127, 192, 384, 64, 40),
    ("shell_L255_Nr97", 255, 384, 768, 96, 8),
]


def mhd_bench(torch):
    """Full MHD step (Dynamobench case 1 physics) at three sizes, FP64."""
    from calypso_amd import mhd
    out = {}
    for name, L_, nth, nphi, nri, steps in MHD_CONFIGS:
        p, r = mhd.dynamobench_case1(l_truncation=L_, nth=nth, nphi=nphi, nri=nri)
        sim = mhd.MhdSimulation(p, r, stream=torch.cuda.current_stream().cuda_stream)
        sim.set_dynamobench_ic(400)
        sim.initialize()
        sim.step(3)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        e = sim.energies()
        sim.close()
        out[name] = {"Lmax": L_, "nth": nth, "nphi": nphi, "Nr": nri + 1, "steps": steps,
                     "ms_per_step": ms, "steps_per_s": 1e3 / ms,
                     "energies_finite": bool(np.all(np.isfinite(e)))}
        print(f"# mhd {name}: {ms:.3f} ms/step", file=sys.stderr, flush=True)
    return out


def mhd_precision_bench(torch, steps=200):
    """The L=127, Nr=65 step with each precision of the Legendre transforms
    (FP64 radial solves, P_l^m recursion and diagnostics in every mode):
    ms/step and the energies after `steps` steps against the FP64 step."""
    from calypso_amd import mhd
    out, ref = {}, None
    for mode in ("hip_fp64", "hip_fp32", "hip_bf16x3", "hip_int8x6"):
        p, r = mhd.dynamobench_case1(l_truncation=127, nth=192, nphi=384, nri=64,
                                     legendre_mode=mode)
        sim = mhd.MhdSimulation(p, r, stream=torch.cuda.current_stream().cuda_stream)
        sim.set_dynamobench_ic(400)
        sim.initialize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.step(steps)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        e = sim.energies()
        sim.close()
        if ref is None:
            ref = e
        out[mode] = {"ms_per_step": ms, "steps_per_s": 1e3 / ms, "energies": [float(x) for x in e],
                     "max_rel_energy_diff_vs_fp64": float(np.max(np.abs(e - ref) / np.abs(ref)))}
        print(f"# mhd L127 {mode}: {ms:.3f} ms/step, energy diff "
              f"{out[mode]['max_rel_energy_diff_vs_fp64']:.2e}", file=sys.stderr, flush=True)
    return {"Lmax": 127, "nth": 192, "nphi": 384, "Nr": 65, "steps": steps, "dt": 5e-5,
            "graph_replayed": True, "modes": out}


def large_legendre_bench(torch, steps=3):
    """Large shell: Lmax=511, nth=768, Nr=128, 8 components (2 vectors + 2
    scalars), one bwd+fwd pair per step in each precision: time, per-field
    error against the FP64 path and the device footprint."""
    import calypso_amd
    L5, nth5, nr5, nv5, ns5 = 511, 768, 128, 2, 2
    nc5 = 3 * nv5 + ns5
    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr5)
    J5 = (L5 + 1) ** 2
    l = torch.from_numpy(np.concatenate([np.arange(abs(m), L5 + 1) for m in range(-L5, L5 + 1)])
                         .astype(np.float64)).cuda()
    gen = torch.Generator(device="cuda").manual_seed(511)
    sp = (torch.randn((nr5, J5, nc5), generator=gen, dtype=torch.float64, device="cuda")
          * ((l + 1.0) ** -1.5)[None, :, None]).reshape(-1).contiguous()
    out, ref = {}, None
    for mode in ("hip_fp64", "hip_fp32", "hip_bf16x3", "hip_int8x6"):
        lt = calypso_amd.LegendreTransform(L5, nth5, radius, nc5, nv5, ns5, mode)
        vr = torch.empty(lt.n_rtm(), dtype=torch.float64, device="cuda")
        sp2 = torch.empty_like(sp)
        lt.backward(sp, vr)
        lt.forward(vr, sp2)
        torch.cuda.synchronize()
        e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        bw = fw = 0.0
        for _ in range(steps):
            e0.record()
            lt.backward(sp, vr)
            e1.record()
            lt.forward(vr, sp2)
            e2.record()
            torch.cuda.synchronize()
            bw += e0.elapsed_time(e1)
            fw += e1.elapsed_time(e2)
        info = lt.info()
        lt.close()
        r = {"bwd_ms": bw / steps, "fwd_ms": fw / steps,
             "legendre_tflops": 2 * legendre_flops(L5, nth5, nr5, nv5, ns5)
             / ((bw + fw) / steps * 1e-3) / 1e12,
             # HBM footprint: P_l^m tables of the mode + spectra in, grid
             # values, spectra out (FP64)
             "table_bytes": info["table_bytes"],
             "field_bytes": int(8 * (2 * sp.numel() + vr.numel()))}
        if ref is None:
            ref = (vr.clone(), sp2.clone())
        else:
            r["err_vs_fp64_bwd"] = field_err(vr, ref[0], nc5)
            r["err_vs_fp64_fwd"] = field_err(sp2, ref[1], nc5)
        out[mode] = r
        del vr, sp2
        print(f"# L511 {mode}: bwd {r['bwd_ms']:.2f} ms fwd {r['fwd_ms']:.2f} ms", file=sys.stderr,
              flush=True)
    return {"Lmax": L5, "nth": nth5, "Nr": nr5, "components": nc5, "vectors": nv5,
            "scalars": ns5, "modes": out}


def ensemble_bench(torch, nshell=64, steps=20, nstream=4, balanced=True):
    """Ensemble of independent shells, Lmax in {31,47,63,95}, Nr in
    {32,48,64,96}, nth = 1.5(L+1), nphi = 2 nth, seeded Ra in [50, 500]: each
    shell's step is a captured graph; the shells are dealt over `nstream`
    streams (default priority; four, the hardware queues a process gets)
    longest first, each to the stream with the
    least work so far (work of a shell ~ Nr nth (L+1)^2, the Legendre
    transforms), and stepped together.  Reported: shell-steps per second of
    the ensemble and of the same shells stepped one after another."""
    from calypso_amd import mhd
    rng = np.random.default_rng(64)
    Ls, Nrs = (31, 47, 63, 95), (32, 48, 64, 96)
    streams = [torch.cuda.Stream() for _ in range(nstream)]
    specs = []
    for q in range(nshell):
        L_, nr_ = Ls[q % 4], Nrs[(q // 4) % 4]
        specs.append((L_, nr_, 3 * (L_ + 1) // 2, float(rng.uniform(50.0, 500.0))))
    load = [0.0] * nstream
    home = [0] * nshell
    for q in sorted(range(nshell), key=lambda q: -specs[q][1] * specs[q][2] * (specs[q][0] + 1) ** 2):
        L_, nr_, nth, _ = specs[q]
        home[q] = min(range(nstream), key=lambda i: load[i]) if balanced else q % nstream
        load[home[q]] += nr_ * nth * (L_ + 1) ** 2
    sims, shells = [], []
    for q in range(nshell):
        L_, nr_, nth, ra = specs[q]
        p, r = mhd.dynamobench_case1(l_truncation=L_, nth=nth, nphi=2 * nth, nri=nr_ - 1)
        p.coef_buo = (1.0 - 0.35) * ra / 1.0e-3   # (1 - r_i/r_o) Ra / E, as dynamobench_case1
        sim = mhd.MhdSimulation(p, r, stream=streams[home[q]].cuda_stream)
        sim.set_dynamobench_ic(400)
        sim.initialize()
        sim.step(2)   # Euler start, then the captured Adams-Bashforth step
        sims.append(sim)
        shells.append({"Lmax": L_, "Nr": nr_, "nth": nth, "nphi": 2 * nth, "Ra": ra})
    torch.cuda.synchronize()
    # one after another: each shell alone for `steps` steps
    t_seq = 0.0
    for sim in sims:
        t0 = time.perf_counter()
        sim.step(steps)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
    # together
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim in sims:
            sim.step(1)
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - t0
    finite = all(bool(np.all(np.isfinite(sim.energies()))) for sim in sims)
    for sim in sims:
        sim.close()
    res = {"shells": nshell, "streams": nstream, "steps": steps,
           "shell_steps_per_s_ensemble": nshell * steps / t_ens,
           "shell_steps_per_s_sequential": nshell * steps / t_seq,
           "speedup": t_seq / t_ens, "energies_finite": finite,
           "sizes": sorted({(d["Lmax"], d["Nr"]) for d in shells})}
    print(f"# ensemble {nshell} shells: {res['shell_steps_per_s_ensemble']:.0f} shell-steps/s "
          f"(sequential {res['shell_steps_per_s_sequential']:.0f})", file=sys.stderr, flush=True)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--modes", default="hip_fp64,hip_fp32,hip_bf16x3,hip_int8x6")
    ap.add_argument("--cpu-sample", type=int, default=96)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mhd", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the Lmax=511 transforms")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the L=127 precision comparison and the ensemble")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 1 GiB device copy (reference for the byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): this build "
              f"runs the step on a single GPU and has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    import torch
    import calypso_amd

    radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, NR)
    J = (L + 1) ** 2
    sp = make_spectrum(torch, J)
    flops = legendre_flops(L, NTH, NR, NV, NS)
    modes = [m for m in args.modes.split(",") if m]
    res = {}
    ref_vr = ref_sp = None
    for mode in modes:
        r = time_mode(torch, calypso_amd, mode, radius, sp, args.steps, args.warmup)
        if mode == "hip_fp64":
            ref_vr, ref_sp = r["vr"].clone(), r["sp2"].clone()
        res[mode] = r
        print(f"# {mode}: bwd {r['bwd_ms']:.3f} ms  fwd {r['fwd_ms']:.3f} ms", file=sys.stderr,
              flush=True)
    out_modes = {}
    for mode, r in res.items():
        step_ms = r["wall_s"] * 1e3 / args.steps
        e = {"steps_per_s": 1e3 / step_ms, "ms_per_step": step_ms,
             "bwd_ms": r["bwd_ms"], "fwd_ms": r["fwd_ms"],
             "legendre_tflops": 2 * flops / ((r["bwd_ms"] + r["fwd_ms"]) * 1e-3) / 1e12,
             "plan_bwd": r["info"]["plan_bwd"], "plan_fwd": r["info"]["plan_fwd"]}
        if ref_vr is not None:
            e["err_vs_fp64_bwd"] = field_err(r["vr"], ref_vr, NCOMP)
            e["err_vs_fp64_fwd"] = field_err(r["sp2"], ref_sp, NCOMP)
        out_modes[mode] = e
    # headline: fastest mode meeting the error bound against the FP64 GPU path
    # (FP64 always qualifies: it is the reference the bound is measured against)
    ok = [m for m in out_modes if m == "hip_fp64" or
          max(out_modes[m].get("err_vs_fp64_bwd", 1), out_modes[m].get("err_vs_fp64_fwd", 1)) <= ERR_BOUND]
    head = max(ok, key=lambda m: out_modes[m]["steps_per_s"]) if ok else modes[0]
    h = out_modes[head]
    dom = "bwd" if h["bwd_ms"] >= h["fwd_ms"] else "fwd"
    dom_ms = h[f"{dom}_ms"]
    achieved = flops * OPS_PER_FLOP[head] / (dom_ms * 1e-3) / 1e12
    dom_kernel = kernel_name(dom, head, h[f"plan_{dom}"])
    line = {
        "metric": "steps/s + Legendre TFLOP/s, mixed-prec vs FP64 GPU, Lmax=255 Nr=96, err<=1e-11",
        "value": h["steps_per_s"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"hip_fp64": "f64", "hip_fp32": "f32", "hip_bf16x3": "bf16x3",
                  "hip_int8x6": "int8x6 (FP64-exact sums)"}[head],
        "data": "synthetic (seeded Gaussian spectra, amplitude (l+1)^-1.5)",
        "config": {"workload": "legendre bwd+fwd pair", "Lmax": L, "nth": NTH, "Nr": NR,
                   "components": NCOMP, "vectors": NV, "scalars": NS, "headline_mode": head,
                   "err_bound": ERR_BOUND},
        "legendre_tflops": h["legendre_tflops"],
        "modes": out_modes,
        "roofline": {"kernel": dom_kernel, "bound": "mfma", "achieved": achieved,
                     "peak": PEAK[head], "unit": "TOP/s" if head == "hip_int8x6" else "TFLOP/s",
                     "frac": achieved / PEAK[head],
                     "traffic": measured_traffic(dom_kernel)},
    }
    if args.copy_calibration:
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
    if not args.no_mhd:
        line["mhd_step"] = mhd_bench(torch)
        if not args.no_extra:
            line["mhd_step_L127_precisions"] = mhd_precision_bench(torch)
            line["ensemble"] = ensemble_bench(torch)
    if not args.no_large:
        line["legendre_L511"] = large_legendre_bench(torch)
    if not args.no_cpu:
        vr_host = ref_vr.cpu().numpy() if ref_vr is not None else res[modes[0]]["vr"].cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(calypso_amd, radius, sp.cpu().numpy(), vr_host,
                                            args.cpu_sample, args.cpu_threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
