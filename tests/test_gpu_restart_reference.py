"""A restart file written by the reference's own MHD program (built from its
sources) after step 10 of Dynamobench case 1's physics on a small grid
(L = 7, 12 x 24, 13 radial points; tests/golden/restart_resume_l7, written by
tests/golden/make_restart_resume.py) read by the library in the reference's
rj order: the resumed run (Adams-Bashforth from the file's previous terms)
continues the reference's run, energies after steps 11-20 within 1e-12, and
the restart it writes after step 20 holds what the reference's own restart
of step 20 holds: the same header, field names, component counts and layout,
values within 1e-10 of each column's largest."""
import gzip
import json
import os

import numpy as np
import pytest

from calypso_amd import grids, mhd

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden", "restart_resume_l7")


def test_resume_from_reference_written_restart(torch_cuda, tmp_path):
    ref = json.load(open(os.path.join(HERE, "energies.json")))
    g = ref["grid"]
    print(f"[restart] reference-written restart at step {ref['restart_step']}, L={g['l_truncation']}",
          flush=True)
    path = tmp_path / "rst.1.0.fst"
    path.write_bytes(gzip.open(os.path.join(HERE, ref["restart"])).read())
    p, r = mhd.dynamobench_case1(l_truncation=g["l_truncation"], nth=g["nth"], nphi=g["nphi"],
                                 nri=g["nri"])
    sim = mhd.MhdSimulation(p, r)
    sim.set_rj_order(grids.rj_mode_order(g["l_truncation"]))
    sim.read_restart(str(path))
    sim.initialize()
    assert sim.i_step == ref["restart_step"]
    worst = 0.0
    for n in range(ref["restart_step"] + 1, 21):
        sim.step(1)
        want = np.array(ref["steps"][str(n)])
        worst = max(worst, float(np.max(np.abs(sim.energies() / want - 1.0))))
    print(f"[restart] steps 11-20: worst relative energy difference {worst:.1e}", flush=True)
    assert worst < 1e-12, worst
    mine = read_field_file(sim.write_restart(str(tmp_path / "out"), 2))
    sim.close()
    theirs = read_field_file(gzip.open(os.path.join(HERE, ref["restart_end"])).read().decode())
    assert mine["head"] == theirs["head"]
    assert [(n, a.shape) for n, a in mine["fields"]] == [(n, a.shape) for n, a in theirs["fields"]]
    for (name, a), (_, b) in zip(mine["fields"], theirs["fields"]):
        for c in range(a.shape[1]):
            scale = max(np.abs(b[:, c]).max(), 1e-300)
            err = np.abs(a[:, c] - b[:, c]).max() / scale
            assert err < 1e-10, (name, c, err)


def read_field_file(src):
    """Header (domain, step, time and dt, node and field counts, component
    counts) and the (nnod, ncomp) blocks of an ASCII field file; src is a
    path or the text."""
    text = open(src).read() if os.path.exists(str(src)) else src
    rows = [ln for ln in text.splitlines() if ln.strip() and not ln.lstrip().startswith("!")]
    dom, step = int(rows[0]), int(rows[1])
    time, dt = (float(x) for x in rows[2].split())
    nnod, nfld = (int(x) for x in rows[3].split())
    i, ncomp = 4, []
    while len(ncomp) < nfld:
        ncomp += [int(x) for x in rows[i].split()]
        i += 1
    fields = []
    for nc in ncomp:
        name = rows[i].strip()
        block = np.array([[float(x) for x in rows[i + 1 + k].split()] for k in range(nnod)])
        fields.append((name, block.reshape(nnod, nc)))
        i += 1 + nnod
    return {"head": (dom, step, round(time, 12), dt, nnod, nfld, tuple(ncomp)), "fields": fields}
