"""The one-domain grid tables against the reference's own grid generator:
the files the library writes are byte for byte those of gen_sph_grids
(tests/golden/sph_grids, written by oracle/_ref/gen_sph_grids built from the
reference sources), and the tables in them fix the buffer orders the
C ABI assumes (host only, no GPU)."""
import gzip
import os
import re

import numpy as np
import pytest

from calypso_amd import grids, mhd

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sph_grids")
CASES = {"l7": (7, 12, 24, 4), "l5": (5, 8, 16, 3)}


def _gold(case, ext):
    with gzip.open(os.path.join(GOLD, case, f"in.0.{ext}.gz"), "rb") as f:
        return f.read()


def _section(text, title, ncol):
    """Integer rows (ncol values) after the comment block ending in `title`."""
    lines = text.decode().splitlines()
    i = next(n for n, l in enumerate(lines) if l.startswith("!") and title in l)
    while lines[i].startswith("!"):
        i += 1
    count = int(lines[i].split()[0])
    rows = [[int(x) for x in lines[i + 1 + q].split()] for q in range(count)]
    assert all(len(r) == ncol for r in rows)
    return np.array(rows)


@pytest.mark.parametrize("case", sorted(CASES))
def test_grid_files_byte_identical(tmp_path, case):
    L, nth, nphi, nri = CASES[case]
    r = mhd.chebyshev_radius(nri, 0.35 / (1.0 - 0.35), 1.0 / (1.0 - 0.35))
    grids.write_grid_files(tmp_path / "in", L, nth, nphi, r)
    for ext in ("rj", "rlm", "rtm", "rtp"):
        got = (tmp_path / f"in.0.{ext}").read_bytes()
        want = _gold(case, ext)
        if got != want:
            g, w = got.decode().splitlines(), want.decode().splitlines()
            bad = next(i for i in range(min(len(g), len(w))) if g[i] != w[i]) \
                if any(a != b for a, b in zip(g, w)) else min(len(g), len(w))
            pytest.fail(f"{ext}: first difference at line {bad + 1}: "
                        f"{g[bad] if bad < len(g) else '<eof>'!r} vs "
                        f"{w[bad] if bad < len(w) else '<eof>'!r}")


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_tables_fix_the_buffer_orders(case):
    """What the C ABI assumes about the reference's one-domain buffers, read
    from the generator's own tables: the rlm and rtm exchanges are the
    identity (WR/WS in node order), rlm nodes are j + J(k-1) with the modes
    in order m = -L..L, l = |m|..L, rtm nodes are t + nth(k-1) + nth nr(m+L),
    and the rj modes come in that same order, not l(l+1)+m."""
    L, nth, nphi, nri = CASES[case]
    nr, J, M = nri + 1, (L + 1) ** 2, 2 * L + 1
    rlm = _gold(case, "rlm")
    modes = _section(rlm, "on sphere surface", 3)
    order = grids.rj_mode_order(L)
    assert np.array_equal(modes[:, 0], order)
    assert np.array_equal(modes[:, 0], modes[:, 1] * (modes[:, 1] + 1) + modes[:, 2])
    assert np.array_equal(modes[:, 2], np.repeat(np.arange(-L, L + 1), L + 1 - np.abs(np.arange(-L, L + 1))))
    nodes = _section(rlm, "global radial ID and wavenumber ID", 3)
    k = np.repeat(np.arange(1, nr + 1), J)
    assert np.array_equal(nodes[:, 1], k) and np.array_equal(nodes[:, 2], np.tile(order, nr))
    assert np.array_equal(_section(rlm, "communication table between spectr", 1)[:, 0],
                          np.arange(1, nr * J + 1))
    rj = _gold(case, "rj")
    assert np.array_equal(_section(rj, "on sphere surface", 3)[:, 0], order)
    assert np.array_equal(_section(rj, "communication table between spectr", 1)[:, 0],
                          np.arange(1, nr * J + 1))
    rtm = _gold(case, "rtm")
    phi = _section(rtm, "phi direction", 2)
    assert np.array_equal(phi[:, 1], np.arange(-L, L + 1))
    nodes = _section(rtm, "global radial ID and grid ID", 4)
    t, kk, mm = np.meshgrid(np.arange(1, nth + 1), np.arange(1, nr + 1), np.arange(M),
                            indexing="ij")
    assert np.array_equal(nodes[:, 1], kk.transpose(2, 1, 0).ravel())
    assert np.array_equal(nodes[:, 2], t.transpose(2, 1, 0).ravel())
    assert np.array_equal(nodes[:, 3], phi[mm.transpose(2, 1, 0).ravel(), 0])
    assert np.array_equal(_section(rtm, "communication table between grid", 1)[:, 0],
                          np.arange(1, nth * nr * M + 1))


def test_rejects_grids_outside_the_generator(tmp_path):
    r = np.linspace(0.5, 1.5, 4)
    with pytest.raises(Exception, match="nphi = 2 nth"):
        grids.write_grid_files(tmp_path / "in", 7, 12, 30, r)
