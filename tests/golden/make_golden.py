"""Writes tests/golden/legendre_ref.npz from the reference's own compiled
routines (oracle/_ref/libcalypso_ref.so, built from the reference sources by
oracle/Makefile): Gauss points, Schmidt P_l^m, and the org-loop backward and
forward Legendre transforms on seeded inputs.

Run in the build container (the reference tree must be present to build the
library):  make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

CASES = [
    # name, L, nth, nr, nvec, nscl
    ("l7", 7, 12, 3, 3, 1),
    ("l10_odd", 10, 17, 2, 1, 2),
]


def spectrum(rng, grid, ncomp):
    """Gaussian coefficients with amplitude (l+1)^-1.5 per mode."""
    l = grid.jlm[:, 1].astype(np.float64)
    amp = (l + 1.0) ** -1.5
    x = rng.standard_normal((grid.nr, grid.J, ncomp))
    return (x * amp[None, :, None]).ravel()


def main():
    out = {}
    for name, L, nth, nr, nv, ns in CASES:
        ncomp = 3 * nv + ns
        rng = np.random.default_rng(1000 + L)
        radius = np.linspace(7.0 / 13.0, 20.0 / 13.0, nr)
        g = oracle.Grid(L, nth, radius)
        pt, wt, ct = oracle.gauss_points(nth, use_ref=True)
        ref = oracle.RefLegendre(g)
        sp = spectrum(rng, g, ncomp)
        vr = rng.standard_normal(g.n_rtm(ncomp))
        P = np.zeros((g.J, nth))
        dP = np.zeros((g.J, nth))
        j = 0
        for mp in range(2 * L + 1):
            m = abs(mp - L)
            for t in range(nth):
                p, d = oracle.schmidt_legendres_m(L, m, ct[t], use_ref=True)
                P[j:j + L - m + 1, t] = p[m:]
                dP[j:j + L - m + 1, t] = d[m:]
            j += L - m + 1
        out.update({
            f"{name}_meta": np.array([L, nth, nr, nv, ns]),
            f"{name}_radius": radius,
            f"{name}_point": pt, f"{name}_weight": wt, f"{name}_colat": ct,
            f"{name}_P": P, f"{name}_dP": dP,
            f"{name}_sp_in": sp, f"{name}_vr_out": ref.bwd(ncomp, nv, ns, sp),
            f"{name}_vr_in": vr, f"{name}_sp_out": ref.fwd(ncomp, nv, ns, vr),
        })
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "legendre_ref.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, sum(v.size for v in out.values()), "values")


if __name__ == "__main__":
    main()
