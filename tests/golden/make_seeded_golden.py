"""Writes tests/golden/seeded_c1_energies.json: the oracle's MHD step
(oracle/mhd_step.py, pinned against the reference's Dynamobench_case1
energies) from the seeded random start of the L=31 shell configuration --
Nth=48, Nphi=96, 49 radial points (48 Chebyshev intervals), E=1e-3, Ra=100,
Pr=1, Pm=5, r_i/r_o=0.35, dt=1e-4, Crank-Nicolson weight 0.5 -- energies
(sph_pwr_volume_s.dat columns) after 0, 50 and 100 steps.

Run:  make -C oracle oracle && python tests/golden/make_seeded_golden.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import mhd_step  # noqa: E402

SEED, AMP_U, AMP_B, AMP_T = 2024, 0.3, 0.03, 0.1


def main():
    case = mhd_step.shell_case(L=31, nth=48, nphi=96, nri=48, dt=1.0e-4, coef_imp=0.5)
    ref = mhd_step.MhdStep(case)
    ref.initialize(seeded=(SEED, AMP_U, AMP_B, AMP_T))
    steps = {"0": [float(x) for x in ref.energies()]}
    for n in range(1, 101):
        ref.step()
        if n % 50 == 0:
            steps[str(n)] = [float(x) for x in ref.energies()]
            print(n, steps[str(n)], flush=True)
    out = {"source": "oracle/mhd_step.py, seeded start (make_seeded_golden.py)",
           "config": {"Lmax": 31, "nth": 48, "nphi": 96, "nri": 48, "dt": 1.0e-4,
                      "coef_imp": 0.5, "E": 1e-3, "Ra": 100.0, "Pr": 1.0, "Pm": 5.0},
           "seeded": {"seed": SEED, "amp_u": AMP_U, "amp_b": AMP_B, "amp_t": AMP_T},
           "columns": ["K_ene_pol", "K_ene_tor", "K_ene", "temperature",
                       "M_ene_pol", "M_ene_tor", "M_ene"],
           "steps": steps}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seeded_c1_energies.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
