"""Writes tests/golden/sph_grids/<case>/in.0.{rj,rlm,rtm,rtp}.gz: the
single-domain index and communication tables written by the reference's own
grid generator gen_sph_grids (programs/mesh_utilities/
main_generate_sph_grids.f90), built from the reference sources by
`make -C oracle grids` into oracle/_ref/gen_sph_grids, for a Chebyshev shell
(r_i/r_o = 0.35, r_o - r_i = 1, no inner-core points).

Run in the build container:  make -C oracle grids && python tests/golden/make_sph_grids.py
"""
import gzip
import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "gen_sph_grids")

# case: (truncation, meridional points, zonal points, Chebyshev intervals)
CASES = {"l7": (7, 12, 24, 4), "l5": (5, 8, 16, 3)}

CONTROL = """begin spherical_shell_ctl
  begin data_files_def
    num_subdomain_ctl    1
    sph_file_prefix             'grid/in'
    excluding_FEM_mesh_ctl      'YES'
  end data_files_def
!
  begin num_domain_ctl
    num_radial_domain_ctl         1
    num_horizontal_domain_ctl     1
  end num_domain_ctl
!
  begin num_grid_sph
    truncation_level_ctl      {L}
!
    ngrid_meridonal_ctl      {nth}
    ngrid_zonal_ctl          {nphi}
!
    radial_grid_type_ctl   Chebyshev
    num_fluid_grid_ctl     {nri}
    fluid_core_size_ctl   1.0
    ICB_to_CMB_ratio_ctl  0.35
  end num_grid_sph
end spherical_shell_ctl
"""


def main():
    for name, (L, nth, nphi, nri) in CASES.items():
        with tempfile.TemporaryDirectory() as tmp:
            os.makedirs(os.path.join(tmp, "grid"))
            with open(os.path.join(tmp, "control_sph_shell"), "w") as f:
                f.write(CONTROL.format(L=L, nth=nth, nphi=nphi, nri=nri))
            subprocess.run([EXE], cwd=tmp, check=True, stdout=subprocess.DEVNULL)
            out = os.path.join(HERE, "sph_grids", name)
            os.makedirs(out, exist_ok=True)
            shutil.copy(os.path.join(tmp, "control_sph_shell"), out)
            for ext in ("rj", "rlm", "rtm", "rtp"):
                data = open(os.path.join(tmp, "grid", f"in.0.{ext}"), "rb").read()
                with gzip.GzipFile(os.path.join(out, f"in.0.{ext}.gz"), "wb", mtime=0) as g:
                    g.write(data)
        print(name, "done")


if __name__ == "__main__":
    main()
