"""TEST INFRASTRUCTURE: builds the reference's own MHD program
(programs/SPH_MHD/main_sph_MHD_w_psf.f90, the `sph_mhd` of its tests) from its
Fortran sources with amdflang, into oracle/_ref/sph_mhd.

The sources stay where they lie under the reference tree.  The compile order
is the module dependency order read from their `module` / `use` statements;
`use` statements inside `#ifdef FFTW3` blocks are not followed, as the
reference's own build without FFTW does not take them.  FFTPACK 5.1d is the
reference's vendored copy (External_libs/FFTPACK5.1D).  As in the reference's
own build, the objects go into a static archive and the program is linked
against it, so only what the program reaches is linked.  MPI comes from the
MPICH under /opt/conda.  Nothing here is part of the product library; the
program is run only by tests/golden/make_case2_steps.py to write fixtures.

Usage:  python oracle/build_ref_mhd.py [REFERENCE_ROOT]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import build_ref_grids as grids  # noqa: E402

OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "mhd_obj")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "src")
    prog = os.path.join(src, "programs", "SPH_MHD")
    main_src = os.path.join(prog, "main_sph_MHD_w_psf.f90")
    if not os.path.exists(main_src):
        sys.exit(f"reference program {main_src} not found")
    mod2file, uses = grids.module_index(src)
    for f in ("SPH_analyzer_MHD.f90", "analyzer_sph_MHD_w_psf.f90"):
        p = os.path.join(prog, f)
        for m in re.findall(r"(?im)^\s*module\s+(\w+)\s*$", open(p).read()):
            mod2file[m.lower()] = p
    for path in list(uses):
        text = open(path, errors="ignore").read()
        if "#ifdef FFTW3" in text:
            text = re.sub(r"(?s)#ifdef FFTW3.*?#(else|endif)", "", text)
            uses[path] = sorted({u.lower() for u in re.findall(r"(?im)^\s*use\s+(\w+)", text)})
    order = grids.compile_order(mod2file, uses, [main_src])
    order.insert(0, os.path.join(src, "External_libs", "FFTPACK5.1D", "fftpack5.1d.f90"))
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for f in order:
        o = os.path.join(OBJ, os.path.basename(f).rsplit(".", 1)[0] + ".o")
        if not (os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(f)):
            subprocess.run([grids.FC, "-O2", "-w", "-cpp", "-I", grids.MPI_INC, "-c", f, "-o", o],
                           cwd=OBJ, check=True)
        objs.append(o)
    main_obj = objs.pop()
    lib = os.path.join(OBJ, "libref_mhd.a")
    if os.path.exists(lib):
        os.remove(lib)
    subprocess.run(["ar", "rcs", lib, *objs], check=True)
    exe = os.path.join(OUT, "sph_mhd")
    subprocess.run([grids.FC, "-o", exe, main_obj, lib, "-L", grids.MPI_LIB, "-lmpifort",
                    "-lmpi", f"-Wl,-rpath,{grids.MPI_LIB}"], check=True)
    print(f"built {exe} from {len(order)} reference sources")


if __name__ == "__main__":
    main()
