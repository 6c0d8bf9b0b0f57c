"""TEST INFRASTRUCTURE: builds the reference's own grid generator
(programs/mesh_utilities/main_generate_sph_grids.f90, the `gen_sph_grids`
program) from its Fortran sources with amdflang, into oracle/_ref/.

The sources stay where they lie under the reference tree; the compile order
is the module dependency order found by reading their `module` / `use`
statements (every module the program needs exists in the reference
sources).  MPI comes from the MPICH under /opt/conda, as for the reference's
own build.  Nothing here is part of the product library.

Usage:  python oracle/build_ref_grids.py [REFERENCE_ROOT]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "grid_obj")
FC = os.environ.get("FORTRAN", "amdflang")
MPI_INC, MPI_LIB = "/opt/conda/include", "/opt/conda/lib"


def module_index(src_root):
    mod2file, uses = {}, {}
    for d, _, files in os.walk(src_root):
        for f in sorted(files):
            if not f.endswith((".f90", ".F90")):
                continue
            p = os.path.join(d, f)
            text = open(p, errors="ignore").read()
            for m in re.findall(r"(?im)^\s*module\s+(\w+)\s*$", text):
                mod2file.setdefault(m.lower(), p)
            uses[p] = sorted({u.lower() for u in re.findall(r"(?im)^\s*use\s+(\w+)", text)})
    return mod2file, uses


def compile_order(mod2file, uses, roots):
    seen, order = set(), []

    def visit(path):
        if path in seen:
            return
        seen.add(path)
        for u in uses[path]:
            if u in mod2file:
                visit(mod2file[u])
        order.append(path)

    for r in roots:
        visit(r)
    return order


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "src")
    main_src = os.path.join(src, "programs", "mesh_utilities", "main_generate_sph_grids.f90")
    if not os.path.exists(main_src):
        sys.exit(f"reference program {main_src} not found")
    mod2file, uses = module_index(src)
    order = compile_order(mod2file, uses, [main_src])
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for f in order:
        o = os.path.join(OBJ, os.path.basename(f).rsplit(".", 1)[0] + ".o")
        if not (os.path.exists(o) and os.path.getmtime(o) > os.path.getmtime(f)):
            subprocess.run([FC, "-O1", "-w", "-cpp", "-I", MPI_INC, "-c", f, "-o", o],
                           cwd=OBJ, check=True)
        objs.append(o)
    exe = os.path.join(OUT, "gen_sph_grids")
    subprocess.run([FC, "-o", exe, *objs, "-L", MPI_LIB, "-lmpifort", "-lmpi",
                    f"-Wl,-rpath,{MPI_LIB}"], check=True)
    print(f"built {exe} from {len(order)} reference sources")


if __name__ == "__main__":
    main()
