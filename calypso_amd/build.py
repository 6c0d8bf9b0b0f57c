"""Build the in-tree libraries for gfx950.

* ``calypso_amd/lib/libcalypso_amd.so``   HIP kernels + C ABI (hipcc)
* ``calypso_amd/lib/libcalypso_amd_f.so`` Fortran module shim (amdflang):
  ``legendre_trans_hip``, ``mhd_fft_select_hip``, ``lubksb_band_hip``,
  ``cal_nonlinear_sph_mhd_hip``,
  ``sph_analyzer_mhd_hip`` and ``calypso_amd_capi`` (.mod files beside it)
* ``calypso_amd/lib/leg_driver``, ``mhd_driver``, ``fft_lu_driver``  Fortran
  programs used by the tests to drive the shim end to end

Run ``python -m calypso_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")

HIP_SOURCES = ["legendre.hip", "leg_ozaki.hip", "fft.hip", "radial.hip", "mhd.hip", "scalar_shell.hip", "sph_io.cpp",
               "sph_grids.cpp", "capi.cpp"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    return out


def build_hip(incremental=False):
    """Compile each source to its own object in parallel (every kernel is
    launched from the file that defines it, so no relocatable device code is
    needed), then link the shared library.  incremental=True skips objects
    newer than their source and every header."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIB, exist_ok=True)
    obj_dir = os.path.join(LIB, "hobj")
    os.makedirs(obj_dir, exist_ok=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    cflags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objs = [os.path.join(obj_dir, s.rsplit(".", 1)[0] + ".o") for s in HIP_SOURCES]

    hdr_time = max(os.path.getmtime(os.path.join(d, f))
                   for d in (CSRC, os.path.join(ROOT, "include"))
                   for f in os.listdir(d) if f.endswith((".hpp", ".h")))

    def compile_one(i):
        src = os.path.join(CSRC, HIP_SOURCES[i])
        if (incremental and os.path.exists(objs[i])
                and os.path.getmtime(objs[i]) > max(os.path.getmtime(src), hdr_time)):
            return
        _run([hipcc, *cflags, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, HIP_SOURCES[i]), "-o", objs[i]])

    with ThreadPoolExecutor(max_workers=min(len(HIP_SOURCES), 6)) as ex:
        list(ex.map(compile_one, range(len(HIP_SOURCES))))
    out = os.path.join(LIB, "libcalypso_amd.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        print("amdflang not found: Fortran shim skipped", flush=True)
        return None
    obj = os.path.join(LIB, "fobj")
    os.makedirs(obj, exist_ok=True)
    srcs = ["calypso_amd_capi.f90", "legendre_trans_hip.f90", "sph_analyzer_mhd_hip.f90",
            "mhd_fft_select_hip.f90", "lubksb_band_hip.f90", "cal_nonlinear_sph_mhd_hip.f90"]
    for s in srcs:
        _run([fc, "-O2", "-fPIC", "-c", os.path.join(FSRC, s), "-module-dir", LIB,
              "-o", os.path.join(obj, s.replace(".f90", ".o"))])
    so = os.path.join(LIB, "libcalypso_amd_f.so")
    _run([fc, "-shared", "-o", so, *[os.path.join(obj, s.replace(".f90", ".o")) for s in srcs],
          "-L", LIB, "-lcalypso_amd", "-Wl,-rpath,$ORIGIN"])
    for prog in ("leg_driver", "mhd_driver", "fft_lu_driver", "scalar_shell_driver"):
        _run([fc, "-O2", os.path.join(FSRC, prog + ".f90"), "-I", LIB,
              "-o", os.path.join(LIB, prog), "-L", LIB, "-lcalypso_amd_f", "-lcalypso_amd",
              "-Wl,-rpath,$ORIGIN"])
    return so


def main():
    build_hip()
    build_fortran()


if __name__ == "__main__":
    sys.exit(main())
