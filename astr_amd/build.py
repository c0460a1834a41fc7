"""Build, in-tree: the HIP library astr_amd/libastr_hip.so for gfx950
(hipcc), and the Fortran host (flang): the ISO_C_BINDING modules of
astr_amd/fortran/astr_hip_mod.F90 and the TGV solver program
astr_amd/fortran/tgvsolver_hip linked against the library."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
SOURCES = ["flow.hip", "chem.hip"]
OUT = os.path.join(PKG, "libastr_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(PKG, "fortran")
FBUILD = os.path.join(FDIR, "_build")
FPROGS = ["tgvsolver_hip", "chemcells_hip"]

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed FP32 instructions stay off (wrong results seen with some forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(PKG, "..", "include"),
]


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for " + " ".join(SOURCES))
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, OUT)
    return OUT


def build_fortran() -> str:
    """Fortran host modules and programs (rpath to the library)."""
    os.makedirs(FBUILD, exist_ok=True)
    mod_o = os.path.join(FBUILD, "astr_hip_mod.o")
    steps = [[FLANG, "-O2", "-cpp", "-fPIC", "-module-dir", FBUILD, "-c",
              os.path.join(FDIR, "astr_hip_mod.F90"), "-o", mod_o]]
    for prog in FPROGS:
        steps.append([FLANG, "-O2", "-cpp", "-module-dir", FBUILD, "-o",
                      os.path.join(FDIR, prog + ".tmp"), os.path.join(FDIR, prog + ".F90"), mod_o,
                      "-L" + PKG, "-lastr_hip", "-Wl,-rpath,$ORIGIN/.."])
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed: " + " ".join(cmd))
    for prog in FPROGS:
        os.replace(os.path.join(FDIR, prog + ".tmp"), os.path.join(FDIR, prog))
    return FDIR


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran())
