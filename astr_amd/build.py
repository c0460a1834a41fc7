"""Build, in-tree: the HIP library astr_amd/libastr_hip.so for gfx950
(hipcc), and the Fortran host (flang): the ISO_C_BINDING modules of
astr_amd/fortran/astr_hip_mod.F90, the miniapp's TGV solver program
astr_amd/fortran/tgvsolver_hip, the chemistry cell program and the main
solver's run program astr_amd/fortran/astr_hip (astr_hip_run.F90), linked
against the library."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
SOURCES = ["flow.hip", "chem.hip", "upwind.hip", "probe.hip"]
# passes.hip is compiled once per launch shape (-DASTR_SHAPE_ID=k)
PASS_SHAPES = 13
OUT = os.path.join(PKG, "libastr_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(PKG, "fortran")
FBUILD = os.path.join(FDIR, "_build")
FPROGS = ["tgvsolver_hip", "chemcells_hip"]
# astr_hip_run.F90 holds the main solver's module and its program astr_hip
FRUN = "astr_hip_run"

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # packed FP32 instructions stay off (wrong results seen with some forms)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(PKG, "..", "include"),
]


def _compile(job):
    src, obj, defs = job
    cmd = [HIPCC, *[f for f in FLAGS if f != "-shared"], *defs, "-c", "-o", obj, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, r


def build(verbose: bool = False) -> str:
    """Compile the translation units in parallel, then link the library."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(CSRC, "_build")
    os.makedirs(objdir, exist_ok=True)
    jobs = [(os.path.join(CSRC, s), os.path.join(objdir, s.replace(".hip", ".o")), [])
            for s in SOURCES]
    jobs += [(os.path.join(CSRC, "passes.hip"), os.path.join(objdir, f"passes_s{k}.o"),
              [f"-DASTR_SHAPE_ID={k}"]) for k in range(PASS_SHAPES)]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(_compile, jobs))
    for cmd, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        if verbose:
            sys.stderr.write(r.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp, *[j[1] for j in jobs]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed: " + " ".join(cmd))
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
    steps.append([FLANG, "-O2", "-cpp", "-module-dir", FBUILD, "-o",
                  os.path.join(FDIR, "astr_hip.tmp"), os.path.join(FDIR, FRUN + ".F90"), mod_o,
                  "-L" + PKG, "-lastr_hip", "-Wl,-rpath,$ORIGIN/.."])
    for cmd in steps:
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed: " + " ".join(cmd))
    for prog in FPROGS + ["astr_hip"]:
        os.replace(os.path.join(FDIR, prog + ".tmp"), os.path.join(FDIR, prog))
    return FDIR


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran())
