"""Build libfall3d_ads_hip.so (gfx950) and the Fortran MODULE ADS shim in-tree.

`python -m fall3d_amd.build` compiles every HIP source under fall3d_amd/csrc
with hipcc for gfx950 (objects in fall3d_amd/_build, library in
fall3d_amd/lib), then the Fortran shim and its drop-in driver with amdflang.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfall3d_ads_hip.so")
ARCH = os.environ.get("FALL3D_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

# -ffp-contract=off: the kernels place every fused multiply-add explicitly
# (fmaf) so that both owners of a shared face compute bit-identical fluxes.
# The packed-FP32 target feature is disabled (no v_pk_*_f32 generation).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(BUILD, f[:-4] + ".o") for f in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4)

    def comp(src, obj):
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(comp, srcs, objs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> list[str]:
    """Compile the MODULE ADS shim (mod_ADS_hip.F90) against the test-harness
    KindType/Domain modules and link the drop-in driver program; compile the
    resident-group interfaces (fall3d_ads_c.F90) and link the resident
    driver program."""
    if not os.path.isdir(FSRC):
        return []
    out = os.path.join(BUILD, "fortran")
    os.makedirs(out, exist_ok=True)
    programs = {"fall3d_dropin_driver": ["harness_kinds.F90", "mod_ADS_hip.F90", "dropin_driver.F90"],
                "fall3d_resident_driver": ["fall3d_ads_c.F90", "resident_driver.F90"]}
    exes = []
    for name, order in programs.items():
        objs = []
        for f in order:
            src = os.path.join(FSRC, f)
            if not os.path.exists(src):
                return exes
            obj = os.path.join(out, f.replace(".F90", ".o"))
            cmd = [FC, "-O2", "-cpp", "-fPIC", "-module-dir", out, "-I", out, "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            _run(cmd)
            objs.append(obj)
        exe = os.path.join(LIBDIR, name)
        _run([FC, "-o", exe, *objs, "-L", LIBDIR, "-lfall3d_ads_hip", f"-Wl,-rpath,$ORIGIN"])
        exes.append(exe)
    return exes


def build(verbose: bool = False) -> None:
    build_hip(verbose=verbose)
    build_fortran(verbose=verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
