"""Builds libpmhip.so (HIP for gfx950) in place, next to this file.

    python -m poromec_amd.build          # or: python poromec_amd/build.py

Every source under csrc/ is compiled by hipcc into an object under
csrc/_build/ (in parallel), then linked into poromec_amd/libpmhip.so.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libpmhip.so"
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 instructions in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # every product and sum rounded on its own, as in the reference's Fortran
    # (no fused multiply-add contraction)
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def build(verbose: bool = False) -> Path:
    obj_dir = CSRC / "_build"
    obj_dir.mkdir(exist_ok=True)

    def compile_one(src: Path) -> Path:
        obj = obj_dir / (src.stem + ".o")
        cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, sources()))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")


def build_fortran() -> Path | None:
    """The Fortran host (iso_c_binding module, deck reader and the
    pm_mefdrive driver), linked against libpmhip.so; skipped when amdflang
    is absent."""
    fdir = HERE / "fortran"
    exe = fdir / "pm_mefdrive"
    if not Path(FC).exists():
        return None
    mod_dir = fdir / "_build"
    mod_dir.mkdir(exist_ok=True)
    cmd = [FC, "-O2", "-J", str(mod_dir), str(fdir / "pmhip.f90"), str(fdir / "pm_deck.f90"), str(fdir / "pm_vtk.f90"),
           str(fdir / "pm_mefdrive.f90"),
           "-L", str(HERE), "-lpmhip", "-Wl,-rpath,$ORIGIN/..", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stderr}")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran())
