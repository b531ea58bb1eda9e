"""The Fortran bind(C) derived types of poromec_amd/fortran/pmhip.f90 must
have the layout of the C structs of include/poromec_hip.h: a tiny Fortran
program prints their storage sizes (no device needed)."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

from poromec_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
FC = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not Path(FC).exists(), reason="amdflang absent")
def test_bind_c_struct_sizes(tmp_path):
    src = tmp_path / "sz.f90"
    src.write_text("""program sz
  use iso_c_binding
  use pmhip
  type(pm_solver_cfg) :: a
  type(pm_solve_info) :: b
  type(pm_step_info) :: c
  print *, c_sizeof(a), c_sizeof(b), c_sizeof(c)
end program
""")
    mod = tmp_path / "mod"
    mod.mkdir()
    exe = tmp_path / "sz"
    r = subprocess.run([FC, "-J", str(mod), str(ROOT / "poromec_amd/fortran/pmhip.f90"), str(src),
                        "-L", str(ROOT / "poromec_amd"), "-lpmhip",
                        f"-Wl,-rpath,{ROOT / 'poromec_amd'}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out] == [C.sizeof(L.SolverCfg), C.sizeof(L.SolveInfo),
                                     C.sizeof(L.StepInfo)]
