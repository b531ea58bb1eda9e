"""Shared test set-up: the `gpu` marker, import paths, and a build of the
oracle (plain C, test infrastructure) when it is missing."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    if not (ROOT / "oracle" / "libpmoracle.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    ref_src = Path("/root/reference/mef/src")
    if ref_src.exists() and not (ROOT / "oracle" / "_ref" / "libpmref.so").exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=False)


@pytest.fixture(scope="session")
def ref():
    from oracle import load_ref
    R = load_ref()
    if R is None:
        pytest.skip("reference-subset library not built (reference sources absent)")
    return R
