"""bench.py's tables of the SpMV shapes agree with the library's (no device
needed): one entry per shape the tuner times, a kernel name for each, and
byte counts that grow with the value width and the fused dot."""
import importlib.util
from pathlib import Path

from poromec_amd.solver import SPMV_SHAPES

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_shape_tables_cover_every_shape():
    b = _bench()
    assert len(b.shape_lanes) == len(b.shape_unr) == len(b.shape_pf) == SPMV_SHAPES
    header = (ROOT / "include" / "poromec_hip.h").read_text()
    assert f"#define PM_SPMV_SHAPES {SPMV_SHAPES}" in header
    spmv_h = (ROOT / "poromec_amd" / "csrc" / "pm_spmv.h").read_text()
    assert f"#define PM_SPMV_NCFG {SPMV_SHAPES}" in spmv_h


def test_kernel_names_and_bytes():
    b = _bench()
    info = {"nnz": 1000, "ncols": 300, "nruns": 90, "nblocks": 10, "neq": 40}
    for shape in range(SPMV_SHAPES):
        name = b.spmv_kernel_name("fp32", shape)
        assert name.startswith("k_spmv") and "float" in name
        b32 = b.spmv_bytes(info, 4, shape=shape)
        b64 = b.spmv_bytes(info, 8, shape=shape)
        assert b64 - b32 == 4 * info["nnz"]
        assert b.spmv_bytes(info, 4, fused_dot=True, shape=shape) - b32 == 8 * info["neq"]
    assert b.spmv_alg_bytes(info, 8) == 12 * 1000 + 8 * 41 + 16 * 40
