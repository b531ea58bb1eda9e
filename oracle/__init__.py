"""Oracle for the mef Biot u-p path: a plain-C restatement of the reference
algorithms (``pm_*.c``) and a ctypes wrapper around it.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, as the checker
or as the timed CPU baseline.  The product (``poromec_amd``) never calls it.

Parity status of each oracle function (how it is pinned) is recorded in the
docstrings of ``tests/test_oracle_*.py``.
"""
from .oracle import (  # noqa: F401
    OracleEnv,
    OracleModel,
    build_structure,
    elmt17,
    lib,
    load_ref,
    ref_elmt17,
)
