"""The product's host structure build (libpmhip pm_mesh_numbering: promotion,
mid-edge restraints, graph, RCM, numbering, lower-pattern size) against the
oracle, bit for bit, and the C ABI's exported symbol set against
include/poromec_hip.h.  No device is needed."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P
from poromec_amd.solver import mesh_numbering

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("shape,reorder", [((1, 1, 10), True), ((3, 2, 4), True),
                                           ((4, 4, 6), True), ((2, 3, 2), False),
                                           ((6, 5, 4), True)])
def test_numbering_matches_oracle(shape, reorder):
    d = P.column_deck(*shape)
    st = O.build_structure(d, reorder=reorder)
    got = mesh_numbering(d, reorder=reorder)
    assert got["nnode"] == st["nnode"]
    assert got["neq"] == st["neq"]
    assert got["nad"] == st["nad"]
    assert np.array_equal(got["ix"], st["ix"])
    assert np.array_equal(got["perm"], st["perm"])
    assert np.array_equal(got["id"], st["id"])


def test_heterogeneous_deck_numbering():
    d = P.heterogeneous_box_deck(4, seed=3)
    st = O.build_structure(d)
    got = mesh_numbering(d)
    assert np.array_equal(got["id"], st["id"]) and got["nad"] == st["nad"]


def test_abi_exports_every_declared_symbol():
    header = (ROOT / "include" / "poromec_hip.h").read_text()
    declared = set(re.findall(r"^\s*(?:const\s+)?(?:int|void|char\s*\*|const char\s*\*)\s*\**\s*(pm_\w+)\s*\(",
                              header, flags=re.M))
    assert "pm_step" in declared and "pm_solve" in declared
    lib = C.CDLL(str(ROOT / "poromec_amd" / "libpmhip.so"))
    for name in sorted(declared):
        assert hasattr(lib, name), name
