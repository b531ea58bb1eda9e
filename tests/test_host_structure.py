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


@pytest.mark.parametrize("shape,reorder", [((1, 1, 10), True), ((3, 2, 4), True),
                                           ((2, 3, 2), False), ((6, 5, 4), True)])
def test_q1_numbering_matches_oracle(shape, reorder):
    """maxno 8 (the Q1/Q1 element): no promotion, every node a vertex with
    4 dofs, the same graph, RCM and numbering rules."""
    d = P.column_deck(*shape, nen=8)
    st = O.build_structure(d, reorder=reorder)
    got = mesh_numbering(d, reorder=reorder)
    c = P.box_counts(*shape, nen=8)
    assert got["nnode"] == st["nnode"] == c["nnode"] == d.nnodev
    assert got["neq"] == st["neq"] == c["dofs"] - int(d.idv.sum())
    assert got["nad"] == st["nad"]
    assert np.array_equal(got["ix"][:, 8:20], np.zeros((d.numel, 12), np.int32))
    assert np.array_equal(got["ix"], st["ix"])
    assert np.array_equal(got["perm"], st["perm"])
    assert np.array_equal(got["id"], st["id"])


def test_q1_pattern_is_the_27_point_stencil():
    """Without restraints a Q1/Q1 box has 4 x (number of vertex neighbours,
    itself included) columns per row: 27 in the interior, the nominal
    ~108 nonzeros per row of the benchmark configurations."""
    d = P.column_deck(4, 4, 4, nen=8)
    d.idv[:] = 0
    got = mesh_numbering(d)
    nv = (d.nx + 1) * (d.ny + 1) * (d.nz + 1)
    cnt = [(min(i, 1) + min(d.nx - i, 1) + 1) for i in range(d.nx + 1)]
    per_vertex = np.einsum("i,j,k->ijk", cnt, cnt, cnt).sum()
    full = 16 * per_vertex  # 4 rows x 4 columns per vertex pair
    assert got["neq"] == 4 * nv
    assert 2 * got["nad"] + got["neq"] == full


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
