"""Pins the oracle's once-per-run structure build.

* hexa8 -> hexa20 promotion (mk_elconn_quad_v1, Adjacency.f:583-693) and the
  equation numbering (numeqpmec2, Numeq.f:86-151): bit-for-bit against the
  reference's own Fortran routines (oracle/_ref);
* counts against the reference's logs: the 1x1x10 column has nnode 128,
  neq 160, nad 2896 (output/poro_mec/elastico/hexa/exemplo1_mesh0_t2_t_1.txt:
  23,26,30) and the 10^3 box nnode 4961 (.../hexa/solo1_t_1.txt:30);
* mk_bound_quad (Rdata.f:2337-2405): the rule "a mid-edge displacement dof is
  fixed iff both edge ends are", checked by brute force;
* graph_v1 (Graph.f:72-133): brute-force union of element cliques;
* genrcm_v1 (Reord.f:45-437, SPARSPAK GENRCM): hand-derived orderings of
  small graphs that exercise the pseudo-peripheral search, the degree sort
  and the component loop, plus the permutation property on real meshes;
* csria_v1/csrja_v1 (Csr.f:631-867): brute-force lower pattern from node
  pairs that share an element.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

lib = O.lib


def _incidence(ix, nnodev):
    numel = ix.shape[0]
    lists = [[] for _ in range(nnodev)]
    for e in range(numel):
        for k in range(8):
            lists[ix[e, k] - 1].append(e + 1)
    maxgrade = max(len(v) for v in lists)
    incid = np.zeros((nnodev, maxgrade), np.int32)  # Fortran incid(maxgrade, nnode)
    nincid = np.zeros(nnodev, np.int32)
    for n, v in enumerate(lists):
        incid[n, :len(v)] = v
        nincid[n] = len(v)
    return incid, nincid, maxgrade


@pytest.mark.parametrize("shape", [(1, 1, 10), (3, 2, 4), (4, 4, 5)])
def test_promotion_bitwise_vs_reference(ref, shape):
    d = P.column_deck(*shape)
    ix_orc = d.ix.copy()
    nnode = lib.orc_promote_hexa20(ix_orc, d.numel, d.nnodev)
    ix_ref = np.ascontiguousarray(d.ix.copy())
    incid, nincid, maxgrade = _incidence(d.ix, d.nnodev)
    nn = C.c_int(0)
    args = [C.c_int(v) for v in (d.numel,)] + [nn] + [C.c_int(v) for v in (d.nnodev, 20, 8, maxgrade)]
    ref.mk_elconn_quad_v1_(ix_ref.ctypes.data_as(C.c_void_p), incid.ctypes.data_as(C.c_void_p),
                           nincid.ctypes.data_as(C.c_void_p), *[C.byref(a) for a in args])
    assert nn.value == nnode
    assert np.array_equal(ix_ref, ix_orc)
    assert nnode == P.box_counts(*shape)["nnode"]


@pytest.mark.parametrize("shape", [(1, 1, 10), (3, 2, 4)])
def test_numbering_bitwise_vs_reference(ref, shape):
    st = O.build_structure(P.column_deck(*shape))
    id_ref = st["idflags"].copy()
    neq = C.c_int(0)
    ints = [C.c_int(v) for v in (st["nnode"], st["nnodev"], 4)]
    ref.numeqpmec2_(id_ref.ctypes.data_as(C.c_void_p), st["perm"].ctypes.data_as(C.c_void_p),
                    id_ref.ctypes.data_as(C.c_void_p), st["fnno"].ctypes.data_as(C.c_void_p),
                    *[C.byref(a) for a in ints], C.byref(neq))
    assert neq.value == st["neq"]
    assert np.array_equal(id_ref, st["id"])


def test_counts_match_reference_logs():
    st = O.build_structure(P.column_deck(1, 1, 10))
    assert (st["nnode"], st["neq"], st["nad"]) == (128, 160, 2896)
    st = O.build_structure(P.column_deck(10, 10, 10))
    assert st["nnode"] == 4961


def test_bound_quad_rule():
    d = P.column_deck(3, 2, 4)
    st = O.build_structure(d)
    ix, fl = st["ix"], st["idflags"]
    edges = [(1, 2, 9), (2, 3, 10), (3, 4, 11), (4, 1, 12), (5, 6, 13), (6, 7, 14),
             (7, 8, 15), (8, 5, 16), (1, 5, 17), (2, 6, 18), (3, 7, 19), (4, 8, 20)]
    for e in range(d.numel):
        for a, b, m in edges:
            na, nb, nm = ix[e, a - 1] - 1, ix[e, b - 1] - 1, ix[e, m - 1] - 1
            for k in range(3):
                assert fl[nm, k] == int(fl[na, k] == 1 and fl[nb, k] == 1)
            assert fl[nm, 3] == 0


def _brute_graph(ix, nnode):
    nb = [set() for _ in range(nnode)]
    for row in ix[:, :20]:
        for a in row:
            nb[a - 1].update(int(v) for v in row if v != a)
    return nb


def test_graph_brute_force():
    st = O.build_structure(P.column_deck(2, 3, 2))
    nb = _brute_graph(st["ix"], st["nnode"])
    ip, ips = st["ip"], st["ips"]
    for n in range(st["nnode"]):
        got = ips[ip[n] - 1: ip[n + 1] - 1]
        assert list(got) == sorted(nb[n])


def _rcm(adj_lists):
    n = len(adj_lists)
    ip = np.ones(n + 1, np.int64)
    for i, a in enumerate(adj_lists):
        ip[i + 1] = ip[i] + len(a)
    ips = np.array([v for a in adj_lists for v in sorted(a)] or [0], np.int32)
    perm = np.zeros(n, np.int32)
    lib.orc_genrcm(n, ip, ips, perm)
    return list(perm)


def test_rcm_hand_derived():
    # path 1-2-3-4-5: root 1 is peripheral, CM order 1..5 reversed
    assert _rcm([[2], [1, 3], [2, 4], [3, 5], [4]]) == [5, 4, 3, 2, 1]
    # path 3-1-2-4-5: the search moves the root to 5; CM 5,4,2,1,3 reversed
    assert _rcm([[2, 3], [1, 4], [1], [2, 5], [4]]) == [3, 1, 2, 4, 5]
    # tree 1-{2,3,4}, 4-{5,6}: roots 1 -> 5 -> 2 (kept when the depth does
    # not grow); neighbours of 1 sorted by degree (3 before 4); reversed
    assert _rcm([[2, 3, 4], [1], [1], [1, 5, 6], [4], [4]]) == [6, 5, 4, 3, 1, 2]
    # two components {1,3} and {2}
    assert _rcm([[3], [], [1]]) == [3, 1, 2]


def test_rcm_is_permutation_on_meshes():
    for shape in [(2, 2, 3), (4, 3, 2)]:
        st = O.build_structure(P.column_deck(*shape))
        assert sorted(st["perm"]) == list(range(1, st["nnode"] + 1))


@pytest.mark.parametrize("reorder", [True, False])
def test_csrc_brute_force(reorder):
    st = O.build_structure(P.column_deck(2, 2, 3), reorder=reorder)
    nb = _brute_graph(st["ix"], st["nnode"])
    idm = st["id"]
    rows = [set() for _ in range(st["neq"])]
    for n in range(st["nnode"]):
        for m in list(nb[n]) + [n + 1]:
            for a in idm[n]:
                for b in idm[m - 1]:
                    if a > 0 and b > 0 and b < a:
                        rows[a - 1].add(int(b))
    ia, ja = st["ia"], st["ja"]
    assert ia[0] == 1 and ia[-1] - 1 == st["nad"]
    for r in range(st["neq"]):
        assert list(ja[ia[r] - 1: ia[r + 1] - 1]) == sorted(rows[r])
