"""The pres macro's results writer (poromec_amd.vtk, and pm_vtk.f90 in the
Fortran host):

* its number formats and section lines reproduce the one results file the
  reference ships (tests/golden/vtk, legacy ASCII), line for line where the
  current writer writes them as that file's writer did;
* list-directed output (VTU time fields, the legacy cell-field headers) is
  amdflang's, value by value;
* Python and the Fortran host write the same bytes for the same fields, in
  the three encodings, over the vertices and over every node;
* with ``quadratic``, edge nodes sit at their edge's midpoint and carry the
  average of its vertices, and each VTK quadratic-hexahedron edge node is the
  midpoint of the two corners VTK's cell definition pairs it with."""
import lzma
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deckcases import extract
from poromec_amd import problem as P
from poromec_amd import vtk as V
from poromec_amd.deck import read_deck
from poromec_amd.ftext import fortran_e, fortran_f, fortran_i, list_record
from poromec_amd.solver import mesh_numbering

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "poromec_amd" / "fortran" / "pm_mefdrive"
FC = Path("/opt/rocm/bin/amdflang")
GOLD = Path(__file__).resolve().parent / "golden" / "vtk" / "solo_cilindro_tetra_1_istep_0.vtk.xz"


def _e155(v):
    return fortran_e(v, 15, 5)


def _e155e3(v):
    return fortran_e(v, 15, 5, 3)


def _i10(v):
    return fortran_i(v, 10)


def test_reference_results_file_formats():
    """Every number line of the shipped file, and the section lines the
    current writer shares with it, rebuilt from the values they hold.  Its
    writer is older than the current source in three places, not checked:
    FieldData has 3 arrays (TIME(s), TIME(h), CYCLE), the cell SCALARS
    headers are another compiler's list-directed output of other items, and
    the field names differ (flux, stress, ...)."""
    text = lzma.open(GOLD).read().decode()
    assert "\r\n" in text  # shipped with CRLF ends; the writer's own end is LF
    L = text.replace("\r\n", "\n").split("\n")
    assert L[-1] == ""
    L = L[:-1]
    i = L.index("FIELD FieldData 3") + 1
    for kind in ("double", "double", "int"):
        assert L[i].endswith(f" 1 1 {kind}")
        v = L[i + 1]
        assert v == (fortran_f(float(v), 16, 6) if kind == "double" else fortran_i(int(v), 9))
        i += 2
    n = int(L[i].split()[1])
    assert L[i] == "POINTS " + _i10(n) + " double"
    for ln in L[i + 1:i + 1 + n]:
        assert ln == "".join(_e155(float(v)) for v in ln.split()), ln
    i += 1 + n
    numel, tot = (int(v) for v in L[i].split()[1:])
    assert L[i] == "CELLS " + _i10(numel) + _i10(tot)
    for ln in L[i + 1:i + 1 + numel]:
        assert ln == "".join(_i10(int(v)) for v in ln.split()), ln
    i += 1 + numel
    assert L[i] == "CELL_TYPES " + _i10(numel)
    assert all(ln == fortran_i(int(ln), 3) for ln in L[i + 1:i + 1 + numel])
    i += 1 + numel
    assert L[i] == "CELL_DATA " + fortran_i(numel, 9)
    i += 1
    for _ in range(2):  # mat, elIdG
        assert L[i].split()[0] == "SCALARS" and L[i + 1] == "LOOKUP_TABLE default"
        assert all(ln == _i10(int(ln)) for ln in L[i + 2:i + 2 + numel])
        i += 2 + numel
    assert L[i] == "POINT_DATA " + fortran_i(n, 9)
    i += 1
    seen = []
    while i < len(L):
        w = L[i].split()
        seen.append(w[1])
        if w[0] == "SCALARS":
            assert L[i] == V._scalar_head(w[1].ljust(15), "int    " if w[2] == "int"
                                          else "double "), L[i]
            assert L[i + 1] == "LOOKUP_TABLE default"
            fmt = _i10 if w[2] == "int" else _e155e3
            conv = int if w[2] == "int" else float
            assert all(ln == fmt(conv(ln)) for ln in L[i + 2:i + 2 + n]), w[1]
            i += 2 + n
        else:
            assert w[0] in ("VECTORS", "TENSORS") and w[2] == "double"
            assert L[i] == w[0] + " " + w[1].ljust(15) + " double"
            rows = L[i + 1:i + 1 + n]
            assert all(ln == "".join(_e155e3(float(v)) for v in ln.split()) for ln in rows), w[1]
            assert all(len(ln.split()) == (3 if w[0] == "VECTORS" else 9) for ln in rows)
            i += 1 + n
    assert seen == ["noIdG", "desloc", "pressao", "deltaPressao", "flux", "stress", "stressE",
                    "stressB"]


@pytest.mark.skipif(not FC.exists(), reason="amdflang absent")
def test_list_directed_matches_amdflang(tmp_path):
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.uniform(-1, 1, 300) * 10.0 ** rng.integers(-20, 20, 300),
        np.array([0.0, -0.0, 0.1, 0.095, 0.0949999, 0.0950001, 9.5e14, 9.4999e14, 1e15,
                  0.5, 1 / 3, 2.0 / 3600.0, 12345.678, 12345.678 / 3600.0,
                  12345.678 / 86400.0, 2.2250738585072014e-308, 1.7976931348623157e308,
                  2.0 ** 52, 2.0 ** 53 + 2, 90407898141504448.0, 0.1 + 0.2]),
    ])
    np.savetxt(tmp_path / "v.txt", vals.reshape(-1, 1), fmt="%.17e")
    (tmp_path / "w.f90").write_text(
        "program w\n  real(8) :: v\n  integer :: i, ios\n"
        "  open(10, file='v.txt')\n  do\n    read(10, *, iostat=ios) v\n"
        "    if (ios /= 0) exit\n    write(*, *) v\n  end do\n"
        "  do i = -3, 3\n    write(*, *) i * 1000003\n  end do\n"
        "  write(*, *) 'SCALARS ', 'elIdG          ', ' int ', 1\nend program\n")
    subprocess.run([str(FC), "w.f90", "-o", "w"], cwd=tmp_path, check=True, capture_output=True)
    got = subprocess.run(["./w"], cwd=tmp_path, check=True, capture_output=True,
                         text=True).stdout.splitlines()
    want = [list_record(float(v)) for v in vals] + [list_record(i * 1000003) for i in range(-3, 4)]
    want.append(list_record("SCALARS ", "elIdG".ljust(15), " int ", 1))
    assert len(got) == len(want)
    # identical text, except that where 17 digits are needed the processor
    # may take another 17-digit string that reads back to the same value
    bad = []
    for g, w in zip(got, want):
        if g == w:
            continue
        gd = g.strip().lstrip("-").split("E")[0].replace(".", "").strip("0")
        same = (len(gd) == 17 and float(g) == float(w) and len(g) == len(w)
                and ("E" in g) == ("E" in w))
        if not same:
            bad.append((g, w))
    assert not bad, bad[:5]
    assert sum(g == w for g, w in zip(got, want)) >= 0.9 * len(got)


def _synthetic(nnode):
    """The vtkcheck fields: f(k, i) = (mod(37 i + 101 k, 2001) - 1000) 2^(2k-20)."""
    i = np.arange(1, nnode + 1)[:, None]
    k = np.arange(1, 28)[None, :]
    return ((i * 37 + k * 101) % 2001 - 1000).astype(float) * 2.0 ** (2 * k - 20)


@pytest.mark.skipif(not EXE.exists(), reason="Fortran host not built (amdflang absent)")
@pytest.mark.parametrize("case", ["solo1", "solo_cilindro_hexa_1"])
def test_python_and_fortran_writers_agree(tmp_path, case):
    path = extract(case, tmp_path)
    r = subprocess.run([str(EXE), str(path), str(tmp_path / "f"), "vtkcheck"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    inp = read_deck(path)
    num = mesh_numbering(inp.deck, inp.reorder)
    ix20 = num["ix"]
    f = _synthetic(num["nnode"])
    for quad, q in ((False, "l"), (True, "q")):
        fl = (quad,) + (True,) * 8 + (False,) * 4
        for enc, legacy, binary in (("a", True, False), ("b", True, True), ("x", False, False)):
            name = V.results_name(str(tmp_path / f"p_{enc}{q}"), 7, legacy)
            V.write_results(name, ix20, inp.deck.x, f[:, 0:4], f[:, 4], f[:, 5], f[:, 6:12],
                            f[:, 12:18], f[:, 18:24], f[:, 24:27], 7, 12345.678, flags=fl,
                            quad=quad, legacy=legacy, binary=binary, nnodev=inp.deck.nnodev)
            fo = V.results_name(str(tmp_path / f"f_{enc}{q}"), 7, legacy)
            a, b = Path(name).read_bytes(), Path(fo).read_bytes()
            if not legacy:  # the list-directed time values: the same double
                la, lb = a.split(b"\n"), b.split(b"\n")
                assert len(la) == len(lb)
                for j in range(len(la)):
                    if b'Name="TIME_' in la[j]:
                        assert float(la[j + 1]) == float(lb[j + 1])
                        la[j + 1] = lb[j + 1]
                a = b"\n".join(la)
            if a != b:
                k = next(j for j in range(min(len(a), len(b))) if a[j] != b[j])
                raise AssertionError((enc, q, len(a), len(b), a[max(0, k - 80):k + 80],
                                      b[max(0, k - 80):k + 80]))


def test_quadratic_nodes_and_vtk_order(tmp_path):
    """On a box mesh: edge nodes at their edges' midpoints, a field linear in
    x reproduced there exactly, and the written hexa20 cells in VTK's order
    (corners 0-3 one face and 4-7 the opposite one, edge nodes 8-11, 12-15,
    16-19 between corners (0,1),(1,2),(2,3),(3,0), (4,5),...,(7,4) and
    (0,4),...,(3,7))."""
    deck = P.column_deck(2, 3, 2)
    num = mesh_numbering(deck, True)
    ix20, nnode, nv = num["ix"], num["nnode"], deck.nnodev
    xf = np.zeros((nnode, 3))
    xf[:nv] = deck.x
    xq = V.quad_field(xf, ix20, nnode, True)
    for a, b, c in V.HEX20_EDGES:
        assert np.array_equal(xq[ix20[:, c] - 1], 0.5 * (xq[ix20[:, a] - 1] + xq[ix20[:, b] - 1]))
    lin = np.zeros(nnode)
    lin[:nv] = 3.0 + 2.0 * deck.x[:, 0] - 4.0 * deck.x[:, 1] + 0.5 * deck.x[:, 2]
    want = 3.0 + 2.0 * xq[:, 0] - 4.0 * xq[:, 1] + 0.5 * xq[:, 2]
    assert np.allclose(V.quad_field(lin, ix20, nnode, True), want, rtol=0, atol=1e-12)
    z = np.zeros((nnode, 6))
    u = np.zeros((nnode, 4))
    p = V.write_results(tmp_path / "q.vtk", ix20, deck.x, u, np.zeros(nnode), np.zeros(nnode), z,
                        z, z, np.zeros((nnode, 3)), 1, 0.5, flags=(True, True, True) + (False,) * 10,
                        quad=True, legacy=True, nnodev=nv)
    L = Path(p).read_text().split("\n")
    i = next(k for k, ln in enumerate(L) if ln.startswith("POINTS"))
    pts = np.array([[float(v) for v in ln.split()] for ln in L[i + 1:i + 1 + nnode]])
    j = i + 1 + nnode
    numel = int(L[j].split()[1])
    cells = np.array([[int(v) for v in ln.split()] for ln in L[j + 1:j + 1 + numel]])
    assert (cells[:, 0] == 20).all()
    c = cells[:, 1:]
    vtk_edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4),
                 (0, 4), (1, 5), (2, 6), (3, 7)]
    for m, (a, b) in enumerate(vtk_edges):
        mid = 0.5 * (pts[c[:, a]] + pts[c[:, b]])
        assert np.allclose(pts[c[:, 8 + m]], mid, rtol=1e-4, atol=1e-5), m
    # the two faces are parallel and opposite: corners 4-7 are 0-3 moved by one edge
    d = pts[c[:, 4:8]] - pts[c[:, 0:4]]
    assert np.allclose(d, d[:, :1], rtol=1e-4, atol=1e-5)
    assert L[j + 1 + numel] == "CELL_TYPES " + _i10(numel)
    assert L[j + 2 + numel] == " 25"


def test_binary_vtu_rejected(tmp_path):
    deck = P.column_deck(1, 1, 1)
    num = mesh_numbering(deck, True)
    n = num["nnode"]
    with pytest.raises(V.VtkError):
        V.write_results(tmp_path / "x.vtu", num["ix"], deck.x, np.zeros((n, 4)), np.zeros(n),
                        np.zeros(n), None, None, None, None, 0, 0.0, legacy=False, binary=True)
