"""The Fortran host's deck reader (poromec_amd/fortran/pm_deck.f90, run as
``pm_mefdrive <deck> <prefix> parse``, no device calls) reads every deck the
same as poromec_amd.deck: the reference's example decks, decks in the
current grammar written by write_deck, and the earlier forms; decks outside
the path stop it with an error."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deckcases import CASES, extract
from poromec_amd import problem as P
from poromec_amd.deck import read_deck, write_deck
from poromec_amd.solver import SOLVERS, SolverConfig

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "poromec_amd" / "fortran" / "pm_mefdrive"
OPS = {1: "solv", 2: "pnup", 3: "pnsf", 4: "dt", 5: "pres"}

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="Fortran host not built (amdflang absent)")


def fparse(path, tmp):
    """What the Fortran reader made of a deck (its <prefix>_deck.bin)."""
    prefix = tmp / "fparse"
    r = subprocess.run([str(EXE), str(path), str(prefix), "parse"], capture_output=True, text=True,
                       timeout=120)
    if r.returncode != 0:
        return None, r.stdout + r.stderr
    b = (tmp / "fparse_deck.bin").read_bytes()
    pos = 0

    def take(dt, n):
        nonlocal pos
        a = np.frombuffer(b, dt, n, pos)
        pos += a.nbytes
        return a

    h = take(np.int32, 17)
    numel, nnodev, numat, reorder = (int(v) for v in h[:4])
    out = dict(numel=numel, nnodev=nnodev, numat=numat, reorder=bool(reorder),
               vprop=tuple(bool(v) for v in h[4:8]), solver=int(h[8]), precond=int(h[9]),
               maxit=int(h[10]), ngram=int(h[11]), maxnlit=int(h[12]), stop_crit=int(h[13]),
               maxnlit_accept=bool(h[14]))
    npn, nprog = int(h[15]), int(h[16])
    out["tol"], out["nltol"] = take(np.float64, 2)
    out["gravity"] = take(np.float64, 3)
    out["gmod"] = take(np.float64, 1)[0]
    out["pnodes"] = take(np.int32, npn).tolist()
    ops = take(np.int32, nprog)
    vals = take(np.float64, nprog)
    out["program"] = [("dt", float(v)) if o == 4 else (OPS[int(o)],) for o, v in zip(ops, vals)]
    out["ix"] = take(np.int32, 21 * numel).reshape(numel, 21)
    out["idv"] = take(np.int32, 4 * nnodev).reshape(nnodev, 4)
    out["ie"] = take(np.int32, numat)
    out["eload"] = take(np.int32, 7 * numel).reshape(numel, 7)
    out["eloadp"] = take(np.int32, 7 * numel).reshape(numel, 7)
    out["load"] = take(np.int32, 600).reshape(200, 3)
    out["x"] = take(np.float64, 3 * nnodev).reshape(nnodev, 3)
    out["e"] = take(np.float64, 15 * numat).reshape(numat, 15)
    out["u0"] = take(np.float64, 4 * nnodev).reshape(nnodev, 4)
    out["fload"] = take(np.float64, 1000 * 5 * 200).reshape(200, 5, 1000)
    fl = take(np.int32, 16)
    out["print_flags"] = tuple(bool(v) for v in fl[:13])
    out["vtk_legacy"], out["vtk_binary"] = bool(fl[13]), bool(fl[14])
    out["nen"] = int(fl[15])
    assert pos == len(b)
    return out, r.stdout


def assert_same(f, inp):
    d, c = inp.deck, inp.cfg
    assert (f["numel"], f["nnodev"], f["numat"]) == (d.numel, d.nnodev, d.e.shape[0])
    assert f["nen"] == d.nen
    assert f["reorder"] == inp.reorder
    assert f["vprop"] == tuple(bool(v) for v in inp.vprop)
    assert f["solver"] == SOLVERS[c.solver]
    assert (f["precond"], f["maxit"], f["ngram"]) == (c.precond, c.maxit, c.ngram)
    assert (f["maxnlit"], f["stop_crit"], f["maxnlit_accept"]) == \
        (c.maxnlit, c.stop_crit, bool(c.maxnlit_accept))
    assert f["tol"] == c.tol and f["nltol"] == c.nltol
    assert np.array_equal(f["gravity"], d.gravity) and f["gmod"] == d.gravity_mod
    assert f["pnodes"] == list(inp.pnodes)
    assert f["program"] == list(inp.program)
    assert np.array_equal(f["ix"], d.ix)
    assert np.array_equal(f["idv"], d.idv)
    assert np.array_equal(f["ie"], inp.ie)
    assert np.array_equal(f["eload"], d.eload) and np.array_equal(f["eloadp"], d.eloadp)
    assert np.array_equal(f["load"], d.load)
    # the numbers are read by two parsers (Fortran list-directed, Python
    # float): correctly rounded both, so equal to the bit
    assert np.array_equal(f["x"], d.x)
    assert np.array_equal(f["e"], d.e)
    assert np.array_equal(f["u0"], inp.u0)
    assert np.array_equal(f["fload"], d.fload)
    assert f["print_flags"] == tuple(inp.print_flags)
    assert (f["vtk_legacy"], f["vtk_binary"]) == (inp.vtk_legacy, inp.vtk_binary)


@pytest.mark.parametrize("case", CASES)
def test_example_decks(tmp_path, case):
    path = extract(case, tmp_path)
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    assert_same(f, read_deck(path))


def _written(tmp_path, nen=20, **kw):
    deck = P.column_deck(2, 2, 3, dt=0.05, nen=nen)
    deck.load[4] = [39, 2, 3]  # a time-table load from a file
    deck.fload[4, 0, :3] = [0.0, 1.0, 2.0]
    deck.fload[4, 1, :3] = [-1.5e3, 2.25e3, 3.0e3]
    deck.eload[0, 2] = 5
    u0 = np.zeros((deck.nnodev, 4))
    u0[:, 3] = np.linspace(1.0e5, 2.0e5, deck.nnodev)
    cfg = SolverConfig(solver="minres", precond=5, tol=1.0e-9, maxit=777, ngram=40, maxnlit=7,
                       nltol=1.0e-5, stop_crit=1)
    return write_deck(tmp_path / "w.dat", deck, cfg, u0=u0, reorder=False, pnodes=(1, 5, 9),
                      nsteps=4, **kw)


def test_written_deck(tmp_path):
    path = _written(tmp_path)
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    inp = read_deck(path)
    assert inp.reorder is False and inp.cfg.solver == "minres"
    assert_same(f, inp)


def test_written_deck_q1(tmp_path):
    """A maxno-8 deck (the Q1/Q1 element, type 7) reads the same in both
    readers; a maxno-8 deck with an element-17 material stops the Fortran
    reader."""
    path = _written(tmp_path, nen=8)
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    inp = read_deck(path)
    assert inp.deck.nen == 8 and f["nen"] == 8 and f["ie"].tolist() == [7]
    assert_same(f, inp)
    path = Path(path)
    text = path.read_text().replace("\n1 7 ", "\n1 17 ")
    path.write_text(text)
    f, log = fparse(path, tmp_path)
    assert f is None and "maxno 8" in log


def test_written_deck_vprop(tmp_path):
    path = _written(tmp_path, element_type=18, vprop=(True, True, False, True))
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    inp = read_deck(path)
    assert inp.vprop == (True, True, False, True)
    assert_same(f, inp)


def _mesh_lines(n=1):
    deck = P.column_deck(1, 1, n)
    L = [f"nnode {deck.nnodev} numel {deck.numel} numat 1 maxno 20 ndf 4 dim 3",
         "materials", "1 7 " + " ".join(repr(float(v)) for v in deck.e[0][:8]), "end materials",
         "coordinates"]
    L += [f"{k + 1} " + " ".join(repr(float(v)) for v in x) for k, x in enumerate(deck.x)]
    L += ["end coordinates", "hexa8"]
    L += [f"{k + 1} " + " ".join(str(v) for v in r[:8]) + " 1" for k, r in enumerate(deck.ix)]
    L += ["end hexa8", "constrainpmec", "1 1 1 1 0", "2 1 1 1 1", "end constrainpmec"]
    return L


def test_earlier_forms(tmp_path):
    """Inline config and gravity, pcg maxit tol with tol 0, maxnlit, inline
    type 39 load, element type 7, an unknown block closed by end <name>,
    insert/return, a loop without next and a gravity after the mesh."""
    (tmp_path / "inc.dat").write_text("initialpres\n3 2.5d4\n4 -1.25D+3\nend initialpres\nreturn\n"
                                      "garbage after return\n")
    (tmp_path / "node.dat").write_text("2\n3\n4\n")
    L = ["config", "reord false", "omp_elmt true", "end config", "gravity 0.0 -1.0d0 -9.8",
         "pcg 1234 0.0", "maxnlit 3", "mesh", *_mesh_lines(2),
         "loads", "5 39 2 1.0d3 -2.0d3", "1 1 0.5", "end loads",
         "faceR1", "1 2 3", "end faceR1", "insert inc.dat", "end mesh",
         "gravity 1.0 2.0 3.0", "setpnode node.dat", "pgeo", "dt 0.25", "pnup", "loop 3", "solv",
         "pnsf", "next", "dt 0.5", "solv", "loop 2", "solv", "stop"]
    path = tmp_path / "old.dat"
    path.write_text("\r\n".join(L) + "\r\n")
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    inp = read_deck(path)
    assert inp.cfg.solver == "pcg" and inp.cfg.tol == 2.0 ** -53 and inp.cfg.maxnlit_accept
    assert inp.reorder is False and inp.ie.tolist() == [17]
    assert inp.program[-2:] == [("dt", 0.5), ("solv",)]
    assert_same(f, inp)


@pytest.mark.parametrize("bad,msg", [
    ("1 5 1.0 0.3 1.0 1.0 1.0 1.0 1.0 1.0", "element type"),
    ("1 17 1.0 0.3 1.0 1.0 1.0 1.0 1.0 1.0\nend materials\nhexa20", "outside"),
])
def test_rejected_decks(tmp_path, bad, msg):
    L = _mesh_lines(1)
    i = L.index("materials")
    L[i + 1] = bad
    path = tmp_path / "bad.dat"
    path.write_text("\n".join(["mesh", *L, "end mesh", "stop"]) + "\n")
    f, log = fparse(path, tmp_path)
    assert f is None and msg in log


def test_conseq_without_elmt18_rejected(tmp_path):
    (tmp_path / "cq.dat").write_text("kozeny_carman\nend\n")
    path = tmp_path / "cq_deck.dat"
    path.write_text("\n".join(["conseq cq.dat", "mesh", *_mesh_lines(1), "end mesh", "stop"]) + "\n")
    f, log = fparse(path, tmp_path)
    assert f is None and "Invalid elements" in log


@pytest.mark.parametrize("before_mesh", [True, False])
def test_setprint_pres_and_vtk_keys(tmp_path, before_mesh):
    """setprint (first word per line until end, every flag cleared first;
    pconsolidation and elplastic dropped at the mesh of an elastic run),
    pres in the program, and the config keys bvtk and legacy_vtk."""
    (tmp_path / "cfg.dat").write_text("reord true\nbvtk true\nlegacy_vtk true\nend\n")
    words = ["totalstress", "darcyflux x", "porosity", "dpressure"]
    if before_mesh:
        words += ["quadratic", "pconsolidation"]
    (tmp_path / "print.dat").write_text("\n".join(words) + "\nend\nbiotstress\n")
    sp = ["setprint print.dat"]
    L = ["config cfg.dat", *(sp if before_mesh else []), "mesh", *_mesh_lines(2), "end mesh",
         *([] if before_mesh else sp), "setpnode node.dat", "pres", "loop 2", "solv", "pres",
         "next", "stop"]
    (tmp_path / "node.dat").write_text("1\n1\n")
    path = tmp_path / "sp.dat"
    path.write_text("\n".join(L) + "\n")
    f, log = fparse(path, tmp_path)
    assert f is not None, log
    inp = read_deck(path)
    fl = inp.print_flags
    assert fl[0] == before_mesh and fl[4] and fl[7] and fl[8] and fl[3]
    assert not (fl[1] or fl[2] or fl[5] or fl[9])
    assert inp.vtk_legacy and inp.vtk_binary
    assert [op[0] for op in inp.program].count("pres") == 3
    assert_same(f, inp)


@pytest.mark.parametrize("text,msg", [("quadratic\nend\n", "quadratic"),
                                      ("elplastic\nend\n", "plastic path")])
def test_setprint_after_mesh_rejected(tmp_path, text, msg):
    (tmp_path / "print.dat").write_text(text)
    path = tmp_path / "spx.dat"
    path.write_text("\n".join(["mesh", *_mesh_lines(1), "end mesh", "setprint print.dat",
                               "stop"]) + "\n")
    f, log = fparse(path, tmp_path)
    assert f is None and msg in log
    with pytest.raises(Exception, match=msg):
        read_deck(path)
