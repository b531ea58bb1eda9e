"""The deck reader and the output writers (poromec_amd.deck, poromec_amd.mef)
on the CPU.

The example decks the reference ships with its outputs
(output/poro_mec/elastico/hexa) pin the reader from two sides: the system
size the reference logged for each (nnode, neq, nad in its _t_1.txt) must
come out of the parsed mesh, and the oracle's march of the parsed deck must
give the reference's own up_node, stress, stressE, stressB and flux
histories, at every one of exemplo1's 600 steps, to the 10 decimals they
were printed with.  A misread coordinate,
connectivity, restraint, face load, material or initial pressure moves
one of them.  Write -> read round trips and hand-checked small decks cover
the current grammar (file-based config/solver/gravity/loads), and amdflang
itself checks the (i9,10e18.10) writer."""
import shutil
import subprocess

import numpy as np
import pytest

from deckcases import CASES, KINDS, extract, oracle_run, print_close, reference
from poromec_amd import problem as P
from poromec_amd.deck import EPS, DeckError, read_deck, truncate, write_deck
from poromec_amd.mef import fortran_e, printnode_line, read_printnode
from poromec_amd.solver import SolverConfig, mesh_numbering


@pytest.mark.parametrize("case", CASES)
def test_system_size_matches_reference_log(tmp_path, case):
    inp = read_deck(extract(case, tmp_path))
    ref = reference(case)
    assert inp.deck.numel == ref["numel"]
    r = mesh_numbering(inp.deck, inp.reorder)
    assert (r["nnode"], r["neq"], r["nad"]) == (ref["nnode"], ref["neq"], ref["nad"])


def test_example_deck_settings(tmp_path):
    a = read_deck(extract("exemplo1_mesh0_t2", tmp_path))
    assert (a.cfg.solver, a.cfg.maxit, a.cfg.tol) == ("pcg", 100000, 1.0e-11)
    assert (a.cfg.maxnlit, a.cfg.stop_crit, a.cfg.maxnlit_accept) == (2, 1, True)
    assert a.reorder is False and a.pnodes == [1, 44]
    assert a.deck.gravity.tolist() == [0.0, 0.0, -10.0] and a.deck.gravity_mod == 10.0
    assert a.program[:5] == [("dt", 3600.0), ("pnup",), ("pnsf",), ("solv",), ("pnup",)]
    assert sum(op == ("solv",) for op in a.program) == 600
    # face 1 of element 1 carries load 1: a constant traction (0, 0, -3.9)
    assert a.deck.eload[0].tolist() == [1, 0, 0, 0, 0, 0, 0] and not a.deck.eload[1:].any()
    assert a.deck.load[0].tolist() == [40, 4, 2]
    assert a.deck.fload[0, :4, 0].tolist() == [0.0, 0.0, 0.0, -3.9]
    # initial pressures, zeroed on the drained (restrained) top
    assert a.u0[1, 3] == 1.0e-3 and a.u0[43, 3] == 1.0e-2 and a.u0[0, 3] == 0.0
    assert "pgeo" in a.skipped and ("pnsf",) in a.program
    b = read_deck(extract("solo1", tmp_path))
    assert b.cfg.tol == EPS  # "pcg 10000 0.0d0": tol 0 -> machine epsilon, as logged
    c = read_deck(extract("solo_cilindro_hexa_1", tmp_path))
    assert c.cfg.maxnlit == 1 and c.deck.eload[:, 1].any()  # loads on face 2 too


@pytest.mark.parametrize("case,nsteps", [("exemplo1_mesh0_t2", 600), ("solo1", 4),
                                         ("solo_cilindro_hexa_1", 4)])
def test_oracle_reproduces_reference_histories(tmp_path, case, nsteps):
    """Displacements and pressure (pnup) and the vertex stresses and flux
    (pnsf: total, Terzaghi, Biot, Darcy) the reference printed."""
    inp = read_deck(extract(case, tmp_path))
    stress = {}
    hist, iters, _ = oracle_run(inp, nsteps, stress=stress)
    ref = reference(case)
    for kind, mine in [("hist", hist)] + [(k, stress[k]) for k in KINDS]:
        for n, h in mine.items():
            r = ref[kind][n][: nsteps + 1]
            assert h.shape == r.shape, (kind, n)
            assert np.array_equal(h[:, 0], r[:, 0])
            assert np.allclose(h[:, 1], r[:, 1], rtol=1e-12)
            ok = print_close(h[:, 2:], r[:, 2:])
            assert ok.all(), (kind, n, np.argwhere(~ok)[:5])


def test_write_read_round_trip(tmp_path):
    d = P.heterogeneous_box_deck(3, seed=1)
    u0 = np.zeros((d.nnodev, 4))
    u0[:, 3] = np.linspace(0.0, 1.0, d.nnodev)
    cfg = SolverConfig(solver="minres", tol=1.0e-9, maxnlit=3, stop_crit=1, precond=5)
    p = write_deck(tmp_path / "box.dat", d, cfg, u0=u0, pnodes=(1, 5), nsteps=4)
    inp = read_deck(p)
    k = inp.deck
    for name in ("ix", "x", "idv", "e", "eload", "eloadp", "fload", "gravity"):
        assert np.array_equal(getattr(k, name), getattr(d, name)), name
    assert np.array_equal(k.load[:, [0, 2]], d.load[:, [0, 2]])
    assert (k.gravity_mod, k.dt) == (d.gravity_mod, d.dt)
    c = inp.cfg
    assert (c.solver, c.tol, c.maxnlit, c.stop_crit, c.precond) == ("minres", 1e-9, 3, 1, 5)
    assert inp.pnodes == [1, 5] and inp.reorder is True
    want = u0.copy()
    want[d.idv > 0] = 0.0
    assert np.array_equal(inp.u0, want)
    assert inp.program == [("dt", d.dt), ("pnup",)] + [("solv",), ("pnup",)] * 4
    assert truncate(inp, 2).program == [("dt", d.dt), ("pnup",)] + [("solv",), ("pnup",)] * 2


def test_q1_deck_round_trip(tmp_path):
    """A maxno-8 deck (the Q1/Q1 element, type 7) written and read back:
    the mesh stays hexa8 (nen 8) and the materials type 7; a maxno-8 deck
    naming another element type is refused."""
    d = P.heterogeneous_box_deck(3, seed=2, nen=8)
    p = write_deck(tmp_path / "q1.dat", d, SolverConfig(stop_crit=1), nsteps=2)
    text = (tmp_path / "q1.dat").read_text()
    assert " maxno 8 " in text
    inp = read_deck(p)
    assert inp.deck.nen == 8 and np.all(inp.ie == 7)
    for name in ("ix", "x", "idv", "e", "eload", "eloadp"):
        assert np.array_equal(getattr(inp.deck, name), getattr(d, name)), name
    from poromec_amd.solver import mesh_numbering
    assert mesh_numbering(inp.deck)["nnode"] == d.nnodev
    (tmp_path / "q1.dat").write_text(text.replace("\n1 7 ", "\n1 17 "))
    with pytest.raises(DeckError):
        read_deck(tmp_path / "q1.dat")


CURRENT = """config cfg.dat
solver slv.dat
gravity grav.dat
nl 7 1.0d-6 2
mesh
nnode 12 numel 2 numat 2 maxno 20 ndf 4 dim 3
fmaterials
2 mat2.dat
end fmaterials
materials
1 17 10.0 0.25 1.0d-10 1000.0 1.0 2.0 1000.0
end materials
insert coor.dat
hexa8
1 2 1 3 4 6 5 7 8 1
2 6 5 7 8 10 9 11 12 2
end hexa8
constrainpmec
1 1 1 1 0
12 0 0 0 1
end constrainpmec
elmtloads
1 3 0 0 0 0 0
end elmtloads
elmtpresloads
2 0 4
end elmtpresloads
loads
3 40 trac.dat
4 43 flux.dat
end loads
hydrostatic hyd.dat
end mesh
setpnode pn.dat
dt 0.5
pnup
loop 3
solv
pnup
next
dt 0.25
solv
stop
"""


def _current_deck(tmp_path):
    (tmp_path / "deck.dat").write_text(CURRENT)
    (tmp_path / "cfg.dat").write_text("memory 100\nreord false\nbvtk true\nend\n")
    (tmp_path / "slv.dat").write_text("name gmres\nit 900\ntol 0.0\nprecond diagm\nnkrylov 40\n"
                                      "histlog false\nend\n")
    (tmp_path / "grav.dat").write_text("0.0 0.0 -9.0 9.5\n")
    (tmp_path / "mat2.dat").write_text("elType 17\nmodE 20.0\npoisson 0.3\npermeability 2.0\n"
                                       "mbiot 500.0\ncbiot 0.9\ndensity 1.5\nfdensity 1000.0\n"
                                       "porosity 0.3\nend\n")
    xs = [(i % 2, (i // 2) % 2, i // 4) for i in range(12)]
    (tmp_path / "coor.dat").write_text(
        "coordinates\n" + "".join(f"{k + 1} {x}.0d0 {y}.0 {z}.0e0\n" for k, (x, y, z) in
                                  enumerate(xs)) + "end coordinates\nreturn\n")
    (tmp_path / "trac.dat").write_text("4 3\n0.0 1.0 2.0 3.0\n1.0 4.0 5.0 6.0\n2.0 7.0 8.0 9.0\n")
    (tmp_path / "flux.dat").write_text("2 2\n0.0 -1.0\n10.0 -2.0\n")
    (tmp_path / "hyd.dat").write_text("1000.0 0.1 2.0 1.0d-6\n")
    (tmp_path / "pn.dat").write_text("2\n3\n12\n")
    return read_deck(tmp_path / "deck.dat")


def test_current_grammar(tmp_path):
    inp = _current_deck(tmp_path)
    d, c = inp.deck, inp.cfg
    assert inp.reorder is False
    assert (c.solver, c.maxit, c.tol, c.precond, c.ngram) == ("gmres", 900, EPS, 5, 40)
    assert (c.maxnlit, c.nltol, c.stop_crit, c.maxnlit_accept) == (7, 1e-6, 2, False)
    assert d.gravity.tolist() == [0.0, 0.0, -9.0] and d.gravity_mod == 9.5
    assert d.e[0, :7].tolist() == [10.0, 0.25, 1e-10, 1000.0, 1.0, 2.0, 1000.0]
    assert d.e[1, :8].tolist() == [20.0, 0.3, 2.0, 500.0, 0.9, 1.5, 1000.0, 0.3]
    assert d.ix[0, :8].tolist() == [2, 1, 3, 4, 6, 5, 7, 8] and d.ix[:, 20].tolist() == [1, 2]
    assert d.x[11].tolist() == [1.0, 1.0, 2.0]
    assert d.idv[0].tolist() == [1, 1, 1, 0] and d.idv[11].tolist() == [0, 0, 0, 1]
    assert d.idv.sum() == 4
    assert d.eload[0].tolist() == [3, 0, 0, 0, 0, 0, 0] and d.eloadp[1, 1] == 4
    assert d.load[2].tolist() == [40, 4, 3] and d.load[3].tolist() == [43, 2, 2]
    # fload[nc-1, term, time]: term 0 the times, then the values
    assert d.fload[2, :4, :3].tolist() == [[0, 1, 2], [1, 4, 7], [2, 5, 8], [3, 6, 9]]
    assert d.fload[3, :2, :2].tolist() == [[0, 10], [-1, -2]]
    # hydrostatic: p = (p0 + rho |g| (h - z)) * scale, zeroed where restrained
    z = d.x[:, 2]
    want = (0.1 + 1000.0 * 9.5 * (2.0 - z)) * 1e-6
    want[11] = 0.0
    assert np.allclose(inp.u0[:, 3], want, rtol=1e-15, atol=0)
    assert inp.pnodes == [3, 12]
    assert inp.program == ([("dt", 0.5), ("pnup",)] + [("solv",), ("pnup",)] * 3
                           + [("dt", 0.25), ("solv",)])


@pytest.mark.parametrize("text,msg", [
    ("initialstress\n1 0 0 0 0 0 0\nend initialstress\n", "initialstress"),
    ("nodalforces\n1 0 0 1 0\nend nodalforces\n", "nodalforces"),
    ("hexa20\n", "hexa20"),
    ("loads\n1 4 1.0\nend loads\n", "load type 4"),
    ("loads\n1 40 missing.dat\nend loads\n", "not found"),
    ("constraindisp\nend constraindisp\n", "constraindisp"),
])
def test_unsupported_mesh_macros(tmp_path, text, msg):
    deck = CURRENT.replace("hydrostatic hyd.dat\n", text)
    _current_deck(tmp_path)
    (tmp_path / "deck.dat").write_text(deck)
    with pytest.raises(DeckError, match=msg):
        read_deck(tmp_path / "deck.dat")


@pytest.mark.parametrize("old,new,msg", [
    ("maxno 20", "maxno 8", "maxno"),
    ("ndf 4", "ndf 3", "ndf"),
    ("1 17 10.0", "1 15 10.0", "element type"),
    ("nl 7 1.0d-6 2", "nl 7 1.0d-6 3", "criterion"),
    ("name gmres", "name pardiso", "pardiso"),
    ("config cfg.dat", "block_pu 2\nconfig cfg.dat", "block_pu"),
])
def test_rejected_settings(tmp_path, old, new, msg):
    _current_deck(tmp_path)
    for f in ("deck.dat", "slv.dat"):
        p = tmp_path / f
        p.write_text(p.read_text().replace(old, new))
    with pytest.raises(DeckError, match=msg):
        read_deck(tmp_path / "deck.dat")


def test_fortran_e_edge_cases():
    assert fortran_e(0.0) == "  0.0000000000E+00"
    assert fortran_e(3600.0) == "  0.3600000000E+04"
    assert fortran_e(-0.0159674582) == " -0.1596745820E-01"
    assert fortran_e(1.0e-120) == "  0.1000000000-119"
    assert fortran_e(0.99999999999) == "  0.1000000000E+01"
    assert printnode_line(7, 25200.0, [0.0, -1.5, 2.0e-5, 3.9]) == (
        "        7  0.2520000000E+05  0.0000000000E+00 -0.1500000000E+01"
        "  0.2000000000E-04  0.3900000000E+01")


@pytest.mark.skipif(not shutil.which("/opt/rocm/bin/amdflang"), reason="amdflang absent")
def test_printnode_format_matches_fortran(tmp_path):
    """The writer against the Fortran compiler's own (i9,10e18.10) output."""
    rng = np.random.default_rng(3)
    vals = np.concatenate([[0.0, 1.0, -1.0, 0.5, 1e-99, 1e-100, -3.25e150, 123456789.0,
                            0.99999999995, 0.12345678905, 2.0 ** -60],
                           rng.standard_normal(40) * 10.0 ** rng.integers(-30, 30, 40)])
    np.savetxt(tmp_path / "v.txt", vals, fmt="%.17e")
    (tmp_path / "w.f90").write_text("""program w
  implicit none
  real(8) :: v(51)
  integer :: i
  open(1, file='v.txt', action='read')
  read(1, *) v
  close(1)
  do i = 1, 51, 5
    write(*, '(i9,10e18.10)') i, v(i:min(i + 4, 51))
  end do
end program
""")
    subprocess.run(["/opt/rocm/bin/amdflang", "w.f90", "-o", "w"], cwd=tmp_path, check=True,
                   capture_output=True)
    out = subprocess.run(["./w"], cwd=tmp_path, check=True, capture_output=True,
                         text=True).stdout.splitlines()
    for k, line in enumerate(out):
        i = 1 + 5 * k
        chunk = vals[i - 1:min(i + 4, 51)]
        mine = f"{i:9d}" + "".join(fortran_e(v) for v in chunk)
        assert mine == line.rstrip(), (mine, line)


def test_read_printnode_round_trip(tmp_path):
    p = tmp_path / "h.txt"
    rows = [(0, 0.0, [0.0, 0.0, 0.0, 0.01]), (1, 3600.0, [1e-3, -2e-7, -0.0159674582, 3.9])]
    p.write_text("# Variacao DeslocAndPress no tempo no ponto        1\n"
                 + "".join(printnode_line(s, t, u) + "\n" for s, t, u in rows))
    st, t, u = read_printnode(p)
    assert st.tolist() == [0, 1] and t.tolist() == [0.0, 3600.0]
    assert np.allclose(u, [r[2] for r in rows], rtol=1e-10, atol=0)


def test_conseq_and_elmt18_round_trip(tmp_path):
    d = P.column_deck(2, 2, 3)
    d.e[:, 7] = 0.3
    p = write_deck(tmp_path / "v.dat", d, SolverConfig(stop_crit=1), element_type=18,
                   vprop=(True, True, False, True), nsteps=2)
    inp = read_deck(p)
    assert inp.ie.tolist() == [18] * d.e.shape[0]
    assert inp.vprop == (True, True, False, True)
    assert np.array_equal(inp.deck.e, d.e)


@pytest.mark.parametrize("etype,flags,msg", [
    (18, None, "Invalid elements"),
    (17, (True, False, False, False), "Invalid elements"),
])
def test_conseq_element_check(tmp_path, etype, flags, msg):
    d = P.column_deck(2, 2, 3)
    p = write_deck(tmp_path / "v.dat", d, element_type=etype, vprop=flags, nsteps=1)
    with pytest.raises(DeckError, match=msg):
        read_deck(p)


def test_conseq_plastic_rejected(tmp_path):
    d = P.column_deck(2, 2, 3)
    p = write_deck(tmp_path / "v.dat", d, element_type=18, vprop=(True, False, False, False),
                   nsteps=1)
    (tmp_path / "v_conseq.dat").write_text("plastic\nvprop\nend\n")
    with pytest.raises(DeckError, match="plastic"):
        read_deck(p)
