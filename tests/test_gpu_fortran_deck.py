"""The Fortran host as the reference's ``mef <input> <output>``: pm_mefdrive
reads a deck with its own reader (pm_deck.f90), drives the GPU through the
pmhip iso_c_binding module and writes the output files with Fortran edit
descriptors.  Both hosts call the same deterministic library in the same
order, so every history and log file must equal, byte for byte, the one the
Python runner (poromec_amd.mef) writes for the same deck -- whose histories
the deck tests hold against the oracle and the reference's printed
results."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from deckcases import extract, reference
from poromec_amd import problem as P
from poromec_amd.deck import read_deck, truncate, write_deck
from poromec_amd.mef import PNSF_FILES, read_printnode, run
from poromec_amd.solver import SolverConfig

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "poromec_amd" / "fortran" / "pm_mefdrive"


def _run_both(path, tmp_path, nsteps, fp32ir=False):
    if not EXE.exists():
        pytest.skip("Fortran host not built (amdflang absent)")
    inp = read_deck(path)
    if nsteps:
        inp = truncate(inp, nsteps)
    py = str(tmp_path / "py")
    run(inp, py, precision="fp32ir" if fp32ir else "fp64")
    fo = str(tmp_path / "fo")
    args = [str(EXE), str(path), fo, str(nsteps or 0)] + (["fp32ir"] if fp32ir else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    names = [f"_up_node_{n}.txt" for n in inp.pnodes]
    if any(op[0] == "pnsf" for op in inp.program):
        names += [f"_{stem}_node_{n}.txt" for stem, _ in PNSF_FILES for n in inp.pnodes]
    names += ["_log_1.txt", "_log_nonlinear_1.txt"]
    for name in names:
        a, b = Path(py + name).read_text(), Path(fo + name).read_text()
        assert a == b, (name, [(x, y) for x, y in zip(a.splitlines(), b.splitlines()) if x != y][:3])
    tf = Path(fo + "_t_1.txt").read_text().splitlines()
    tp = Path(py + "_t_1.txt").read_text().splitlines()
    assert tf[0] == tp[0] and tf[-5:] == tp[-5:]  # the system sizes
    return inp, fo


@pytest.mark.parametrize("case,nsteps", [("exemplo1_mesh0_t2", 0), ("solo1", 6),
                                         ("solo_cilindro_hexa_1", 6)])
def test_fortran_host_deck(tmp_path, case, nsteps):
    print(f"deck {case} steps {nsteps or 'all'}", flush=True)
    path = extract(case, tmp_path)
    inp, fo = _run_both(path, tmp_path, nsteps)
    ref = reference(case)
    # the reference's printed histories, as printed by this host
    for n in inp.pnodes:
        st, t, u = read_printnode(f"{fo}_up_node_{n}.txt")
        r = ref["hist"][n][: len(st)]
        assert np.array_equal(st, r[:, 0].astype(int))
        assert np.allclose(u, r[:, 2:], rtol=5e-9, atol=1e-9 * np.abs(r[:, 2:]).max())


def test_fortran_host_fp32ir(tmp_path):
    """FP32 matrix values inside FP64 refinement (SQMR, tol 1e-10)."""
    print("written column deck, fp32ir, 4 steps", flush=True)
    deck = P.column_deck(2, 2, 6, dt=0.05, load_mpa=0.5)
    path = write_deck(tmp_path / "c.dat", deck, SolverConfig(stop_crit=1), pnodes=(1, 7),
                      nsteps=4)
    _run_both(path, tmp_path, 0, fp32ir=True)


def test_fortran_host_q1_deck(tmp_path):
    """A maxno-8 deck (the Q1/Q1 element, type 7) through the Fortran host:
    histories, stresses and logs as the Python runner writes them, and a
    results file with the setprint word quadratic still in linear hexa8
    cells (the mesh has no mid-edge nodes)."""
    from poromec_amd import vtk as V
    print("written Q1/Q1 deck, 4 steps", flush=True)
    deck = P.heterogeneous_box_deck(3, seed=4, dt=0.05, nsteps=4, nen=8)
    path = write_deck(tmp_path / "q1.dat", deck, SolverConfig(stop_crit=1), pnodes=(1, 20),
                      nsteps=4)
    (tmp_path / "q1_print.dat").write_text("quadratic\ndesloc\npressure\ntotalstress\nend\n")
    text = Path(path).read_text().replace("mesh\n", "setprint q1_print.dat\nmesh\n", 1)
    Path(path).write_text(text.replace("pnup\nnext", "pnup\npnsf\npres\nnext"))
    inp, fo = _run_both(path, tmp_path, 0)
    assert inp.deck.nen == 8 and inp.print_flags[0]
    py = str(tmp_path / "py")
    a = Path(V.results_name(py, 4, False)).read_bytes()
    b = Path(V.results_name(fo, 4, False)).read_bytes()
    la, lb = a.split(b"\n"), b.split(b"\n")
    for j in range(len(la)):
        if b'Name="TIME_' in la[j]:
            assert float(la[j + 1]) == float(lb[j + 1])
            la[j + 1] = lb[j + 1]
    assert b"\n".join(la) == b
    t = b.decode()
    types = t.split('Name="types" format="ascii">')[1].split("</DataArray>")[0].split()
    assert types == ["12"] * deck.numel


def test_fortran_host_vprop_deck(tmp_path):
    """elmt18 (properties at the integration points, Kozeny-Carman and
    Hashin-Shtrikman) and a time-table load through the Fortran host."""
    print("written elmt18 deck, 5 steps", flush=True)
    deck = P.column_deck(2, 2, 6, dt=0.05, load_mpa=0.5)
    deck.e[:, 5] = 2000.0
    deck.e[:, 7] = 0.3
    u0 = np.zeros((deck.nnodev, 4))
    u0[:, 3] = 1.0e3
    path = write_deck(tmp_path / "v18.dat", deck, SolverConfig(stop_crit=1), u0=u0,
                      pnodes=(1, deck.nnodev), nsteps=5, element_type=18,
                      vprop=(True, True, False, True))
    with open(path) as f:
        text = f.read()
    Path(path).write_text(text.replace("pnup\nnext", "pnup\npnsf\nnext"))
    inp, _ = _run_both(path, tmp_path, 0)
    assert inp.vprop[0] and (inp.ie == 18).all()


@pytest.mark.parametrize("legacy,binary,quad", [(False, False, True), (True, False, False),
                                                (True, True, True)])
def test_fortran_host_results_files(tmp_path, legacy, binary, quad):
    """pres through both hosts: every field setprint offers, over the
    vertices or every node, in the three encodings; the same bytes (the VTU
    time values: the same double), and the pressures and stresses in the
    files are the printed histories' values to the 5 digits written."""
    from poromec_amd import vtk as V
    print(f"pres legacy={legacy} binary={binary} quad={quad}", flush=True)
    deck = P.column_deck(2, 2, 5, dt=0.05, load_mpa=0.5)
    path = write_deck(tmp_path / "r.dat", deck, SolverConfig(stop_crit=1), pnodes=(1, 9),
                      nsteps=3)
    (tmp_path / "r_config.dat").write_text(
        f"reord true\nlegacy_vtk {str(legacy).lower()}\nbvtk {str(binary).lower()}\nend\n")
    words = [w for w in V.PRINT_WORDS[1:9]] + (["quadratic"] if quad else [])
    (tmp_path / "r_print.dat").write_text("\n".join(words) + "\nend\n")
    text = Path(path).read_text().replace("mesh\n", "setprint r_print.dat\nmesh\n", 1)
    text = text.replace("pnup\nnext", "pnup\npnsf\npres\nnext")
    Path(path).write_text(text)
    inp, fo = _run_both(path, tmp_path, 0)
    assert inp.print_flags[0] == quad and inp.vtk_legacy == legacy
    py = str(tmp_path / "py")
    for istep in (1, 2, 3):
        a = Path(V.results_name(py, istep, legacy)).read_bytes()
        b = Path(V.results_name(fo, istep, legacy)).read_bytes()
        if not legacy:
            la, lb = a.split(b"\n"), b.split(b"\n")
            for j in range(len(la)):
                if b'Name="TIME_' in la[j]:
                    assert float(la[j + 1]) == float(lb[j + 1])
                    la[j + 1] = lb[j + 1]
            a = b"\n".join(la)
        assert a == b, istep
    if legacy and not binary:
        L = Path(V.results_name(py, 3, True)).read_text().split("\n")
        i = L.index(V._scalar_head("pressao".ljust(15), "double ")) + 2
        p = np.array([float(v) for v in L[i:i + deck.nnodev]])
        j = L.index("TENSORS " + "totalStress".ljust(15) + " double") + 1
        sx = np.array([[float(v) for v in ln.split()] for ln in L[j:j + deck.nnodev]])
        for n in inp.pnodes:
            _, _, up = read_printnode(f"{py}_up_node_{n}.txt")
            assert np.isclose(p[n - 1], up[-1, 3], rtol=1e-4, atol=1e-9 * abs(up[:, 3]).max())
            _, _, st = read_printnode(f"{py}_stress_node_{n}.txt")
            full = st[-1][[0, 3, 5, 3, 1, 4, 5, 4, 2]]
            assert np.allclose(sx[n - 1], full, rtol=1e-4, atol=1e-6 * abs(st).max())
