"""The reference's example decks run end to end on the GPU through the deck
reader and the output writers (poromec_amd.mef): the up_node files must hold
the oracle's histories of the same deck (1e-9, and bit for bit in reference
summation order) and the reference's own printed results."""
import dataclasses

import numpy as np
import pytest

from deckcases import KINDS, extract, oracle_run, print_close, reference
from poromec_amd.deck import read_deck, truncate
from poromec_amd.mef import PNSF_FILES, read_printnode, run

pytestmark = pytest.mark.gpu


def _check_files(prefix, inp, nsteps, res):
    hist = {}
    for n in inp.pnodes:
        path = f"{prefix}_up_node_{n}.txt"
        lines = open(path).read().splitlines()
        assert lines[0] == f"# Variacao DeslocAndPress no tempo no ponto{n:9d}"
        assert len(lines) == nsteps + 2
        st, t, u = read_printnode(path)
        assert st.tolist() == list(range(nsteps + 1))
        # the text is the run's history to the 10 significant digits printed
        assert np.allclose(u, res["hist"][n][:, 2:], rtol=5.01e-10, atol=1e-300)
        hist[n] = np.column_stack([st, t, u])
    for stem, name in PNSF_FILES:
        for n in inp.pnodes:
            path = f"{prefix}_{stem}_node_{n}.txt"
            lines = open(path).read().splitlines()
            assert lines[0] == f"# Variacao {name} no tempo no ponto{n:9d}"
            assert len(lines) == nsteps + 2
            st, t, v = read_printnode(path)
            assert np.allclose(v, res["stress"][stem][n][:, 2:], rtol=5.01e-10, atol=1e-300)
    for suffix in ("_log_1.txt", "_log_nonlinear_1.txt", "_t_1.txt"):
        assert open(prefix + suffix).readline()
    return hist


@pytest.mark.parametrize("case,nsteps", [("exemplo1_mesh0_t2", 600), ("solo1", 8),
                                         ("solo_cilindro_hexa_1", 8)])
def test_deck_run_matches_oracle_and_reference(tmp_path, case, nsteps):
    print(f"deck {case} steps {nsteps}", flush=True)
    inp = truncate(read_deck(extract(case, tmp_path)), nsteps)
    prefix = str(tmp_path / case)
    res = run(inp, prefix)
    hist = _check_files(prefix, inp, nsteps, res)
    ostress = {}
    ohist, oiters, _ = oracle_run(inp, nsteps, stress=ostress)
    ref = reference(case)
    # stresses and flux: 1e-9 of the largest value of each kind the print
    # nodes see; the reference's printed values to their 10 decimals
    for kind in KINDS + ("poro",):
        allo = np.concatenate([ostress[kind][n][:, 2:] for n in inp.pnodes])
        tol = 1e-9 * np.abs(allo).max() + 1e-300
        for n in inp.pnodes:
            g = res["stress"][kind][n]
            assert (np.abs(g[:, 2:] - ostress[kind][n][:, 2:]) <= tol).all(), (kind, n)
            if kind in ref:
                ok = print_close(g[:, 2:], ref[kind][n][: nsteps + 1, 2:])
                assert ok.all(), (kind, n, np.argwhere(~ok)[:5])
    # 1e-9 of the largest displacement / pressure the print nodes see (a
    # component that is zero by symmetry carries rounding noise only)
    allo = np.concatenate([ohist[n][:, 2:] for n in inp.pnodes])
    scale = np.array([np.abs(allo[:, :3]).max()] * 3 + [np.abs(allo[:, 3]).max()])
    for n in inp.pnodes:
        g, o, r = res["hist"][n], ohist[n], ref["hist"][n][: nsteps + 1]
        assert (np.abs(g[:, 2:] - o[:, 2:]) <= 1e-9 * scale).all(), n
        ok = print_close(hist[n][:, 2:], r[:, 2:])
        assert ok.all(), (n, np.argwhere(~ok)[:5])
    # the same number of solves per step as the oracle; PCG on this
    # indefinite u-p system moves its count by more than SQMR does at the
    # rounding level (the reference's own log differs from the oracle's at
    # step 2: 52 against 49), so the Krylov work is compared in total
    assert [len(s["iters"]) for s in res["stats"]] == [len(i) for i in oiters]
    tg = sum(sum(s["iters"]) for s in res["stats"])
    to = sum(sum(i) for i in oiters)
    worst = max(abs(a - b) for s, i in zip(res["stats"], oiters) for a, b in zip(s["iters"], i))
    print(f"Krylov iterations: GPU {tg} oracle {to}, largest per-solve difference {worst}")
    assert abs(tg - to) <= 0.1 * to


def test_deck_ordered_bitwise(tmp_path):
    """In reference summation order the printed histories are the oracle's,
    digit for digit, and the state bit for bit."""
    nsteps = 40
    inp = truncate(read_deck(extract("exemplo1_mesh0_t2", tmp_path)), nsteps)
    res = run(inp, str(tmp_path / "ord"), ordered=True)
    ostress = {}
    ohist, oiters, _ = oracle_run(inp, nsteps, stress=ostress)
    for n in inp.pnodes:
        assert np.array_equal(res["hist"][n], ohist[n]), n
        for kind in KINDS + ("poro",):
            assert np.array_equal(res["stress"][kind][n], ostress[kind][n]), (kind, n)
    assert [s["iters"] for s in res["stats"]] == oiters


def test_deck_fp32ir(tmp_path):
    """FP32 matrix values inside FP64 refinement reach the same histories."""
    nsteps = 8
    inp = truncate(read_deck(extract("solo1", tmp_path)), nsteps)
    inp.cfg.tol = 1e-11  # the deck's tolerance is machine epsilon, beyond FP32+IR's reach
    res = run(inp, str(tmp_path / "ir"), precision="fp32ir")
    ref = reference("solo1")
    allr = np.concatenate([ref["hist"][n][: nsteps + 1, 2:] for n in inp.pnodes])
    scale = np.array([np.abs(allr[:, :3]).max()] * 3 + [np.abs(allr[:, 3]).max()])
    for n in inp.pnodes:
        g, r = res["hist"][n], ref["hist"][n][: nsteps + 1]
        assert (np.abs(g[:, 2:] - r[:, 2:]) <= 1e-8 * scale + 1e-9).all(), n


def test_q1_deck_run(tmp_path):
    """A maxno-8 deck (the Q1/Q1 element, type 7) written by write_deck and
    run end to end by the deck runner: the up_node histories hold the
    oracle's march of the same deck to 1e-9, and a results file is written
    with linear hexa8 cells."""
    from poromec_amd import problem as P
    from poromec_amd.deck import write_deck
    from poromec_amd.solver import SolverConfig
    d = P.heterogeneous_box_deck(4, seed=7, dt=0.1, nsteps=4, nen=8)
    path = write_deck(tmp_path / "q1.dat", d, SolverConfig(stop_crit=1, tol=1e-11),
                      pnodes=(1, 30, 60), nsteps=4)
    inp = read_deck(path)
    assert inp.deck.nen == 8
    inp = dataclasses.replace(inp, program=inp.program + [("pres",)])
    prefix = str(tmp_path / "q1")
    res = run(inp, prefix)
    ohist, oiters, _ = oracle_run(inp, 4)
    allo = np.concatenate([ohist[n][:, 2:] for n in inp.pnodes])
    scale = np.array([np.abs(allo[:, :3]).max()] * 3 + [np.abs(allo[:, 3]).max()])
    for n in inp.pnodes:
        assert (np.abs(res["hist"][n][:, 2:] - ohist[n][:, 2:]) <= 1e-9 * scale).all(), n
    import glob
    vtu = sorted(glob.glob(prefix + "_istep_*.vtu"))
    assert vtu
    text = open(vtu[-1]).read()
    types = text.split('Name="types" format="ascii">')[1].split("</DataArray>")[0].split()
    assert types == ["12"] * d.numel  # VTK_HEXAHEDRON: 8 vertices per cell
    assert f'NumberOfPoints="{d.nnodev:15d}"' in text
