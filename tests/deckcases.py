"""The reference's example decks (tests/golden/decks, written by
tests/golden/make_deck_fixtures.py) and the oracle's run of a deck's macro
program, for the deck tests."""
from __future__ import annotations

import tarfile
from pathlib import Path

import numpy as np

from oracle import oracle as O
from poromec_amd.solver import SOLVERS

GOLD = Path(__file__).resolve().parent / "golden" / "decks"
CASES = ("exemplo1_mesh0_t2", "solo1", "solo_cilindro_hexa_1")


def extract(case: str, dest: Path) -> Path:
    """Unpacks a case's deck files into dest; returns the main deck's path."""
    with tarfile.open(GOLD / f"{case}.tar.gz") as tar:
        for m in tar.getmembers():
            if not m.isfile() or "/" in m.name or m.name.startswith("."):
                raise ValueError(f"unexpected member {m.name}")
            (dest / m.name).write_bytes(tar.extractfile(m).read())
    return dest / f"{case}.dat"


KINDS = ("stress", "stressE", "stressB", "flux")


def reference(case: str) -> dict:
    """The reference's results for a case: its logged sizes and, per print
    node, its printed histories: "hist" the up_node rows (istep, t, u, v, w,
    p), and per kind of KINDS the pnsf rows (istep, t, components)."""
    z = np.load(GOLD / f"{case}_ref.npz")
    out = {k: int(z[k]) for k in ("nnode", "neq", "nad", "numel")}
    out["hist"] = {int(k[3:]): z[k] for k in z.files if k.startswith("up_")}
    for kind in KINDS:
        out[kind] = {int(k[len(kind) + 1:]): z[k] for k in z.files
                     if k.startswith(kind + "_") and k[len(kind) + 1:].isdigit()}
    return out


def oracle_run(inp, nsteps=None, stress=None):
    """The oracle's run of the deck's program: {node: (n, 6) history}, the
    per-step iteration counts and the final nodal u.  A dict passed as
    `stress` receives the pnsf histories ({kind: {node: rows}}, kinds
    stress, stressE, stressB, flux, poro)."""
    c = inp.cfg
    deck = inp.deck
    st = O.build_structure(deck, reorder=inp.reorder)
    vp = inp.vprop if inp.vprop[0] else None
    m = O.OracleModel(deck, st=st, solver=SOLVERS[c.solver], precond=c.precond, solvtol=c.tol,
                      maxit=c.maxit, ngram=c.ngram, maxnlit=c.maxnlit, tol=c.nltol,
                      stop_crit=c.stop_crit, maxnlit_accept=c.maxnlit_accept,
                      ie=inp.ie if vp else None, vprop=vp)
    nv = deck.nnodev
    m.u0[:nv] = inp.u0
    m.u[:nv] = inp.u0
    hist = {n: [] for n in inp.pnodes}
    iters, istep, t = [], 0, 0.0
    m.env.dt = 1.0
    for op in inp.program:
        if op[0] == "dt":
            m.env.dt = op[1]
        elif op[0] == "pnup":
            for n in inp.pnodes:
                hist[n].append([istep, t, *m.u[n - 1]])
        elif op[0] == "pnsf" and stress is not None:
            arrs = (*m.tform(), m.porosity[:, None])
            for kind, a in zip(KINDS + ("poro",), arrs):
                d = stress.setdefault(kind, {})
                for n in inp.pnodes:
                    d.setdefault(n, []).append([istep, t, *a[n - 1]])
        elif op[0] == "solv":
            if nsteps is not None and istep == nsteps:
                break
            istep += 1
            t = t + m.env.dt
            info = m.step()
            iters.append(info["iters"])
    if stress is not None:
        for kind in stress:
            stress[kind] = {n: np.array(v) for n, v in stress[kind].items()}
    return {n: np.array(v) for n, v in hist.items()}, iters, m.u.copy()


def print_close(a, b, atol=1e-9, rtol=1e-9):
    """Agreement within what a 10-decimal text print resolves."""
    return np.abs(a - b) <= atol + rtol * np.abs(b)
