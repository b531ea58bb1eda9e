"""Runs a reference input deck on the MI355X and writes the reference's
output files: the ``mef`` program (``Main.f``) for the macros that
:mod:`poromec_amd.deck` reads.

    python -m poromec_amd.mef <deck> <output prefix> [--fp32ir] [--ordered]

mirrors ``mef <input> <output>`` (``Main.f:336-402``: the input file, then
the prefix of every output file).  Files written, named as ``name``
(``Filenames.f``) names them:

* ``<prefix>_up_node_<n>.txt`` for every node of ``setpnode``, one line per
  ``pnup``: ``istep, t, u, v, w, p`` in ``(i9,10e18.10)`` under the header
  line of ``printnode`` (``Pdata.f:733-760``);
* for every ``pnsf``, the node histories of the total stress
  ``_stress_node_<n>``, Terzaghi's effective stress ``_stressE_node_<n>``,
  Biot's effective stress ``_stressB_node_<n>`` (xx yy zz xy yz xz), the
  Darcy flux ``_flux_node_<n>`` and the porosity ``_poro_node_<n>``
  (``Main.f:2310-2481``, ``Filenames.f`` codes 31-34, 37);
* ``<prefix>_log_1.txt``: the solver log, one line per linear solve;
* ``<prefix>_log_nonlinear_1.txt``: ``step istep t`` per ``solv`` macro
  (``Main.f:892``) and the nonlinear iterations it took;
* ``<prefix>_t_1.txt``: device times and the sizes of the system;
* for every ``pres``, ``<prefix>_istep_<n>.vtu`` (``.vtk`` with
  ``legacy_vtk``): the mesh and the fields ``setprint`` selects
  (``write_mesh_res_pm``, ``Write_vtk.f:912``; :mod:`poromec_amd.vtk`).

Every step runs in the HIP library (``pm_step``); there is no CPU path.
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
import time

import numpy as np

from . import vtk
from .deck import MefInput, read_deck
from .ftext import fortran_e
from .problem import NDF
from .solver import SOLVER_NAMES, SOLVERS, MefSolver


def printnode_line(istep: int, t: float, u) -> str:
    """One line of a ``printnode`` history: ``(i9,10e18.10)``."""
    return f"{istep:9d}" + fortran_e(t) + "".join(fortran_e(x) for x in u)


def printnode_header(node: int, name: str = "DeslocAndPress") -> str:
    return f"# Variacao {name} no tempo no ponto{node:9d}"


# pnsf outputs: file stem, printnode name (Main.f:2375-2481)
PNSF_FILES = (("stress", "totalStress"), ("stressE", "eSstress"), ("stressB", "biotStress"),
              ("flux", "darcyFlux"), ("poro", "poro"))


def read_printnode(path):
    """(istep, t, values) arrays of a printnode history file."""
    rows = [ln.split() for ln in open(path) if ln.strip() and not ln.startswith("#")]
    a = np.array([[float(v.replace("D", "E")) for v in r] for r in rows])
    return a[:, 0].astype(int), a[:, 1], a[:, 2:]


def run(inp: MefInput, prefix: str, precision: str = "fp64", ordered: bool = False,
        device: int = 0, echo=None) -> dict:
    """Runs the deck's macro program; returns the pnup histories of the
    print nodes ({node: array of rows istep, t, u, v, w, p}), the pnsf ones
    ({"stress" | "stressE" | "stressB" | "flux" | "poro": {node: rows}}) and
    per-step solver statistics."""
    wall0 = time.perf_counter()
    deck = inp.deck
    cfg = dataclasses.replace(inp.cfg, precision=precision)
    s = MefSolver(deck, device=device, reorder=inp.reorder)
    try:
        if ordered:
            s.set_ordered(True)
        if inp.vprop[0]:
            s.set_vprop(inp.ie, inp.vprop)
        nv = deck.nnodev
        u0 = np.zeros((s.nnode, NDF))
        u0[:nv] = inp.u0
        s.set_state(u=u0, u0=u0, t=0.0)
        for n in inp.pnodes:
            if not 1 <= n <= s.nnode:
                raise ValueError(f"print node {n} outside 1..{s.nnode}")
        hist = {n: [] for n in inp.pnodes}
        shist = {k: {n: [] for n in inp.pnodes} for k, _ in PNSF_FILES}
        stats = []
        dt, t, istep = 1.0, 0.0, 0
        new_file = True
        new_sf = True
        ix20 = None
        ms_elem = ms_solve = ms_total = 0.0
        with open(f"{prefix}_log_1.txt", "w") as flog, \
                open(f"{prefix}_log_nonlinear_1.txt", "w") as fnl:
            flog.write("Solver control flop.\n")
            fnl.write("Non-linear control flop.\n")
            for op in inp.program:
                if op[0] == "dt":
                    dt = op[1]
                elif op[0] == "pnup":
                    if not inp.pnodes:
                        raise ValueError("No node to print for PNUP!")
                    u = s.state()["u"]
                    for n in inp.pnodes:
                        mode = "w" if new_file else "a"
                        with open(f"{prefix}_up_node_{n}.txt", mode) as f:
                            if new_file:
                                f.write(printnode_header(n) + "\n")
                            f.write(printnode_line(istep, t, u[n - 1]) + "\n")
                        hist[n].append([istep, t, *u[n - 1]])
                    new_file = False
                elif op[0] == "pnsf":
                    if not inp.pnodes:
                        raise ValueError("No node to print for PNSF!")
                    tx, txe, txb, flux = s.tform()
                    poro = s.state()["porosity"][:, None]
                    mode = "w" if new_sf else "a"
                    for (stem, name), arr in zip(PNSF_FILES, (tx, txe, txb, flux, poro)):
                        for n in inp.pnodes:
                            with open(f"{prefix}_{stem}_node_{n}.txt", mode) as f:
                                if new_sf:
                                    f.write(printnode_header(n, name) + "\n")
                                f.write(printnode_line(istep, t, arr[n - 1]) + "\n")
                            shist[stem][n].append([istep, t, *arr[n - 1]])
                    new_sf = False
                elif op[0] == "pres":
                    fl = inp.print_flags
                    st = s.state()
                    if any(fl[4:8]):
                        tx, txe, txb, flux = s.tform()
                    else:
                        tx = txe = txb = np.zeros((s.nnode, 6))
                        flux = np.zeros((s.nnode, 3))
                    if ix20 is None:
                        ix20 = s.numbering()[0]
                    vtk.write_results(vtk.results_name(prefix, istep, inp.vtk_legacy), ix20,
                                      deck.x, st["u"], st["dp"], st["porosity"], tx, txb, txe,
                                      flux, istep, t, flags=fl,
                                      quad=fl[0] and deck.nen == 20,  # maxno 8: vertices only
                                      legacy=inp.vtk_legacy,
                                      binary=inp.vtk_binary, nnodev=deck.nnodev)
                elif op[0] == "solv":
                    istep += 1
                    t = t + dt
                    fnl.write(f"step{istep:9d}{fortran_e(t, 13, 6)}\n")
                    info = s.step(cfg, dt)
                    ms_elem += info["ms_element"]
                    ms_solve += info["ms_solve"]
                    ms_total += info["ms_total"]
                    for k in range(min(info["nsolve"], 16)):
                        name = SOLVER_NAMES[info["used"][k]] if info["used"][k] in SOLVER_NAMES \
                            else SOLVER_NAMES[SOLVERS[cfg.solver]]
                        flog.write(f"{name.upper()}:  it{info['iters'][k]:10d} true rel resid "
                                   f"{fortran_e(info['true_rel'][k], 17, 10)} tol "
                                   f"{fortran_e(cfg.tol, 17, 10)}\n")
                    fnl.write(f"nlit{info['nlit']:7d}\n")
                    stats.append(dict(nlit=info["nlit"], iters=info["iters"]))
                    if echo:
                        echo(f"step {istep:6d} t {t:.6e} nlit {info['nlit']} "
                             f"iters {info['iters']}")
        m = s.matrix_info()
        with open(f"{prefix}_t_1.txt", "w") as f:
            f.write("# Arquivo de log do poro mecanico (MI355X)\n")
            f.write("Tempos (seg):\n")
            for key, v in (("ELMT", ms_elem), ("SOLVR", ms_solve), ("STEPS", ms_total)):
                f.write(f"{key:<12}{v / 1e3:12.6f}\n")
            f.write(f"{'TOTAL':<12}{time.perf_counter() - wall0:12.6f}\n")
            f.write("Malha e sistema linear:\n")
            for key, v in (("neq", s.neq), ("nad", s.nad), ("nnz", m["nnz"]),
                           ("nnode", s.nnode), ("numel", deck.numel)):
                f.write(f"{key:<10}{v:12d}\n")
        return dict(hist={n: np.array(v) for n, v in hist.items()}, stats=stats,
                    stress={k: {n: np.array(v) for n, v in d.items()} for k, d in shist.items()},
                    nnode=s.nnode, neq=s.neq, nad=s.nad)
    finally:
        s.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("deck")
    ap.add_argument("prefix")
    ap.add_argument("--fp32ir", action="store_true",
                    help="FP32 matrix values inside FP64 iterative refinement")
    ap.add_argument("--ordered", action="store_true",
                    help="reference summation order (bitwise against the oracle; slow)")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    inp = read_deck(a.deck)
    if inp.skipped and not a.quiet:
        print("macros without output here:", " ".join(sorted(set(inp.skipped))))
    r = run(inp, a.prefix, precision="fp32ir" if a.fp32ir else "fp64", ordered=a.ordered,
            device=a.device, echo=None if a.quiet else print)
    print(f"done: {len(r['stats'])} steps, nnode {r['nnode']} neq {r['neq']} nad {r['nad']}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
