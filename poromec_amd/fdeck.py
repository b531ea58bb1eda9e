"""Writes a deck in the stream layout read by fortran/pm_mefdrive.f90."""
from __future__ import annotations

import numpy as np

from .problem import Deck
from .solver import SOLVERS, SolverConfig


def write_fdeck(path, deck: Deck, cfg: SolverConfig, nsteps: int, pnode: int = 1,
                reorder: bool = True):
    i32 = lambda *v: np.asarray(v, dtype="<i4").tobytes()  # noqa: E731
    f64 = lambda *v: np.asarray(v, dtype="<f8").tobytes()  # noqa: E731
    with open(path, "wb") as f:
        f.write(i32(deck.numel, deck.nnodev, deck.e.shape[0], nsteps, pnode, int(reorder)))
        f.write(f64(deck.dt, *deck.gravity, deck.gravity_mod))
        f.write(i32(SOLVERS[cfg.solver], cfg.precond, cfg.maxit, cfg.ngram,
                    0 if cfg.precision == "fp64" else 1, cfg.maxnlit, cfg.stop_crit))
        f.write(f64(cfg.tol, cfg.nltol))
        for a, dt in ((deck.ix, "<i4"), (deck.idv, "<i4"), (deck.x, "<f8"), (deck.e, "<f8"),
                      (deck.eload, "<i4"), (deck.eloadp, "<i4"), (deck.load, "<i4"),
                      (deck.fload, "<f8")):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def read_state(path, nnode):
    raw = np.fromfile(path, dtype="<f8")
    u = raw[: 4 * nnode].reshape(nnode, 4)
    dp = raw[4 * nnode: 5 * nnode]
    return u, dp
