"""Rounding envelope of the oracle: how far its own Krylov iteration counts
move when its input moves by one unit in the last place.

On the Biot saddle-point systems the iteration count of a Krylov solve is
chaotic at the rounding level: the oracle's count changes by up to ~8 % when
the matrix values, or the material constants they are built from, change by
one ulp.  A GPU path that sums in another order (tree reductions, colour-order
assembly) cannot reproduce one oracle run's count to +-2; its counts are held
instead to the envelope of a dozen oracle runs on 1-ulp perturbed inputs,
widened by 2 iterations plus 2 % (the sampled envelope underestimates the
spread).  The reference-order mode is held to exact counts and bitwise
results.
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np

from oracle import oracle as O

ULP = 2.0 ** -52


def _perturb(a, rng):
    return a * (1.0 + ULP * rng.choice([-1.0, 0.0, 1.0], size=np.shape(a)))


def oracle_solve(st, ad, al, solver, precond, b, tol, maxit, ngram=150, nrestart=0):
    neq = st["neq"]
    mvec = np.zeros(neq)
    O.lib.orc_precond(precond, neq, ad, mvec)
    x = np.zeros(neq)
    res = C.c_double()
    args = [neq, st["ia"], st["ja"], ad, al, mvec, b, x]
    if solver == "sqmr":
        fail = C.c_int()
        it = O.lib.orc_rpsqmr(*args, tol, maxit, C.byref(res), C.byref(fail))
    elif solver == "minres":
        it = O.lib.orc_pminres(*args, tol, maxit, nrestart, C.byref(res))
    elif solver == "gmres":
        it = O.lib.orc_gmres2(*args, ngram, tol, maxit, C.byref(res))
    elif solver == "pcg":
        it = O.lib.orc_pcg(*args, tol, maxit, C.byref(res))
    else:
        it = O.lib.orc_pbicgstab(*args, tol, maxit, C.byref(res))
    return it, x


def solve_envelope(st, m, solver, precond, b, tol, maxit, runs=12, seed=0):
    """(iterations, x) of the unperturbed oracle solve, the (min, max) of the
    iteration counts over `runs` solves with the matrix values perturbed by
    up to one ulp, and the largest relative change of x among them."""
    rng = np.random.default_rng(seed)
    it0, x0 = oracle_solve(st, m.ad, m.al, solver, precond, b, tol, maxit)
    its, dx = [it0], 0.0
    for _ in range(runs):
        it, x = oracle_solve(st, _perturb(m.ad, rng), _perturb(m.al, rng), solver, precond, b,
                             tol, maxit)
        its.append(it)
        dx = max(dx, np.abs(x - x0).max() / np.abs(x0).max())
    return it0, x0, (min(its), max(its)), dx


def march_envelope(deck, nsteps, runs=12, seed=0, **oracle_kw):
    """Per step, the (min, max) iteration count of each solve over the oracle
    march and `runs` marches whose material constants E and k are perturbed
    by up to one ulp.  Steps whose solve counts differ between runs keep the
    unperturbed run's count for the extra solves."""
    rng = np.random.default_rng(seed)
    decks = [deck]
    for _ in range(runs):
        d = copy.deepcopy(deck)
        d.e = d.e.copy()
        d.e[:, 0] = _perturb(d.e[:, 0], rng)
        d.e[:, 2] = _perturb(d.e[:, 2], rng)
        decks.append(d)
    hist = []
    for d in decks:
        m = O.OracleModel(d, **oracle_kw)
        hist.append([m.step()["iters"] for _ in range(nsteps)])
    env = []
    for k in range(nsteps):
        base = hist[0][k]
        lo, hi = list(base), list(base)
        for h in hist[1:]:
            for q, it in enumerate(h[k][: len(base)]):
                lo[q] = min(lo[q], it)
                hi[q] = max(hi[q], it)
        env.append(list(zip(lo, hi)))
    return env


def within(it, lo_hi, slack=2, rel=0.02):
    """`it` inside the envelope widened by `slack` + `rel` of its top: a
    dozen perturbed runs sample the spread, they do not bound it."""
    lo, hi = lo_hi
    s = slack + rel * hi
    return lo - s <= it <= hi + s
