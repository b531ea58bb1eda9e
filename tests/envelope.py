"""Rounding sensitivity of the oracle's Krylov iteration counts.

On the Biot saddle-point systems the iteration count of a Krylov solve is
chaotic at the rounding level.  With the oracle alone: perturbing the matrix
values, or E, k and the vertex coordinates, by one ulp moves the count by up
to ~8 % (column: 82-97 around 92; 4^3 heterogeneous box: 683-738 around 715,
820-871 around 845).  The fast GPU path sums in another order -- tree
reductions, which are also more accurate than the reference's sequential
sums and so tend to converge a little sooner -- and cannot reproduce one
oracle run's count to +-2.  Its counts are held to the oracle's within
2 iterations + 10 %; exact counts and bitwise results are the bar of the
reference-order mode (tests/test_gpu_ordered.py).
"""
from __future__ import annotations

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


def within(it, lo_hi, slack=2, rel=0.10):
    """`it` inside the sampled envelope widened by `slack` + `rel` of its
    top: a dozen perturbed runs sample the spread, they do not bound it."""
    lo, hi = lo_hi
    s = slack + rel * hi
    return lo - s <= it <= hi + s


def close_count(it_gpu, it_oracle, slack=2, rel=0.10):
    """A fast-path iteration count against one oracle run's."""
    return abs(it_gpu - it_oracle) <= slack + rel * it_oracle
