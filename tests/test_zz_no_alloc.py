"""pm_solve_noalloc: the same solve as pm_solve with its scratch in the
library's static device array (no hipMalloc in the call).  Its results must
be pm_solve's, bit for bit (same kernels, same summation order), on two
calls in a row on one GPU, the second on a smaller system (the array keeps
the first call's contents), through both set-up paths (the device-assembled
matrix and the caller's CSRC); and it must refuse what would need more
memory than the array holds."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import _lib as L
from poromec_amd import problem as P

pytestmark = pytest.mark.gpu

MAX_NEQ = 524288  # PM_NOALLOC_MAX_NEQ


def _noalloc(obj, b, cfg):
    info = L.SolveInfo()
    x = np.zeros(obj.neq)
    rc = obj.lib.pm_solve_noalloc(obj.ctx, C.byref(cfg.c()), L.ptr(L.as_f64(b)), L.ptr(x),
                                  C.byref(info))
    return rc, x, info


def test_noalloc_equals_pm_solve_twice_larger_then_smaller():
    from poromec_amd.solver import MefSolver, SolverConfig
    rng = np.random.default_rng(21)
    decks = [P.heterogeneous_box_deck(5, seed=2, sigma_lnk=1.0, corr_len=2.0, dt=0.1),
             P.column_deck(2, 2, 4, dt=0.01)]
    sizes = []
    for d in decks:
        s = MefSolver(d)
        s.pform(2, True, False, d.dt, d.dt)
        sizes.append(s.neq)
        b = rng.standard_normal(s.neq)
        for cfg in (SolverConfig(tol=1e-10), SolverConfig(solver="pcg", tol=1e-10),
                    SolverConfig(tol=1e-10, precision="fp32ir")):
            x_ref, i_ref = s.solve(b, cfg)
            rc, x, info = _noalloc(s, b, cfg)
            assert rc == 0, (cfg, L.PM_ERR.get(rc))
            assert info.iters == i_ref["iters"] and info.used_solver == i_ref["used"]
            assert np.array_equal(x, x_ref), (cfg, np.abs(x - x_ref).max())
        s.close()
    assert sizes[1] < sizes[0]


def test_noalloc_csrc_dropin():
    from poromec_amd.solver import CsrcSolver, SolverConfig
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    cs = CsrcSolver(st["ia"], st["ja"])
    cs.set_values(m.ad, m.al[: st["nad"]])
    b = np.random.default_rng(3).standard_normal(st["neq"])
    cfg = SolverConfig(tol=1e-10)
    x_ref, i_ref = cs.solve(b, cfg)
    rc, x, info = _noalloc(cs, b, cfg)
    assert rc == 0 and info.iters == i_ref["iters"]
    assert np.array_equal(x, x_ref)
    cs.close()


def test_noalloc_refusals():
    from poromec_amd.solver import CsrcSolver, MefSolver, SolverConfig
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    s = MefSolver(d)
    s.pform(2, True, False, d.dt, d.dt)
    b = np.random.default_rng(4).standard_normal(s.neq)
    # FP32 values do not exist before the first FP32 use
    rc, _, _ = _noalloc(s, b, SolverConfig(tol=1e-10, precision="fp32ir"))
    assert rc == -8
    # GMRES needs its basis
    rc, _, _ = _noalloc(s, b, SolverConfig(solver="gmres", tol=1e-10))
    assert rc == -9
    # SQMR at maxit: pm_solve falls back to GMRES, the no-alloc solve stops
    _, i_ref = s.solve(b, SolverConfig(tol=1e-10, maxit=5))
    assert i_ref["used"] == 2
    rc, _, _ = _noalloc(s, b, SolverConfig(tol=1e-10, maxit=5))
    assert rc == -4
    s.close()
    # a system larger than the array: a diagonal one, no stored off-diagonal
    n = MAX_NEQ + 1
    ia = np.ones(n + 1, np.int64)
    cs = CsrcSolver(ia, np.zeros(1, np.int32))
    cs.set_values(np.full(n, 2.0), np.zeros(1))
    rc, _, _ = _noalloc(cs, np.ones(n), SolverConfig(tol=1e-10))
    assert rc == -2
    cs.close()
