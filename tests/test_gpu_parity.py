"""GPU parity of the HIP path (through the C ABI) against the oracle.

Bars (from the task statement of this path):
* CSRC sparsity bit-exact (ia, ja);
* matrix values within 1e-13 relative (|a_gpu - a_orc| <= 1e-13 * max|a| of
  the row: the GPU contracts FMAs and sums element contributions in colour
  order, so entries that cancel to ~0 are compared on the row's scale);
* right-hand sides within 1e-13 of max|b|;
* Krylov iteration counts: these tests run the fast path (tree reductions,
  colour-order assembly), whose counts are held to the oracle's within
  2 + 10 % (tests/envelope.py: the oracle's own counts move by up to ~8 %
  under 1-ulp input perturbations); solutions within ten times the spread of
  1-ulp perturbed oracle solutions.  Exact counts and bitwise solutions and
  histories are the bar of the reference-order mode (test_gpu_ordered.py);
* displacement and pressure histories within 1e-9 relative at every step.
"""
import numpy as np
import pytest

import envelope as E
from oracle import oracle as O
from poromec_amd import problem as P

pytestmark = pytest.mark.gpu


def _solver(deck, **kw):
    from poromec_amd.solver import MefSolver
    return MefSolver(deck, **kw)


def _rowwise_close(ia, ad_g, al_g, ad_o, al_o, rtol=1e-13):
    neq = len(ad_o)
    worst = 0.0
    for r in range(neq):
        s, e = ia[r] - 1, ia[r + 1] - 1
        scale = max(abs(ad_o[r]), np.abs(al_o[s:e]).max() if e > s else 0.0)
        err = max(abs(ad_g[r] - ad_o[r]), np.abs(al_g[s:e] - al_o[s:e]).max() if e > s else 0.0)
        worst = max(worst, err / scale if scale > 0 else err)
    return worst


@pytest.fixture(scope="module")
def column():
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    return d, st


def test_csrc_pattern_and_values(column):
    d, st = column
    s = _solver(d)
    s.pform(2, True, True, d.dt, d.dt)
    ia, ja, ad, al = s.export_csrc()
    assert np.array_equal(ia, st["ia"])
    assert np.array_equal(ja, st["ja"])
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=True)
    assert _rowwise_close(st["ia"], ad, al, m.ad, m.al[: st["nad"]]) < 1e-13


def test_internal_force_and_rhs(column):
    d, st = column
    rng = np.random.default_rng(11)
    s = _solver(d)
    m = O.OracleModel(d, st)
    u = rng.standard_normal((st["nnode"], 4)) * 1e-3
    dp = rng.standard_normal(st["nnode"]) * 1e-4
    u[d.nnodev:, 3] = 0.0
    dp[d.nnodev:] = 0.0
    m.u[:] = u
    m.dp[:] = dp
    s.set_state(u=u, u0=np.zeros_like(u), dp=dp, t=0.0)
    for isw in (2, 4):
        m.b[:] = 0.0
        m.env.t = 0.5
        m.assemble(isw=isw, lhs=(isw == 2), rhs=True)
        s.set_vectors(b=np.zeros(st["neq"]))
        s.pform(isw, isw == 2, True, d.dt, 0.5, target="b")
        _, b, _ = s.vectors()
        scale = np.abs(m.b).max()
        assert np.abs(b - m.b).max() <= 1e-13 * scale, isw


def test_spmv(column):
    d, st = column
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    x = np.random.default_rng(2).standard_normal(st["neq"])
    y_o = np.zeros(st["neq"])
    O.lib.orc_matvec_csrc(st["neq"], st["ia"], st["ja"], m.ad, m.al, x, y_o)
    y = s.spmv(x)
    assert np.abs(y - y_o).max() <= 1e-13 * np.abs(y_o).max()


@pytest.fixture(scope="module")
def column_spd():
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    d.idv[:, 3] = 1  # every pressure prescribed: K = K_uu, positive definite
    st = O.build_structure(d)
    return d, st


@pytest.mark.parametrize("system,solver,precond", [
    ("column", "sqmr", 2), ("column", "pcg", 2), ("column", "gmres", 2),
    ("column_spd", "bicgstab", 2), ("column_spd", "minres", 5), ("column_spd", "pcg", 2)])
def test_solvers_vs_oracle(request, system, solver, precond):
    from poromec_amd.solver import SolverConfig
    d, st = request.getfixturevalue(system)
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    b = np.random.default_rng(4).standard_normal(st["neq"])
    maxit = 200 if solver == "gmres" else 20000
    it_o, x_o, env, dx = E.solve_envelope(st, m, solver, precond, b, 1e-10, maxit)
    assert it_o > 0, f"oracle {solver} did not converge ({it_o})"
    x, info = s.solve(b, SolverConfig(solver=solver, precond=precond, tol=1e-10, maxit=maxit))
    assert E.within(info["iters"], env), (info["iters"], it_o, env)
    err = np.abs(x - x_o).max() / np.abs(x_o).max()
    assert err <= max(10 * dx, 1e-12), (err, dx)


def test_fp32ir_reaches_tolerance(column):
    from poromec_amd.solver import SolverConfig
    d, st = column
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    b = np.random.default_rng(5).standard_normal(st["neq"])
    x64, i64 = s.solve(b, SolverConfig(tol=1e-10))
    x32, i32 = s.solve(b, SolverConfig(tol=1e-10, precision="fp32ir"))
    assert i32["true_rel"] < 1e-10
    assert i32["spmv32"] > 0
    assert np.abs(x32 - x64).max() <= 1e-8 * np.abs(x64).max()


def test_fp32ir_every_spmv_shape():
    """Every SpMV shape, the index-prefetch ones included, on a box whose
    interior node blocks are wider than one pass of the narrower shapes
    (270 columns): FP32+IR reaches the FP64 solution, and the shapes' inner
    iteration counts agree up to the drift that their different summation
    orders cause (an operator that is not the FP32 matrix, or not the same
    one every iteration, shows as a far larger count)."""
    from poromec_amd.solver import SPMV_SHAPES, SolverConfig
    d = P.heterogeneous_box_deck(4, seed=2, sigma_lnk=0.5, corr_len=2.0)
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    assert s.matrix_info()["ncols"] > 0
    b = np.random.default_rng(6).standard_normal(s.neq)
    x64, _ = s.solve(b, SolverConfig(tol=1e-10))
    iters = []
    for shape in range(SPMV_SHAPES):
        s.set_spmv_shape(True, shape)
        x32, i32 = s.solve(b, SolverConfig(tol=1e-10, precision="fp32ir"))
        assert i32["true_rel"] < 1e-10, shape
        assert np.abs(x32 - x64).max() <= 1e-8 * np.abs(x64).max(), shape
        iters.append(i32["iters"])
    for shape, it in enumerate(iters):
        assert abs(it - iters[0]) <= 0.1 * iters[0] + 5, (shape, iters)


def test_fp64_spmv_every_shape():
    """Every FP64 SpMV shape, the index-prefetch ones included, against the
    exported CSRC matrix applied on the host (A = diag + L + L^T), on a box
    whose interior node blocks need two passes of the (64, 3) shape."""
    import scipy.sparse as sp
    from poromec_amd.solver import SPMV_SHAPES
    d = P.heterogeneous_box_deck(4, seed=3, sigma_lnk=0.5, corr_len=2.0)
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    ia, ja, ad, al = s.export_csrc()
    n = s.neq
    rows = np.repeat(np.arange(n), np.diff(ia))
    L = sp.csr_matrix((al, (rows, ja - 1)), shape=(n, n))
    x = np.random.default_rng(8).standard_normal(n)
    y0 = ad * x + L @ x + L.T @ x
    for shape in range(SPMV_SHAPES):
        s.set_spmv_shape(False, shape)
        y = s.spmv(x)
        assert np.abs(y - y0).max() <= 1e-13 * np.abs(y0).max(), shape


def test_csrc_dropin_solve(column):
    from poromec_amd.solver import CsrcSolver, SolverConfig
    d, st = column
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    cs = CsrcSolver(st["ia"], st["ja"])
    cs.set_values(m.ad, m.al[: st["nad"]])
    x = np.random.default_rng(9).standard_normal(st["neq"])
    y_o = np.zeros(st["neq"])
    O.lib.orc_matvec_csrc(st["neq"], st["ia"], st["ja"], m.ad, m.al, x, y_o)
    assert np.abs(cs.spmv(x) - y_o).max() <= 1e-14 * np.abs(y_o).max()
    xs, info = cs.solve(y_o, SolverConfig(tol=1e-10))
    assert info["true_rel"] < 1e-9


def _history_errors(d, nsteps, cfg_kw=None, oracle_kw=None):
    from poromec_amd.solver import SolverConfig
    st = O.build_structure(d)
    m = O.OracleModel(d, st, **(oracle_kw or {}))
    s = _solver(d)
    cfg = SolverConfig(**(cfg_kw or {}))
    worst_u = worst_p = 0.0
    its = []
    for k in range(nsteps):
        io = m.step()
        ig = s.step(cfg)
        g = s.state()
        nv = d.nnodev
        eu = np.abs(g["u"][:, :3] - m.u[:, :3]).max() / max(np.abs(m.u[:, :3]).max(), 1e-300)
        ep = np.abs(g["u"][:nv, 3] - m.u[:nv, 3]).max() / max(np.abs(m.u[:nv, 3]).max(), 1e-300)
        worst_u, worst_p = max(worst_u, eu), max(worst_p, ep)
        its.append((io["iters"], ig["iters"]))
        assert io["nlit"] == ig["nlit"], (k, io, ig)
        for a, b in zip(io["iters"], ig["iters"]):
            assert E.close_count(b, a), (k, io["iters"], ig["iters"])
        dpe = np.abs(g["dp"][:nv] - m.dp[:nv]).max() / max(np.abs(m.dp[:nv]).max(), 1e-300)
        assert dpe < 1e-9, (k, dpe)
    return worst_u, worst_p, its


def test_time_march_histories_column():
    d = P.column_deck(2, 2, 8, dt=0.01, nsteps=8)
    eu, ep, its = _history_errors(d, 8)
    assert eu < 1e-9 and ep < 1e-9, (eu, ep, its)


def test_time_march_histories_heterogeneous():
    d = P.heterogeneous_box_deck(4, seed=7, dt=0.1, nsteps=4, corr_len=2.0)
    eu, ep, its = _history_errors(d, 4)
    assert eu < 1e-9 and ep < 1e-9, (eu, ep, its)


def test_sqmr_gmres_fallback_fast_mode(column):
    """The SQMR -> GMRES fallback on the assembled device matrix in the
    default (fast) mode: GMRES finishes the solve and reaches the tolerance
    the oracle's fallback reaches."""
    from poromec_amd.solver import SolverConfig
    d, st = column
    s = _solver(d)
    s.pform(2, True, False, d.dt, d.dt)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    b = np.random.default_rng(6).standard_normal(st["neq"])
    it_o, x_o, used_o = O.solv(st, m.ad, m.al[: st["nad"]], b, solver=11, tol=1e-10, maxit=5)
    assert used_o == 2 and it_o > 0
    x, info = s.solve(b, SolverConfig(solver="sqmr", tol=1e-10, maxit=5))
    assert info["used"] == 2
    assert abs((info["iters"] - 5) - it_o) <= 2 + 0.1 * it_o, (info["iters"], it_o)
    # GMRES stops on the left-Jacobi-preconditioned residual (Iters.f:1411-1775)
    minv = 1.0 / m.ad
    r = minv * (b - s.spmv(x))
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(minv * b)
    assert np.abs(x - x_o).max() <= 1e-6 * np.abs(x_o).max()
