"""Reference summation order on the GPU: bit-for-bit solver parity.

Krylov iteration counts on these saddle-point systems are chaotic at the
rounding level: the oracle's own count moves by up to +-25 when the matrix
values are perturbed by one unit in the last place.  With pm_set_ordered the
library sums every dot product sequentially and every SpMV row in the order
of matvec_csrc_sym_pm, and nothing is contracted into fused multiply-adds, so
given the reference's own CSRC matrix (the solv_pm drop-in,
pm_struct_upload + pm_values_upload) each solver must return exactly the
oracle's iteration count and the oracle's solution bit for bit."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

pytestmark = pytest.mark.gpu


def _system(spd: bool):
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    if spd:  # every pressure prescribed: K = K_uu, symmetric positive definite
        d.idv[:, 3] = 1
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    return st, m


@pytest.fixture(scope="module")
def saddle():
    return _system(False)


@pytest.fixture(scope="module")
def spd():
    return _system(True)


def _csrc(st, m):
    from poromec_amd.solver import CsrcSolver
    cs = CsrcSolver(st["ia"], st["ja"])
    cs.set_values(m.ad, m.al[: st["nad"]])
    cs.set_ordered(True)
    return cs


def _oracle_solve(st, m, solver, precond, b, tol, maxit, ngram=150):
    neq = st["neq"]
    mvec = np.zeros(neq)
    O.lib.orc_precond(precond, neq, m.ad, mvec)
    x = np.zeros(neq)
    res = C.c_double()
    args = [neq, st["ia"], st["ja"], m.ad, m.al, mvec, b, x]
    if solver == "sqmr":
        fail = C.c_int()
        it = O.lib.orc_rpsqmr(*args, tol, maxit, C.byref(res), C.byref(fail))
    elif solver == "minres":
        it = O.lib.orc_pminres(*args, tol, maxit, 0, C.byref(res))
    elif solver == "gmres":
        it = O.lib.orc_gmres2(*args, ngram, tol, maxit, C.byref(res))
    elif solver == "pcg":
        it = O.lib.orc_pcg(*args, tol, maxit, C.byref(res))
    else:
        it = O.lib.orc_pbicgstab(*args, tol, maxit, C.byref(res))
    return it, x


def test_ordered_spmv_bitwise(saddle):
    st, m = saddle
    cs = _csrc(st, m)
    x = np.random.default_rng(2).standard_normal(st["neq"])
    y_o = np.zeros(st["neq"])
    O.lib.orc_matvec_csrc(st["neq"], st["ia"], st["ja"], m.ad, m.al, x, y_o)
    assert np.array_equal(cs.spmv(x), y_o)
    cs.close()


@pytest.mark.parametrize("system,solver,precond,tol", [
    ("saddle", "sqmr", 2, 1e-10), ("saddle", "pcg", 2, 1e-10), ("saddle", "gmres", 2, 1e-10),
    ("spd", "pcg", 2, 1e-10), ("spd", "minres", 5, 1e-10), ("spd", "bicgstab", 2, 1e-10),
    ("spd", "gmres", 2, 1e-10)])
def test_ordered_solvers_bitwise(request, system, solver, precond, tol):
    from poromec_amd.solver import SolverConfig
    st, m = request.getfixturevalue(system)
    b = np.random.default_rng(4).standard_normal(st["neq"])
    maxit = 200 if solver == "gmres" else 20000
    it_o, x_o = _oracle_solve(st, m, solver, precond, b, tol, maxit)
    assert it_o > 0, f"oracle {solver} did not converge ({it_o})"
    cs = _csrc(st, m)
    x, info = cs.solve(b, SolverConfig(solver=solver, precond=precond, tol=tol, maxit=maxit))
    cs.close()
    assert info["iters"] == it_o, (info["iters"], it_o)
    assert np.array_equal(x, x_o), np.abs(x - x_o).max() / np.abs(x_o).max()


def _gravity_box():
    d = P.heterogeneous_box_deck(3, seed=3, sigma_lnk=0.8, corr_len=2.0, dt=0.05, nsteps=4)
    d.gravity = np.array([0.0, 0.0, -1.0])
    d.e[:, 5] = 2000.0
    d.e[:, 6] = 1000.0
    return d


def _face_load_box():
    """Gravity box with all three face-load types: a type-41 normal pressure
    and a type-43 hydraulic flux on the top faces, a type-40 traction on the
    faces of one side."""
    d = _gravity_box()
    load, fload = P.empty_load_table()
    P.set_traction_load(load, fload, 1, (1e-4, -2e-4, -1e-3))
    load[1] = (41, 2, 2)
    fload[1, 0, :2] = (0.0, 1e30)
    fload[1, 1, :2] = (5e-4, 5e-4)
    load[2] = (43, 2, 2)
    fload[2, 0, :2] = (0.0, 1e30)
    fload[2, 1, :2] = (1e-6, 1e-6)
    d.load, d.fload = load, fload
    ne = d.numel
    top = np.arange(ne) >= d.nx * d.ny * (d.nz - 1)
    side = (np.arange(ne) % d.nx) == 0
    d.eload = np.zeros((ne, 7), np.int32)
    d.eload[top, 0] = 2
    d.eload[side, 4] = 1
    d.eloadp = np.zeros((ne, 7), np.int32)
    d.eloadp[top, 0] = 3
    return d


def _deck_of(which):
    if which == "column":
        return P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    return _gravity_box() if which == "gravity_box" else _face_load_box()


@pytest.mark.parametrize("which", ["column", "gravity_box", "face_load_box"])
def test_ordered_assembly_bitwise(which):
    """isw 2 matrix and right-hand side, and isw 4 residual (face loads,
    gravity), summed in element order: equal to the reference's bit for bit."""
    from poromec_amd.solver import MefSolver
    d = _deck_of(which)
    st = O.build_structure(d)
    s = MefSolver(d)
    s.set_ordered(True)
    m = O.OracleModel(d, st)
    rng = np.random.default_rng(11)
    u = rng.standard_normal((st["nnode"], 4)) * 1e-3
    dp = rng.standard_normal(st["nnode"]) * 1e-4
    u[d.nnodev:, 3] = 0.0
    dp[d.nnodev:] = 0.0
    m.u[:] = u
    m.dp[:] = dp
    s.set_state(u=u, u0=np.zeros_like(u), dp=dp, t=0.0)
    for isw in (2, 4):
        b0 = rng.standard_normal(st["neq"])
        m.b[:] = b0
        m.env.t = 0.5
        m.assemble(isw=isw, lhs=(isw == 2), rhs=True)
        s.set_vectors(b=b0)
        s.pform(isw, isw == 2, True, d.dt, 0.5, target="b")
        _, b, _ = s.vectors()
        assert np.array_equal(b, m.b), (isw, np.abs(b - m.b).max())
        if isw == 2:
            ia, ja, ad, al = s.export_csrc()
            assert np.array_equal(ad, m.ad), np.abs(ad - m.ad).max()
            assert np.array_equal(al, m.al[: st["nad"]]), np.abs(al - m.al[: st["nad"]]).max()
    s.close()


@pytest.mark.parametrize("which", ["column", "gravity_box", "face_load_box"])
def test_ordered_march_bitwise(which):
    """The whole backward-Euler march in the reference's order: the oracle's
    nonlinear and Krylov iteration counts exactly, and u, dp and porosity
    bit for bit after every step."""
    from poromec_amd.solver import MefSolver, SolverConfig
    d = P.column_deck(2, 2, 8, dt=0.01, nsteps=6) if which == "column" else _deck_of(which)
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    s = MefSolver(d)
    s.set_ordered(True)
    for k in range(4):
        io = m.step()
        ig = s.step(SolverConfig())
        assert (io["nlit"], list(io["iters"])) == (ig["nlit"], ig["iters"]), (k, io, ig)
        g = s.state()
        assert np.array_equal(g["u"], m.u), (k, np.abs(g["u"] - m.u).max())
        assert np.array_equal(g["dp"], m.dp), k
        assert np.array_equal(g["porosity"], m.porosity), k
    s.close()


@pytest.mark.parametrize("which", ["column", "gravity_box"])
def test_tform_bitwise(which):
    """Vertex stresses and Darcy flux (the pnsf quantities) of a random
    state: the element-order gather sums as tform_pm does, so the kernel's
    results are the oracle's bit for bit, in the default mode too."""
    from poromec_amd.solver import MefSolver
    d = _deck_of(which)
    st = O.build_structure(d)
    s = MefSolver(d)
    m = O.OracleModel(d, st)
    rng = np.random.default_rng(5)
    u = rng.standard_normal((st["nnode"], 4)) * 1e-3
    dp = rng.standard_normal(st["nnode"]) * 1e-4
    u[d.nnodev:, 3] = 0.0
    dp[d.nnodev:] = 0.0
    m.u[:] = u
    m.dp[:] = dp
    s.set_state(u=u, u0=np.zeros_like(u), dp=dp, t=0.0)
    for name, g, o in zip(("tx", "txe", "txb", "flux"), s.tform(), m.tform()):
        assert np.array_equal(g, o), (name, np.abs(g - o).max())
        assert np.abs(o[: d.nnodev]).max() > 0
    s.close()


def test_ordered_march_criterion2():
    """cal_residuo_pm criterion 2 (Residuo.f:84-136, the reference default
    with the resp0 = max(resp0, gresu) quirk of line 119) on the GPU in
    reference order: the load step converges at the oracle's nonlinear
    iteration with its Krylov counts and state bit for bit; the next step
    reaches maxnlit like the oracle (PM_ERR_NLIT for its -10, the state at
    that point equal bit for bit), and with maxnlit_accept both keep it."""
    from poromec_amd._lib import PmError
    from poromec_amd.solver import MefSolver, SolverConfig
    d = P.column_deck(2, 2, 8, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    for accept in (False, True):
        cfg = SolverConfig(stop_crit=2, maxnlit=4, maxnlit_accept=accept)
        m = O.OracleModel(d, st, stop_crit=2, maxnlit=4, maxnlit_accept=accept)
        s = MefSolver(d)
        s.set_ordered(True)
        io = m.step(check=False)
        ig = s.step(cfg)
        assert io["rc"] == 0 and io["nlit"] == 3, io
        assert (io["nlit"], list(io["iters"])) == (ig["nlit"], ig["iters"]), (io, ig)
        g = s.state()
        assert np.array_equal(g["u"], m.u) and np.array_equal(g["dp"], m.dp)
        io = m.step(check=False)
        assert io["nlit"] == 4
        if accept:
            assert io["rc"] == 0
            ig = s.step(cfg)
            assert (io["nlit"], list(io["iters"])) == (ig["nlit"], ig["iters"]), (io, ig)
        else:
            assert io["rc"] == -10
            with pytest.raises(PmError) as e:
                s.step(cfg)
            assert e.value.code == -7  # PM_ERR_NLIT
        g = s.state()
        assert np.array_equal(g["u"], m.u), np.abs(g["u"] - m.u).max()
        assert np.array_equal(g["dp"], m.dp)
        s.close()


def test_ordered_sqmr_gmres_fallback(saddle):
    """solv_pm's fallback (Solv.f:92-94,765; Iters.f:6343-6349): SQMR stopped
    at maxit hands the system to GMRES(ngram) from x = 0 under the same maxit.
    In reference order the GPU reports GMRES as the finishing solver, takes
    the oracle's GMRES iterations after its maxit SQMR ones, and returns the
    oracle's solution bit for bit."""
    from poromec_amd.solver import SolverConfig
    st, m = saddle
    b = np.random.default_rng(9).standard_normal(st["neq"])
    it_o, x_o, used_o = O.solv(st, m.ad, m.al[: st["nad"]], b, solver=11, tol=1e-10, maxit=5)
    assert used_o == 2 and it_o > 0
    cs = _csrc(st, m)
    x, info = cs.solve(b, SolverConfig(solver="sqmr", tol=1e-10, maxit=5))
    cs.close()
    assert info["used"] == 2
    assert info["iters"] == 5 + it_o, (info["iters"], it_o)
    assert np.array_equal(x, x_o), np.abs(x - x_o).max() / np.abs(x_o).max()
