"""The equal-order Q1/Q1 hexa8 element (element type 7) on the GPU, through
the C ABI, against the oracle's statement of it (oracle/pm_elmt07.c, pinned
without a reference by tests/test_oracle_elmt07.py).

Bars, as for the hexa20/8 path:
* CSRC sparsity bit-exact (ia, ja); matrix values within 1e-13 of the row's
  largest entry; isw 2 and isw 4 right-hand sides within 1e-13 of max|b|
  (distorted elements, gravity, tractions of types 40 and 41, a type-43
  flux);
* vertex stresses and Darcy flux within 1e-13 of the oracle's from the same
  state; porosity with the march (1e-9);
* the march: u and p within 1e-9 at every step, the same nonlinear
  iterations, Krylov counts within the oracle's rounding envelope
  (2 + 10 %); FP32+IR to the FP64 tolerance;
* the adjoint gradient against central differences (1e-6 relative);
* a batch of Q1 boxes against every box stepped alone.
"""
import dataclasses

import numpy as np
import pytest

import envelope as E
from oracle import oracle as O
from poromec_amd import problem as P

pytestmark = pytest.mark.gpu

RTOL_HIST = 1e-9


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def q1_deck(n=3, seed=0, loads=True):
    """A heterogeneous Q1/Q1 box with its interior vertices moved, gravity,
    the type-40 top traction on most top faces, a type-41 pressure on the
    others (x = 0 column) and a type-43 flux through the x = 0 wall (face 6
    of an element is its x-min side)."""
    d = P.heterogeneous_box_deck(n, seed=seed, sigma_lnk=0.8, corr_len=2.0, dt=0.05, nsteps=6,
                                 nen=8)
    rng = np.random.default_rng(seed + 1)
    x = d.x.copy()
    inner = np.all((x > 0) & (x < n), axis=1)
    x[inner] += 0.08 * rng.standard_normal((inner.sum(), 3))
    d = dataclasses.replace(d, x=x)
    if loads:
        d.gravity = np.array([0.0, 0.0, -10.0])
        d.e[:, 5] = 2000.0
        d.load[1] = (41, 2, 2)
        d.fload[1, 0, :2] = (0.0, 1e30)
        d.fload[1, 1, :2] = (-2e-4, -2e-4)
        d.load[2] = (43, 2, 2)
        d.fload[2, 0, :2] = (0.0, 1e30)
        d.fload[2, 1, :2] = (1e-3, 1e-3)
        xs = np.arange(d.numel) % n
        top = d.eload[:, 0] == 1
        d.eload[top & (xs == 0), 0] = 2
        d.eloadp[xs == 0, 5] = 3
    return d


def _rowwise(ia, ad_g, al_g, ad_o, al_o):
    worst = 0.0
    for r in range(len(ad_o)):
        s, e = ia[r] - 1, ia[r + 1] - 1
        scale = max(abs(ad_o[r]), np.abs(al_o[s:e]).max() if e > s else 0.0)
        err = max(abs(ad_g[r] - ad_o[r]), np.abs(al_g[s:e] - al_o[s:e]).max() if e > s else 0.0)
        worst = max(worst, err / scale if scale > 0 else err)
    return worst


def test_q1_structure_counts():
    """The configurations' nominal counts: 4 x 4 x 40 column = 4100 dofs
    before restraints; the oracle's CSRC pattern bit for bit."""
    from poromec_amd.solver import MefSolver
    d = P.column_deck(4, 4, 40, nen=8)
    assert P.box_counts(4, 4, 40, nen=8)["dofs"] == 4100
    st = O.build_structure(d)
    s = MefSolver(d)
    assert (s.nnode, s.neq, s.nad) == (d.nnodev, st["neq"], st["nad"])
    s.pform(2, True, False, d.dt, d.dt)
    ia, ja, _, _ = s.export_csrc()
    assert np.array_equal(ia, st["ia"]) and np.array_equal(ja, st["ja"])
    s.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_q1_matrix_and_rhs(seed):
    from poromec_amd.solver import MefSolver
    d = q1_deck(3, seed)
    st = O.build_structure(d)
    s = MefSolver(d)
    m = O.OracleModel(d, st)
    rng = np.random.default_rng(seed + 5)
    u = rng.standard_normal((st["nnode"], 4)) * 1e-3
    dp = rng.standard_normal(st["nnode"]) * 1e-4
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
        assert np.abs(b - m.b).max() <= 1e-13 * np.abs(m.b).max(), isw
        if isw == 2:
            ia, ja, ad, al = s.export_csrc()
            assert np.array_equal(ia, st["ia"]) and np.array_equal(ja, st["ja"])
            assert _rowwise(st["ia"], ad, al, m.ad, m.al[: st["nad"]]) < 1e-13
    s.close()


def test_q1_march_tform_porosity():
    """Six steps of the loaded, distorted box: histories, nonlinear and
    Krylov iterations, then stresses, flux and porosity."""
    from poromec_amd.solver import MefSolver, SolverConfig
    d = q1_deck(3, 0)
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    s = MefSolver(d)
    for k in range(6):
        io = m.step()
        ig = s.step(SolverConfig(tol=1e-10))
        assert io["nlit"] == ig["nlit"], (k, io, ig)
        for a, b in zip(io["iters"], ig["iters"]):
            assert E.close_count(b, a), (k, io["iters"], ig["iters"])
        g = s.state()
        assert _rel(g["u"][:, :3], m.u[:, :3]) <= RTOL_HIST, k
        assert _rel(g["u"][:, 3], m.u[:, 3]) <= RTOL_HIST, k
        assert _rel(g["porosity"], m.porosity) <= RTOL_HIST, k  # a function of the state
    tx, txe, txb, fl = s.tform()
    otx, otxe, otxb, ofl = m.tform()
    for a, b in ((tx, otx), (txe, otxe), (txb, otxb), (fl, ofl)):
        assert _rel(a, b) <= RTOL_HIST
    # the same state through both tform paths: 1e-13
    s.set_state(u=m.u, u0=m.u0, dp=m.dp, porosity=m.porosity, t=m.t)
    tx, txe, txb, fl = s.tform()
    for a, b in ((tx, otx), (txe, otxe), (txb, otxb), (fl, ofl)):
        assert _rel(a, b) <= 1e-13
    s.close()


def test_q1_stabilized_column():
    """The benchmark column (4 x 4 x 40, 4100 dofs, M / Kv = 74) with the
    pressure stabilization beta = 1/4: 20 steps against the oracle's march,
    and the undrained p0 to 1e-3 below 2 m after the first step.  SQMR runs
    with the |diag| Jacobi preconditioner (precond 5): with the signed
    diagonal (precond 2) its recurrence diverges on this system's load
    vector (the reference then switches to GMRES after maxit, Solv.f:765;
    random right-hand sides converge), while |diag| converges in ~95
    iterations."""
    from poromec_amd.solver import MefSolver, SolverConfig
    d = P.column_deck(4, 4, 40, dt=0.01, nsteps=20, nen=8)
    d.e[:, 8] = 0.25
    m = O.OracleModel(d, precond=5)
    s = MefSolver(d)
    E_, nu, M, q = 10.0, 0.3, 1000.0, 1.0e-3
    Kv = E_ * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    p0 = M * q / (Kv + M)
    z = d.x[:, 2]
    for k in range(20):
        io = m.step()
        ig = s.step(SolverConfig(tol=1e-10, precond=5))
        assert io["nlit"] == ig["nlit"], (k, io, ig)
        assert io["used"] == ig["used"][: len(io["used"])] == [11] * len(io["used"])
        for a, b in zip(io["iters"], ig["iters"]):
            assert E.close_count(b, a), (k, io["iters"], ig["iters"])
        u = s.state()["u"]
        assert _rel(u[:, :3], m.u[:, :3]) <= RTOL_HIST, k
        assert _rel(u[:, 3], m.u[:, 3]) <= RTOL_HIST, k
        if k == 0:
            assert np.abs(u[z <= z.max() - 2.0, 3] / p0 - 1.0).max() < 1e-3
    s.close()


def test_q1_every_spmv_shape():
    """Every SpMV shape on a Q1/Q1 box (interior node blocks of 4 rows x 108
    columns, 27 runs): FP64 against the exported CSRC matrix applied on the
    host, FP32+IR to the FP64 solution."""
    import scipy.sparse as sp
    from poromec_amd.solver import SPMV_SHAPES, MefSolver, SolverConfig
    d = q1_deck(6, 3, loads=False)
    s = MefSolver(d)
    s.pform(2, True, False, d.dt, d.dt)
    ia, ja, ad, al = s.export_csrc()
    n = s.neq
    rows = np.repeat(np.arange(n), np.diff(ia))
    Lm = sp.csr_matrix((al, (rows, ja - 1)), shape=(n, n))
    x = np.random.default_rng(8).standard_normal(n)
    y0 = ad * x + Lm @ x + Lm.T @ x
    b = np.random.default_rng(6).standard_normal(n)
    x64, _ = s.solve(b, SolverConfig(tol=1e-10))
    for shape in range(SPMV_SHAPES):
        s.set_spmv_shape(False, shape)
        y = s.spmv(x)
        assert np.abs(y - y0).max() <= 1e-13 * np.abs(y0).max(), shape
        s.set_spmv_shape(True, shape)
        x32, i32 = s.solve(b, SolverConfig(tol=1e-10, precision="fp32ir"))
        assert i32["true_rel"] < 1e-10, shape
        assert np.abs(x32 - x64).max() <= 1e-8 * np.abs(x64).max(), shape
    s.close()


def test_q1_fp32ir_step():
    from poromec_amd.solver import MefSolver, SolverConfig
    d = q1_deck(6, 2)
    s = MefSolver(d)
    i64 = s.step(SolverConfig(tol=1e-10))
    u64 = s.state()["u"].copy()
    z4, zn = np.zeros((s.nnode, 4)), np.zeros(s.nnode)
    s.set_state(u=z4, u0=z4, dp=zn, t=0.0)
    s.set_materials(d)
    i32 = s.step(SolverConfig(tol=1e-10, precision="fp32ir"))
    assert max(i32["true_rel"]) < 1e-10 and i32["nlit"] == i64["nlit"]
    u32 = s.state()["u"]
    assert _rel(u32[:, :3], u64[:, :3]) <= RTOL_HIST
    assert _rel(u32[:, 3], u64[:, 3]) <= RTOL_HIST
    s.close()


def test_q1_adjoint_matches_finite_differences():
    from poromec_amd.adjoint import MarchMisfit
    from poromec_amd.solver import MefSolver, SolverConfig
    nsteps = 5
    cfg = SolverConfig(tol=1e-13, maxit=200000)
    d = q1_deck(3, 4, loads=False)
    d = dataclasses.replace(d, nsteps=nsteps)
    probe = MefSolver(d)
    _, idm, _ = probe.numbering()
    rng = np.random.default_rng(9)
    pfree = np.nonzero(idm[:, 3] > 0)[0]
    ufree = np.nonzero(idm[:, 2] > 0)[0]
    sens = [(int(v), 3) for v in rng.choice(pfree, 5, replace=False)] + \
           [(int(v), 2) for v in rng.choice(ufree, 2, replace=False)]
    e = d.e.copy()
    d.e = e * np.where(np.arange(e.shape[1]) == 2, 1.5, 1.0)
    probe.set_materials(d)
    d.e = e
    obs = np.zeros((nsteps, len(sens)))
    for k in range(nsteps):
        probe.step(cfg)
        u = probe.state()["u"]
        obs[k] = [u[a, b] for a, b in sens]
    probe.close()
    mm = MarchMisfit(d, sens, obs, nsteps=nsteps, cfg=cfg, ckpt_every=2)
    lE, lk = (t.numpy() for t in mm.params())
    J, gE, gk = mm._evaluate(lE, lk)
    assert J > 0
    g = np.concatenate([gE, gk])
    h = 1e-5
    for _ in range(3):
        v = rng.standard_normal(2 * len(lE))
        vE, vk = v[: len(lE)], v[len(lE):]
        fd = (mm._evaluate(lE + h * vE, lk + h * vk)[0] -
              mm._evaluate(lE - h * vE, lk - h * vk)[0]) / (2 * h)
        ad = float(g @ v)
        assert abs(fd) > 0 and abs(fd - ad) <= 1e-6 * abs(fd), (fd, ad)
    mm.close()


def test_q1_batch_against_single_boxes():
    from poromec_amd.solver import MefSolver, SolverConfig
    decks = P.batch_decks(nbox=4, seed=7, nmin=3, nmax=6, nen=8)
    big, off = P.concat_decks(decks)
    sb = MefSolver(big)
    sb.batch_define(off)
    cfg = SolverConfig(tol=1e-10)
    sb.step(cfg)
    its = sb.batch_iters()
    u_all = sb.state()["u"]
    o = 0
    for k, d in enumerate(decks):
        one = MefSolver(d)
        info = one.step(cfg)
        assert E.close_count(int(its[k]), info["iters"][0]), (k, its[k], info["iters"])
        u = one.state()["u"]
        part = u_all[o: o + d.nnodev]
        o += d.nnodev
        assert _rel(part[:, :3], u[:, :3]) <= RTOL_HIST, k
        assert _rel(part[:, 3], u[:, 3]) <= RTOL_HIST, k
        one.close()
    sb.close()


def test_q1_rejects_what_it_does_not_have():
    from poromec_amd._lib import PmError
    from poromec_amd.solver import MefSolver
    d = q1_deck(2, 0, loads=False)
    s = MefSolver(d)
    with pytest.raises(PmError):
        s.set_vprop(np.full(d.e.shape[0], 18, np.int32), (1, 0, 0, 0))
    s.set_vprop(np.full(d.e.shape[0], 7, np.int32), (0, 0, 0, 0))  # accepted: nothing to do
    s.set_ordered(True)
    with pytest.raises(PmError):
        s.pform(2, True, True, d.dt, d.dt)
    s.close()
    d.e[:, 8] = 0.25  # the adjoint has no stabilized form
    s = MefSolver(d)
    _, idm, _ = s.numbering()
    free = int(np.nonzero(idm[:, 3] > 0)[0][0])
    with pytest.raises(PmError):
        s.adjoint(1, [(free, 3)], np.zeros((1, 1)))
    s.close()
