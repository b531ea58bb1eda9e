"""The batching form on the GPU: boxes of different sizes and materials held
as one block-diagonal system must step exactly like each box stepped on its
own (histories within 1e-9; Krylov iterations within 2 + 10 %, the width of
the oracle's own spread under 1-ulp input perturbations on these systems,
since the per-box reductions sum in another order), replaying the
captured hipGraph chunks must give the eager launches' results bit for bit,
and the FP32+IR batched solve must reach the FP64 tolerance in every box."""
import numpy as np
import pytest

from poromec_amd import problem as P

pytestmark = pytest.mark.gpu


def _boxes():
    decks = [P.heterogeneous_box_deck(n, seed=10 + n, sigma_lnk=s, dt=0.1, corr_len=2.0)
             for n, s in ((3, 0.5), (5, 1.2), (4, 1.0))]
    return decks, *P.concat_decks(decks)


def _split(u_all, decks):
    """per-box node arrays: vertices first, then mid-edge nodes, box by box"""
    nv_tot = sum(d.nnodev for d in decks)
    out, vo, eo = [], 0, nv_tot
    for d in decks:
        c = P.box_counts(d.nx, d.ny, d.nz)
        ne = c["nnode"] - c["nnodev"]
        out.append(np.concatenate([u_all[vo: vo + d.nnodev], u_all[eo: eo + ne]]))
        vo += d.nnodev
        eo += ne
    return out


def test_batched_steps_match_single_boxes():
    from poromec_amd.solver import MefSolver, SolverConfig
    decks, big, off = _boxes()
    cfg = SolverConfig(tol=1e-10)
    sb = MefSolver(big)
    sb.batch_define(off)
    singles = [MefSolver(d) for d in decks]
    for step in range(3):
        sb.step(cfg)
        its = sb.batch_iters()
        for k, s in enumerate(singles):
            info = s.step(cfg)
            a, b = info["iters"][0], int(its[k])
            assert abs(a - b) <= 2 + 0.10 * max(a, b), (step, k, info["iters"], its)
        parts = _split(sb.state()["u"], decks)
        for k, s in enumerate(singles):
            u = s.state()["u"]
            nv = decks[k].nnodev
            assert np.abs(parts[k][:, :3] - u[:, :3]).max() <= 1e-9 * np.abs(u[:, :3]).max()
            assert np.abs(parts[k][:nv, 3] - u[:nv, 3]).max() <= 1e-9 * np.abs(u[:nv, 3]).max()


def test_graph_replays_match_eager():
    from poromec_amd.solver import MefSolver, SolverConfig
    decks, big, off = _boxes()
    cfg = SolverConfig(tol=1e-10, solver="sqmr")
    runs = []
    for graphs in (True, False):
        s = MefSolver(big)
        s.batch_define(off)
        s.batch_graphs(graphs)
        hist = []
        for _ in range(3):  # several replays, eager host work in between
            s.step(cfg)
            hist.append(s.state()["u"].copy())
        runs.append(hist)
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_batched_fp32ir():
    from poromec_amd.solver import MefSolver, SolverConfig
    decks, big, off = _boxes()
    s = MefSolver(big)
    s.batch_define(off)
    s.pform(2, True, False, big.dt, big.dt)
    b = np.random.default_rng(3).standard_normal(s.neq)
    x64, i64 = s.solve(b, SolverConfig(tol=1e-10))
    x32, i32 = s.solve(b, SolverConfig(tol=1e-10, precision="fp32ir"))
    assert i64["true_rel"] < 1e-9 and i32["true_rel"] < 1e-10
    assert np.abs(x32 - x64).max() <= 1e-8 * np.abs(x64).max()


def test_batched_every_spmv_shape():
    """Every SpMV shape of the batched kernel (the run-per-lane ones
    included), FP64 and FP32+IR: each box reaches the tolerance and the
    solution of the default shape."""
    from poromec_amd.solver import SPMV_SHAPES, MefSolver, SolverConfig
    decks, big, off = _boxes()
    s = MefSolver(big)
    s.batch_define(off)
    s.pform(2, True, False, big.dt, big.dt)
    b = np.random.default_rng(4).standard_normal(s.neq)
    x0, _ = s.solve(b, SolverConfig(tol=1e-10))
    for shape in range(SPMV_SHAPES):
        s.set_spmv_shape(False, shape)
        s.set_spmv_shape(True, shape)
        for prec in ("fp64", "fp32ir"):
            x, info = s.solve(b, SolverConfig(tol=1e-10, precision=prec))
            assert info["true_rel"] < 1e-9, (shape, prec, info["true_rel"])
            assert np.abs(x - x0).max() <= 1e-8 * np.abs(x0).max(), (shape, prec)


def test_batched_fp32ir_straggler_fallback():
    """One box far larger than the others: once half the boxes have converged
    at L inner iterations, the FP32 inner solve stops at 4 L and the boxes
    still open get their correction from an FP64 solve.  The result still
    reaches the tolerance and the FP64 solution."""
    from poromec_amd.solver import MefSolver, SolverConfig
    decks = [P.heterogeneous_box_deck(n, seed=30 + n + k, sigma_lnk=s, dt=0.1, corr_len=2.0)
             for k, (n, s) in enumerate(((2, 0.3), (2, 0.3), (2, 0.3), (7, 1.5)))]
    big, off = P.concat_decks(decks)
    s = MefSolver(big)
    s.batch_define(off)
    s.pform(2, True, False, big.dt, big.dt)
    b = np.random.default_rng(5).standard_normal(s.neq)
    x64, _ = s.solve(b, SolverConfig(tol=1e-10))
    x32, info = s.solve(b, SolverConfig(tol=1e-10, precision="fp32ir"))
    assert info["spmv64"] > info["outer"], info  # FP64 fallback products beyond the residuals
    assert info["true_rel"] < 1e-10
    assert np.abs(x32 - x64).max() <= 1e-8 * np.abs(x64).max()


def test_batched_criterion2():
    """Criterion 2 (Residuo.f:84-136) run per box in the batching form: the
    load step converges in every box at the oracle's nonlinear iteration
    (3), with each box's state within 1e-9 of the box stepped alone; the next
    step reaches maxnlit in every box, as the oracle does (its -10), and the
    batched step returns PM_ERR_NLIT."""
    from oracle import oracle as O
    from poromec_amd._lib import PmError
    from poromec_amd.solver import MefSolver, SolverConfig
    decks, big, off = _boxes()
    cfg = SolverConfig(tol=1e-10, stop_crit=2, maxnlit=4)
    for d in decks:
        m = O.OracleModel(d, stop_crit=2, maxnlit=4)
        assert [(i["rc"], i["nlit"]) for i in (m.step(check=False), m.step(check=False))] == \
            [(0, 3), (-10, 4)]
    sb = MefSolver(big)
    sb.batch_define(off)
    info = sb.step(cfg)
    assert info["nlit"] == 3, info
    singles = [MefSolver(d) for d in decks]
    for s in singles:
        assert s.step(cfg)["nlit"] == 3
    parts = _split(sb.state()["u"], decks)
    for k, s in enumerate(singles):
        u = s.state()["u"]
        nv = decks[k].nnodev
        assert np.abs(parts[k][:, :3] - u[:, :3]).max() <= 1e-9 * np.abs(u[:, :3]).max()
        assert np.abs(parts[k][:nv, 3] - u[:nv, 3]).max() <= 1e-9 * np.abs(u[:nv, 3]).max()
    with pytest.raises(PmError) as e:
        sb.step(cfg)
    assert e.value.code == -7
    for s in singles:
        with pytest.raises(PmError) as e:
            s.step(cfg)
        assert e.value.code == -7
        s.close()
    sb.close()


def test_resetup_drops_the_batch():
    """A new mesh on a context that held a batch definition drops it (the box
    ranges and the captured graph refer to the old matrix): afterwards the
    context solves as one system and the batch queries report PM_ERR_STATE."""
    from poromec_amd import _lib as L
    from poromec_amd.solver import MefSolver, SolverConfig
    import ctypes as C
    decks, big, off = _boxes()
    sb = MefSolver(big)
    sb.batch_define(off)
    sb.step(SolverConfig(tol=1e-10))
    d = decks[1]
    ix, x, idv = L.as_i32(d.ix), L.as_f64(d.x), L.as_i32(d.idv)
    nnode, neq, nad = C.c_int(), C.c_int(), C.c_int64()
    L.check(sb.lib.pm_mesh_setup(sb.ctx, d.numel, d.nnodev, L.ptr(ix), L.ptr(x), L.ptr(idv), 1,
                                 C.byref(nnode), C.byref(neq), C.byref(nad)), "pm_mesh_setup")
    it = np.zeros(sb.nbox, np.int32)
    assert sb.lib.pm_batch_iters(sb.ctx, L.ptr(it)) == -8
    sb.deck, sb.nnode, sb.neq = d, nnode.value, neq.value
    z4, zn = np.zeros((sb.nnode, 4)), np.zeros(sb.nnode)
    sb.set_state(u=z4, u0=z4, dp=zn, t=0.0)
    sb.set_materials(d)
    info = sb.step(SolverConfig(tol=1e-10))
    one = MefSolver(d)
    ref = one.step(SolverConfig(tol=1e-10))
    assert abs(info["iters"][0] - ref["iters"][0]) <= 2 + 0.1 * ref["iters"][0]
    u, u1 = sb.state()["u"], one.state()["u"]
    assert np.abs(u - u1).max() <= 1e-9 * np.abs(u1).max()
    one.close()
    sb.close()
