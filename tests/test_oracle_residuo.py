"""The oracle's nonlinear stop test (cal_residuo_pm, Residuo.f:39-136) and
solv_pm's SQMR -> GMRES fallback (Solv.f:92-94,157,765; Iters.f:6343-6349),
pinned by known answers worked out by hand from the reference's text.

Criterion 2 (the reference default, Main.f:273) keeps separate scales for
the displacement and the pressure residual.  At nli = 1 they are the
current norms; at nli = 2 both are raised by max() -- and the pressure scale
by the *displacement* norm, resp0 = max(resp0, gresu) (Residuo.f:119).  The
vectors below are chosen so that the quirk, its "corrected" form
max(resp0, gresp), an update at every nli, a swapped u/p partition and a
guard against a zero scale each give a different istop."""
import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

TOL = 1e-4


@pytest.fixture(scope="module")
def column():
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    return d, st, O.OracleModel(d, st)


def _vec(st, ru, rp, seed=0):
    """b with ||b_u|| = ru over the displacement equations and ||b_p|| = rp
    over the pressure equations (id(1..3, node) and id(4, node), Residuo.f:86-101)."""
    rng = np.random.default_rng(seed)
    idm = st["id"]
    qu = idm[:, :3][idm[:, :3] > 0] - 1
    qp = idm[:, 3][idm[:, 3] > 0] - 1
    b = np.zeros(st["neq"])
    vu = rng.standard_normal(len(qu))
    vp = rng.standard_normal(len(qp))
    b[qu] = ru * vu / np.linalg.norm(vu)
    b[qp] = rp * vp / np.linalg.norm(vp)
    return b


def test_partition_norms(column):
    """state[2], state[3] are ||R_u|| and ||R_p|| over the id map; u and p
    equations together are every equation."""
    d, st, m = column
    idm = st["id"]
    assert np.count_nonzero(idm[:, :3]) + np.count_nonzero(idm[:, 3]) == st["neq"]
    assert np.count_nonzero(idm[d.nnodev:, 3]) == 0  # mid-edge nodes carry no pressure
    b = _vec(st, 3.0, 0.25)
    state = np.zeros(4)
    m.residuo(b, 1, TOL, state, 2)
    assert state[2] == pytest.approx(3.0, rel=1e-14)
    assert state[3] == pytest.approx(0.25, rel=1e-14)


def test_criterion2_known_answers(column):
    """Three nonlinear iterations worked out by hand from Residuo.f:112-136.

    nli 1: |Ru| = 1,    |Rp| = 1e-8  -> resu0 = 1, resp0 = 1e-8; ratios 1, 1: istop 0
    nli 2: |Ru| = 1e-3, |Rp| = 1e-9  -> resu0 = max(1, 1e-3) = 1,
           resp0 = max(1e-8, gresu = 1e-3) = 1e-3; u: 1e-3 >= tol, p: 1e-6 < tol: istop 1
           (with max(resp0, gresp): resp0 = 1e-8, p ratio 0.1, istop 0)
    nli 3: |Ru| = 5e-5, |Rp| = 2e-8  -> scales kept; u: 5e-5 < tol, p: 2e-5 < tol: istop 2
           (with the scales reset to the current norms at every nli: istop 0)"""
    d, st, m = column
    state = np.zeros(4)
    seq = [((1.0, 1e-8), 0, (1.0, 1e-8)),
           ((1e-3, 1e-9), 1, (1.0, 1e-3)),
           ((5e-5, 2e-8), 2, (1.0, 1e-3))]
    for nli, ((ru, rp), istop, (s0, s1)) in enumerate(seq, start=1):
        got = m.residuo(_vec(st, ru, rp, seed=nli), nli, TOL, state, 2)
        assert got == istop, (nli, got)
        assert state[0] == pytest.approx(s0, rel=1e-13), (nli, state)
        assert state[1] == pytest.approx(s1, rel=1e-13), (nli, state)


def test_criterion2_zero_displacement_scale(column):
    """|Ru| = 0 at nli 1 makes resu0 = 0; gresu/resu0 is then 0/0 (NaN) and
    never below tol, so istop stays below 2 however small the residual:
    the mechanism that keeps criterion 2 from converging after the load step."""
    d, st, m = column
    state = np.zeros(4)
    assert m.residuo(_vec(st, 0.0, 1.0), 1, TOL, state, 2) == 0
    assert m.residuo(_vec(st, 0.0, 1e-12), 2, TOL, state, 2) == 1
    assert m.residuo(_vec(st, 0.0, 0.0), 3, TOL, state, 2) == 1


def test_criterion1_known_answers(column):
    """Criterion 1 (Residuo.f:59-75): one norm over all equations, res0 set at
    nli 1 and raised by max() at nli 2; istop is 0 or 2."""
    d, st, m = column
    state = np.zeros(4)
    assert m.residuo(_vec(st, 0.6, 0.8), 1, TOL, state, 1) == 0          # res 1
    assert state[0] == pytest.approx(1.0, rel=1e-14)
    assert m.residuo(_vec(st, 3.0, 4.0), 2, TOL, state, 1) == 0          # res0 = max(1, 5)
    assert state[0] == pytest.approx(5.0, rel=1e-14)
    assert m.residuo(_vec(st, 6e-4, 8e-4), 3, TOL, state, 1) == 0        # 2e-4 is not < tol
    assert state[0] == pytest.approx(5.0, rel=1e-14)                     # kept after nli 2
    assert m.residuo(_vec(st, 2.4e-4, 3.2e-4), 4, TOL, state, 1) == 2    # 8e-5 < tol


def test_criterion2_march_reaches_maxnlit():
    """The column march under criterion 2: the load step converges (nli 3),
    the next step cannot (its displacement residual is round-off from the
    start) and ends at maxnlit -- -10 (stop_mef, Main.f:1046-1053), or a kept
    step with maxnlit_accept.  Criterion 1 converges in both steps."""
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    m1 = O.OracleModel(d, st, stop_crit=1, maxnlit=4)
    assert [m1.step()["nlit"] for _ in range(2)] == [2, 2]
    m2 = O.OracleModel(d, st, stop_crit=2, maxnlit=4)
    s1 = m2.step(check=False)
    assert (s1["rc"], s1["nlit"], s1["nsolve"]) == (0, 3, 2)
    s2 = m2.step(check=False)
    assert (s2["rc"], s2["nlit"], s2["nsolve"]) == (-10, 4, 4)
    m3 = O.OracleModel(d, st, stop_crit=2, maxnlit=4, maxnlit_accept=True)
    m3.step()
    s2a = m3.step()
    assert (s2a["rc"], s2a["nlit"], s2a["nsolve"]) == (0, 4, 4)
    assert s2a["iters"] == s2["iters"]


def _saddle():
    d = P.column_deck(2, 2, 6, dt=0.01, nsteps=6)
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    return st, m


def _dense(st, m):
    n = st["neq"]
    A = np.diag(m.ad.copy())
    for i in range(n):
        for k in range(st["ia"][i] - 1, st["ia"][i + 1] - 1):
            j = st["ja"][k] - 1
            A[i, j] = A[j, i] = m.al[k]
    return A


def test_sqmr_falls_back_to_gmres():
    """SQMR that reaches maxit sets fsqmr, and solv_pm solves again with
    GMRES(ngram) from x = 0 (gmres2 is called with fnew = .true.,
    Solv.f:2292-2320) under the same maxit.  The fallback's result is
    therefore a fresh GMRES solve bit for bit, and it solves the system."""
    st, m = _saddle()
    b = np.random.default_rng(7).standard_normal(st["neq"])
    A = _dense(st, m)
    it_s, x_s, used_s = O.solv(st, m.ad, m.al, b, solver=11, tol=1e-10, maxit=5)
    assert used_s == 2
    it_g, x_g, used_g = O.solv(st, m.ad, m.al, b, solver=2, tol=1e-10, maxit=5)
    assert used_g == 2 and it_g > 0
    assert it_s == it_g
    assert np.array_equal(x_s, x_g)
    # gmres2 stops on the left-Jacobi-preconditioned residual (Iters.f:1411-1775)
    minv = 1.0 / m.ad
    assert np.linalg.norm(minv * (b - A @ x_s)) < 1e-8 * np.linalg.norm(minv * b)
    # without the limit SQMR finishes on its own
    it_q, x_q, used_q = O.solv(st, m.ad, m.al, b, solver=11, tol=1e-10, maxit=50000)
    assert used_q == 11 and it_q > 5
    assert np.linalg.norm(b - A @ x_q) < 1e-8 * np.linalg.norm(b)
