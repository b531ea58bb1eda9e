"""Pins the oracle's hexa20/8 element (elmt17_pm, Elmlib03.f:1822-2636).

* bit-for-bit against the reference's own Fortran elmt17_pm (built from the
  reference sources into oracle/_ref) for isw 2, 3, 4 and 7, with gravity,
  type-40 tractions on two faces and a type-43 flux;
* closed forms that do not go through either code: rigid-body modes of Kuu,
  partition of unity in Kpu, the exact trilinear mass and Laplacian of a
  unit cube in Kpp, the serendipity face-load split (-1/12 corners, 1/3
  midsides), isw 2 / isw 4 consistency, isw 7 on a linear field and isw 3
  (vertex stresses and Darcy flux) on linear displacement and pressure
  fields.
"""
import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

CORNERS = np.array([(1, 1, 1), (0, 1, 1), (0, 0, 1), (1, 0, 1),
                    (1, 1, 0), (0, 1, 0), (0, 0, 0), (1, 0, 0)], float)
EDGES = [(1, 2), (2, 3), (3, 4), (4, 1), (5, 6), (6, 7), (7, 8), (8, 5),
         (1, 5), (2, 6), (3, 7), (4, 8)]


def node_coords(xl):
    """coordinates of all 20 nodes of the subparametric element"""
    mids = np.array([(xl[a - 1] + xl[b - 1]) / 2 for a, b in EDGES])
    return np.vstack([xl, mids])


def distorted(rng, amp=0.08):
    return CORNERS * np.array([1.3, 0.9, 1.1]) + amp * rng.standard_normal((8, 3))


def env(dt=0.07, t=0.3, gravity=(0.0, 0.0, 0.0), gmod=10.0, load=None, fload=None):
    if load is None:
        load, fload = P.empty_load_table()
    return O.make_env(load, fload, dt, t, gravity, gmod)


@pytest.mark.parametrize("isw", [2, 3, 4, 7])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_element_bitwise_vs_reference(ref, isw, seed):
    rng = np.random.default_rng(seed)
    xl = distorted(rng)
    e = P.material_vector([rng.uniform(5, 50)], rng.uniform(0.25, 0.35), 0.1 * rng.uniform(0.5, 2))[0]
    e[5] = 2000.0
    ul = rng.standard_normal(68)
    dpl = rng.standard_normal(8)
    load, fload = P.empty_load_table()
    P.set_traction_load(load, fload, 1, (0.1, -0.2, -1e-3))
    load[1] = (43, 2, 2)
    fload[1, 0, :2] = (0.0, 1e30)
    fload[1, 1, :2] = (0.5, 0.5)
    iqm = np.array([1, 0, 0, 0, 1, 0, 0], np.int32)
    iqh = np.array([0, 0, 1, 0, 0, 0, 0], np.int32)
    grav = (0.1, -0.3, -9.81)
    p0 = np.zeros(120) if isw == 3 else None  # isw 3 fills 8 x (6 + 6 + 3) values
    rc, p, s = O.elmt17(e, iqm, iqh, xl, ul, dpl, isw, env(0.07, 0.3, grav, 9.81, load, fload),
                        p=p0)
    assert rc == 0
    pr, sr = O.ref_elmt17(e, iqm, iqh, xl, ul, dpl, isw, 0.07, 0.3, grav, 9.81, load, fload,
                          p=p0)
    assert np.array_equal(p, pr)
    if isw == 2:
        assert np.array_equal(s, sr)


def _stiff(xl, E=20.0, nu=0.3, kmu=0.1, dt=0.05, M=1000.0, alpha=1.0):
    e = P.material_vector([E], nu, kmu, M=M, alpha=alpha)[0]
    z7 = np.zeros(7, np.int32)
    rc, p, s = O.elmt17(e, z7, z7, xl, np.zeros(68), np.zeros(8), 2, env(dt=dt))
    assert rc == 0
    return e, s


def test_kuu_rigid_body_modes():
    xl = distorted(np.random.default_rng(3))
    _, s = _stiff(xl)
    kuu = s[:60, :60]
    scale = np.abs(kuu).max()
    X = node_coords(xl)
    for d in range(3):
        t = np.zeros((20, 3))
        t[:, d] = 1.0
        assert np.abs(kuu @ t.ravel()).max() < 1e-12 * scale
    for w in np.eye(3):
        rot = np.cross(w, X)
        assert np.abs(kuu @ rot.ravel()).max() < 1e-12 * scale
    assert np.abs(kuu - kuu.T).max() < 1e-13 * scale
    assert np.linalg.eigvalsh((kuu + kuu.T) / 2)[6] > 0  # only 6 zero modes


def test_kpu_partition_of_unity_and_divergence():
    _, s = _stiff(CORNERS.copy(), alpha=0.8)
    kpu = s[60:, :60]
    for d in range(3):
        t = np.zeros((20, 3))
        t[:, d] = 1.0
        assert np.abs(kpu @ t.ravel()).max() < 1e-14
    # u = (0, 0, z) has unit divergence: Kpu u = -alpha * int(N_i) = -alpha/8
    X = node_coords(CORNERS)
    u = np.zeros((20, 3))
    u[:, 2] = X[:, 2]
    assert np.allclose(kpu @ u.ravel(), -0.8 / 8, rtol=0, atol=1e-14)
    assert np.array_equal(s[:60, 60:], kpu.T)


def test_kpp_unit_cube_closed_form():
    M, kmu, dt = 700.0, 0.37, 0.05
    _, s = _stiff(CORNERS.copy(), kmu=kmu, dt=dt, M=M)
    kpp = s[60:, 60:]
    # exact trilinear mass and Laplacian of [0,1]^3 from 1-D factors
    m1 = {True: 1 / 3, False: 1 / 6}
    k1 = {True: 1.0, False: -1.0}
    mass = np.zeros((8, 8))
    lap = np.zeros((8, 8))
    for i in range(8):
        for j in range(8):
            same = [CORNERS[i, d] == CORNERS[j, d] for d in range(3)]
            mass[i, j] = m1[same[0]] * m1[same[1]] * m1[same[2]]
            lap[i, j] = sum(k1[same[d]] * np.prod([m1[same[q]] for q in range(3) if q != d])
                            for d in range(3))
    expect = -mass / M - dt * kmu * lap
    assert np.allclose(kpp, expect, rtol=1e-13, atol=1e-16)


def test_isw4_consistent_with_isw2():
    rng = np.random.default_rng(5)
    xl = distorted(rng)
    e, s = _stiff(xl)
    z7 = np.zeros(7, np.int32)
    u = rng.standard_normal(68)
    rc, p4, _ = O.elmt17(e, z7, z7, xl, u, np.zeros(8), 4, env(dt=0.05))
    scale = np.abs(s).max() * np.abs(u).max()
    # u rows: int B^T D eps(u) = Kuu u;  p rows: -dt perm Lap p
    assert np.allclose(p4[:60], s[:60, :60] @ u[:60], atol=1e-12 * scale, rtol=0)
    e_ = e.copy()
    e_[3] = 1e300  # 1/M -> 0 isolates the flow term of Kpp
    _, s_ = _stiff(xl, E=e[0], nu=e[1], kmu=0.1, dt=0.05, M=1e300)
    assert np.allclose(p4[60:], s_[60:, 60:] @ u[60:], atol=1e-12 * scale, rtol=0)
    # dp term: u rows equal Kup dp
    dp = rng.standard_normal(8)
    rc, p4d, _ = O.elmt17(e, z7, z7, xl, np.zeros(68), dp, 4, env(dt=0.05))
    assert np.allclose(p4d[:60], s[:60, 60:] @ dp, atol=1e-12 * np.abs(s).max(), rtol=0)


def test_face_load_serendipity_split():
    load, fload = P.empty_load_table()
    q = 2.5e-3
    P.set_traction_load(load, fload, 1, (0.0, 0.0, -q))
    e = P.material_vector([20.0], 0.3, 0.1)[0]
    iqm = np.array([1, 0, 0, 0, 0, 0, 0], np.int32)
    z7 = np.zeros(7, np.int32)
    rc, p, _ = O.elmt17(e, iqm, z7, CORNERS, np.zeros(68), np.zeros(8), 4,
                        env(load=load, fload=fload))
    f = p[:60].reshape(20, 3)
    top_corners = [0, 1, 2, 3]
    top_mids = [8, 9, 10, 11]
    assert np.allclose(f[top_corners, 2], -q / 12, rtol=1e-13)
    assert np.allclose(f[top_mids, 2], q / 3, rtol=1e-13)
    assert np.abs(np.delete(f[:, 2], top_corners + top_mids)).max() == 0
    assert np.abs(f[:, :2]).max() == 0


def test_isw7_linear_field():
    rng = np.random.default_rng(7)
    xl = distorted(rng)
    e = P.material_vector([20.0], 0.3, 0.1, M=500.0, alpha=0.9)[0]
    a, b, c = 0.3, -0.2, 0.45
    X = node_coords(xl)
    u = np.zeros(68)
    u[:60] = (X * np.array([a, b, c])).ravel()
    u[60:] = rng.standard_normal(8)
    z7 = np.zeros(7, np.int32)
    rc, p, _ = O.elmt17(e, z7, z7, xl, u, np.zeros(8), 7, env())
    assert np.allclose(p[:8], 0.9 * (a + b + c) + u[60:] / 500.0, rtol=1e-12, atol=1e-14)


def test_isw3_linear_fields():
    """A linear displacement field has the constant strain sym(G) at every
    vertex, so the vertex stress is D sym(G) - alpha dp (total) and that plus
    alpha p (Biot); a linear pressure g.x + p0 has the Darcy flux
    perm (rho_f grav - g) everywhere."""
    rng = np.random.default_rng(11)
    xl = distorted(rng)
    E, nu, alpha, kw = 20.0, 0.3, 0.9, 0.1
    e = P.material_vector([E], nu, kw, M=500.0, alpha=alpha)[0]
    G = rng.standard_normal((3, 3)) * 1e-3
    X = node_coords(xl)
    gp, p0 = np.array([0.3, -0.2, 0.45]), 0.7
    u = np.zeros(68)
    u[:60] = (X @ G.T).ravel()
    u[60:] = xl @ gp + p0
    dp = rng.standard_normal(8)
    grav = (0.1, -0.3, -9.81)
    z7 = np.zeros(7, np.int32)
    rc, p, _ = O.elmt17(e, z7, z7, xl, u, dp, 3, env(gravity=grav, gmod=9.81), p=np.zeros(120))
    assert rc == 0
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[1, 2] + G[2, 1],
                    G[0, 2] + G[2, 0]])
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    sig = np.concatenate([lam * eps[:3].sum() + 2 * mu * eps[:3], mu * eps[3:]])
    perm = e[2] / (e[6] * 1e-6 * 9.81)
    q = perm * (e[6] * 1e-6 * np.array(grav) - gp)
    for i in range(8):
        tot = sig - alpha * dp[i] * np.array([1, 1, 1, 0, 0, 0])
        biot = tot + alpha * u[60 + i] * np.array([1, 1, 1, 0, 0, 0])
        assert np.allclose(p[6 * i:6 * i + 6], tot, rtol=1e-11, atol=1e-13)
        assert np.allclose(p[48 + 6 * i:54 + 6 * i], biot, rtol=1e-11, atol=1e-13)
        assert np.allclose(p[96 + 3 * i:99 + 3 * i], q, rtol=1e-11, atol=1e-20)
