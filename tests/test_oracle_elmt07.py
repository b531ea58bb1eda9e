"""Pins the oracle's equal-order Q1/Q1 hexa8 element (element type 7,
oracle/pm_elmt07.c).

The reference reserves this element (ELMT07_PM, Elmlib03.f:22-23) without
implementing it, so nothing of the reference can be compared with it; every
check below is a closed form or an invariant that a plausible slip (a
transposed Jacobian, a dropped storage or flow term, a wrong sign or index in
the coupling, a mis-ordered face) breaks:

* on bricks, the whole 32 x 32 matrix against exact integrals assembled from
  Kronecker products of 1-D mass, stiffness and gradient matrices with the
  Lame form of D (not the element's a, b, c form, not Gauss quadrature);
* rotation covariance of Kuu and Kpu, rigid-body and constant-pressure null
  spaces and div of a linear field in Kpu on distorted elements;
* the patch test on an assembled distorted mesh: a linear displacement field
  leaves zero force on the interior node;
* isw 4 against isw 2 on the same state, isw 3 and isw 7 on linear fields,
  and face loads against traction x area and flux x area;
* the consolidation column against Terzaghi's undrained response, series
  solution and final settlement.
"""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

CORNERS = np.array([(1, 1, 1), (0, 1, 1), (0, 0, 1), (1, 0, 1),
                    (1, 1, 0), (0, 1, 0), (0, 0, 0), (1, 0, 0)], float)
FACES = [(1, 2, 3, 4), (5, 6, 7, 8), (1, 5, 6, 2), (4, 3, 7, 8), (1, 4, 8, 5), (2, 6, 7, 3)]


def env(dt=0.07, t=0.3, gravity=(0.0, 0.0, 0.0), gmod=10.0, load=None, fload=None):
    if load is None:
        load, fload = P.empty_load_table()
    return O.make_env(load, fload, dt, t, gravity, gmod)


def material(E=20.0, nu=0.3, k=0.1, M=50.0, alpha=0.8, rho=2000.0):
    return P.material_vector([E], nu, k, M=M, alpha=alpha, rho=rho)[0]


def lame(E, nu):
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def element(xl, e, isw, ev, ul=None, dpl=None, iqm=None, iqh=None):
    ul = np.zeros(32) if ul is None else ul
    dpl = np.zeros(8) if dpl is None else dpl
    iqm = np.zeros(7, np.int32) if iqm is None else iqm
    iqh = np.zeros(7, np.int32) if iqh is None else iqh
    rc, p, s = O.elmt07(e, iqm, iqh, xl, ul, dpl, isw, ev)
    assert rc == 0
    return p, s


def brick_exact(h, E, nu, alpha, M, perm, dt):
    """The Q1/Q1 element matrix of a brick from exact 1-D integrals."""
    lam, mu = lame(E, nu)
    idx = CORNERS.astype(int)
    m1 = [L / 6 * np.array([[2.0, 1.0], [1.0, 2.0]]) for L in h]
    k1 = [np.array([[1.0, -1.0], [-1.0, 1.0]]) / L for L in h]
    g1 = [np.array([[-0.5, 0.5], [-0.5, 0.5]]) for L in h]  # g1[a][b] = int phi_a phi_b'

    def integ(i, j, di, dj):  # int (d_di N_i)(d_dj N_j), None = no derivative
        v = 1.0
        for d in range(3):
            a, b = idx[i, d], idx[j, d]
            if di == d and dj == d:
                v *= k1[d][a, b]
            elif di == d:
                v *= g1[d][b, a]
            elif dj == d:
                v *= g1[d][a, b]
            else:
                v *= m1[d][a, b]
        return v

    K = np.zeros((32, 32))
    for i in range(8):
        for j in range(8):
            lap = sum(integ(i, j, c, c) for c in range(3))
            for a in range(3):
                for b in range(3):
                    K[3 * i + a, 3 * j + b] = (lam * integ(i, j, a, b) + mu * integ(i, j, b, a)
                                               + (mu * lap if a == b else 0.0))
                K[24 + i, 3 * j + a] = -alpha * integ(i, j, None, a)
                K[3 * j + a, 24 + i] = K[24 + i, 3 * j + a]
            K[24 + i, 24 + j] = -integ(i, j, None, None) / M - dt * perm * lap
    return K


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (1.3, 0.7, 2.1)])
def test_brick_matrix_exact(h):
    E, nu, k, M, alpha, dt = 20.0, 0.3, 0.1, 50.0, 0.8, 0.07
    xl = CORNERS * np.array(h) + np.array([0.4, -1.2, 3.0])
    _, s = element(xl, material(E, nu, k, M, alpha), 2, env(dt))
    ref = brick_exact(h, E, nu, alpha, M, k, dt)
    np.testing.assert_allclose(s, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    np.testing.assert_array_equal(s[:24, 24:], s[24:, :24].T)  # Kup stored as Kpu^T
    np.testing.assert_allclose(s, s.T, rtol=0, atol=1e-15 * np.abs(s).max())


def test_stabilization_term():
    """e(9) = beta adds -tau int grad N . grad N to Kpp only, tau = beta h^2 /
    (lambda + 2 mu) with h the cube root of the element's volume."""
    h = (1.3, 0.7, 2.1)
    E, nu, k, M, alpha, dt, beta = 20.0, 0.3, 0.1, 50.0, 0.8, 0.07, 0.25
    xl = CORNERS * np.array(h)
    e0 = material(E, nu, k, M, alpha)
    e1 = e0.copy()
    e1[8] = beta
    _, s0 = element(xl, e0, 2, env(dt))
    _, s1 = element(xl, e1, 2, env(dt))
    lam, mu = lame(E, nu)
    tau = beta * np.prod(h) ** (2.0 / 3.0) / (lam + 2 * mu)
    lap = (brick_exact(h, E, nu, alpha, 1e300, 1.0, 1.0) - brick_exact(h, E, nu, alpha, 1e300, 0.0, 1.0))[24:, 24:]
    d = s1 - s0
    np.testing.assert_array_equal(d[:24], 0.0)
    np.testing.assert_array_equal(d[:, :24], 0.0)
    np.testing.assert_allclose(d[24:, 24:], tau * lap, atol=1e-14 * np.abs(tau * lap).max())


def test_stabilized_column_undrained_step():
    """The benchmark column's near-undrained first step (M / Kv = 74): the
    equal-order pressure oscillates (over 20 % of p0 at 2 m below the drained
    top, 8 % at 5 m); with beta = 1/4 it is p0 to 1e-3 below 2 m."""
    E, nu, M, q = 10.0, 0.3, 1000.0, 1.0e-3
    Kv = E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    p0 = M * q / (Kv + M)
    errs = {}
    for beta in (0.0, 0.25):
        d = P.column_deck(4, 4, 40, dt=0.01, nsteps=1, nen=8)
        d.e[:, 8] = beta
        m = O.OracleModel(d, precond=5)  # |diag|: signed-diagonal SQMR stagnates here
        m.step()
        z = d.x[:, 2]
        errs[beta] = [np.abs(m.u[z <= z.max() - th, 3] / p0 - 1.0).max() for th in (2.0, 5.0)]
    assert errs[0.0][0] > 0.2 and errs[0.0][1] > 0.05
    assert errs[0.25][0] < 1e-3 and errs[0.25][1] < 1e-6


def rotation(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.linalg.det(q))


def distorted(rng, amp=0.12):
    return CORNERS * np.array([1.3, 0.9, 1.1]) + amp * rng.standard_normal((8, 3))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rotation_covariance(seed):
    rng = np.random.default_rng(seed)
    xl = distorted(rng)
    R = rotation(rng)
    e = material()
    _, s = element(xl, e, 2, env())
    _, sr = element(xl @ R.T, e, 2, env())
    Q = np.kron(np.eye(8), R)
    scale = np.abs(s).max()
    np.testing.assert_allclose(sr[:24, :24], Q @ s[:24, :24] @ Q.T, atol=1e-12 * scale)
    np.testing.assert_allclose(sr[24:, :24], s[24:, :24] @ Q.T, atol=1e-12 * scale)
    np.testing.assert_allclose(sr[24:, 24:], s[24:, 24:], atol=1e-12 * scale)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_null_spaces_and_divergence(seed):
    rng = np.random.default_rng(seed)
    xl = distorted(rng)
    e = material(M=50.0, alpha=0.8)
    _, s = element(xl, e, 2, env())
    scale = np.abs(s).max()
    # rigid-body modes: 3 translations, 3 linearised rotations
    modes = []
    for d in range(3):
        t = np.zeros((8, 3))
        t[:, d] = 1.0
        modes.append(t.ravel())
    for w in np.eye(3):
        modes.append(np.cross(w, xl).ravel())
    for r in modes:  # no force, and no volume change (div = 0) for the pressure rows
        assert np.abs(s[:24, :24] @ r).max() < 1e-12 * scale
        assert np.abs(s[24:, :24] @ r).max() < 1e-12 * scale
    # constant pressure: no flux (Kpp + S) 1 = 0, with S from dt = 0
    _, s0 = element(xl, e, 2, env(dt=0.0))
    flow = s[24:, 24:] - s0[24:, 24:]
    assert np.abs(flow @ np.ones(8)).max() < 1e-12 * np.abs(flow).max()
    # Kpu on a linear field u = G x: -alpha tr(G) int N_i = -alpha tr(G) M (S 1)_i
    G = rng.standard_normal((3, 3))
    u_lin = (xl @ G.T).ravel()
    intN = -50.0 * (s0[24:, 24:] @ np.ones(8))
    np.testing.assert_allclose(s[24:, :24] @ u_lin, -0.8 * np.trace(G) * intN,
                               atol=1e-12 * np.abs(intN).max() * np.abs(G).max())
    # and int N over the element is its volume
    vol = abs(np.linalg.det(np.array([xl[1] - xl[0], xl[3] - xl[0], xl[4] - xl[0]])))
    assert 0.5 * vol < intN.sum() < 2.0 * vol


def distorted_mesh(seed):
    """2 x 2 x 2 hexa8 mesh, no restraints, with the interior vertex moved:
    all 8 elements are distorted."""
    d = P.column_deck(2, 2, 2, nen=8)
    x = d.x.copy()
    inner = np.all((x > 0) & (x < 2), axis=1)
    x[inner] += 0.25 * np.random.default_rng(seed).standard_normal((inner.sum(), 3))
    return dataclasses.replace(d, x=x, idv=np.zeros_like(d.idv), eload=np.zeros_like(d.eload)), inner


@pytest.mark.parametrize("seed", [0, 1])
def test_patch_test_linear_field(seed):
    """Assembled over a distorted patch, a linear displacement field leaves
    no force on the interior vertex, and a linear pressure field no flux
    (its pressure row differs from the dt = 0 assembly, storage only, by
    nothing)."""
    d, inner = distorted_mesh(seed)
    (n,) = np.nonzero(inner)[0]
    G = np.random.default_rng(seed + 7).standard_normal((3, 3)) * 1e-3
    gp = np.array([0.3, -0.2, 0.5])

    def residual(u_fn, p_fn, dt):
        m = O.OracleModel(d)
        m.env.dt = dt
        m.u[:, :3] = u_fn(d.x)
        m.u[:, 3] = p_fn(d.x)
        b = np.zeros(m.st["neq"])
        m.assemble(isw=2, lhs=True, rhs=True, bvec=b)
        return b, m.st["id"]

    b, idm = residual(lambda x: x @ G.T, lambda x: 0 * x[:, 0], 0.07)
    f = np.array([b[q - 1] for q in idm[n, :3]])
    assert np.abs(f).max() < 1e-12 * np.abs(b).max(), f
    b1, idm = residual(lambda x: 0 * x, lambda x: x @ gp, 0.07)
    b0, _ = residual(lambda x: 0 * x, lambda x: x @ gp, 0.0)
    q = idm[n, 3] - 1
    assert abs(b1[q] - b0[q]) < 1e-12 * np.abs(b1 - b0).max()
    assert np.abs(b1 - b0).max() > 0  # the flux does act on the boundary vertices


@pytest.mark.parametrize("seed", [0, 1])
def test_isw4_against_isw2(seed):
    """isw 4 (no gravity, no face loads) is [Kuu u + Kup dp ; L p] with L the
    flow part of Kpp, both from isw 2."""
    rng = np.random.default_rng(seed)
    xl = distorted(rng)
    e = material()
    ul = rng.standard_normal(32)
    dpl = rng.standard_normal(8)
    ev = env()
    p4, _ = element(xl, e, 4, ev, ul, dpl)
    _, s = element(xl, e, 2, ev)
    _, s0 = element(xl, e, 2, env(dt=0.0))
    u_dp = ul.copy()
    u_dp[24:] = dpl
    ref_u = s[:24] @ u_dp
    ref_p = (s[24:, 24:] - s0[24:, 24:]) @ ul[24:]
    np.testing.assert_allclose(p4[:24], ref_u, atol=1e-12 * np.abs(ref_u).max())
    np.testing.assert_allclose(p4[24:], ref_p, atol=1e-12 * np.abs(ref_p).max())


def test_isw4_gravity_terms():
    """Gravity alone: -rho g int N_i on the displacement rows (sum: -rho g V)
    and dt rho_f perm int grad N_i . g on the pressure rows (sum: zero)."""
    rng = np.random.default_rng(4)
    xl = distorted(rng)
    e = material(rho=2000.0)
    g = np.array([0.1, -0.3, -9.8])
    p4, _ = element(xl, e, 4, env(gravity=g, gmod=10.0))
    _, s0 = element(xl, e, 2, env(dt=0.0))
    intN = -e[3] * (s0[24:, 24:] @ np.ones(8))
    ref = -(2000.0e-6) * np.outer(intN, g).ravel()
    np.testing.assert_allclose(p4[:24], ref, atol=1e-12 * np.abs(ref).max())
    assert abs(p4[24:].sum()) < 1e-12 * np.abs(p4[24:]).max()


@pytest.mark.parametrize("seed", [0, 1])
def test_isw3_isw7_linear_fields(seed):
    rng = np.random.default_rng(seed)
    xl = distorted(rng)
    E, nu, alpha, M, k = 20.0, 0.3, 0.8, 50.0, 0.1
    e = material(E, nu, k, M, alpha)
    G = rng.standard_normal((3, 3))
    gp = rng.standard_normal(3)
    ul = np.zeros(32)
    ul[:24] = (xl @ G.T).ravel()
    ul[24:] = xl @ gp
    dpl = rng.standard_normal(8)
    grav = np.array([0.0, 0.0, -10.0])
    p3, _ = element(xl, e, 3, env(gravity=grav, gmod=10.0), ul, dpl)
    lam, mu = lame(E, nu)
    eps = 0.5 * (G + G.T)
    sig = lam * np.trace(G) * np.eye(3) + 2 * mu * eps
    voigt = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[1, 2], sig[0, 2]])
    flux = k * (1000.0e-6 * grav - gp)
    for i in range(8):
        tot = voigt - alpha * dpl[i] * np.array([1, 1, 1, 0, 0, 0])
        np.testing.assert_allclose(p3[6 * i:6 * i + 6], tot, atol=1e-10 * np.abs(voigt).max())
        biot = tot + alpha * ul[24 + i] * np.array([1, 1, 1, 0, 0, 0])
        np.testing.assert_allclose(p3[48 + 6 * i:54 + 6 * i], biot, atol=1e-10 * np.abs(voigt).max())
        np.testing.assert_allclose(p3[96 + 3 * i:99 + 3 * i], flux, atol=1e-12 * np.abs(flux).max())
    p7, _ = element(xl, e, 7, env(), ul, dpl)
    np.testing.assert_allclose(p7[:8], alpha * np.trace(G) + ul[24:] / M,
                               atol=1e-12 * (abs(np.trace(G)) + np.abs(ul[24:]).max()))


def face_area(xf):
    return 0.5 * np.linalg.norm(np.cross(xf[2] - xf[0], xf[3] - xf[1]))


@pytest.mark.parametrize("face", range(6))
def test_face_loads(face):
    """Type 40: the nodal forces of a face sum to traction x area, on its 4
    vertices only; type 41: the same along the outward normal; type 43: the
    pressure rows sum to dt x flux x area."""
    rng = np.random.default_rng(face)
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))  # a parallelepiped: planar faces
    assert np.linalg.det(A) > 0
    xl = CORNERS @ A.T + np.array([0.2, 0.1, 0.3])
    load, fload = P.empty_load_table()
    trac = rng.standard_normal(3)
    P.set_traction_load(load, fload, 1, trac)
    load[1] = (41, 2, 2)
    fload[1, 0, :2] = (0.0, 1e30)
    fload[1, 1, :2] = (0.7, 0.7)
    load[2] = (43, 2, 2)
    fload[2, 0, :2] = (0.0, 1e30)
    fload[2, 1, :2] = (0.25, 0.25)
    ev = env(dt=0.07, load=load, fload=fload)
    e = material()
    verts = [v - 1 for v in FACES[face]]
    xf = xl[verts]
    area = face_area(xf)
    nrm = np.cross(xf[1] - xf[0], xf[2] - xf[0])
    nrm /= np.linalg.norm(nrm)
    centre = xl.mean(axis=0)
    if nrm @ (xf.mean(axis=0) - centre) < 0:
        nrm = -nrm  # outward
    base, _ = element(xl, e, 4, ev)
    for nc, kind in ((1, "trac"), (2, "press")):
        iqm = np.zeros(7, np.int32)
        iqm[face] = nc
        p4, _ = element(xl, e, 4, ev, iqm=iqm)
        f = -(p4 - base)[:24].reshape(8, 3)
        expect = trac * area if kind == "trac" else 0.7 * nrm * area
        np.testing.assert_allclose(f.sum(axis=0), expect, atol=1e-12)
        others = [i for i in range(8) if i not in verts]
        assert np.abs(f[others]).max() == 0.0
        np.testing.assert_allclose(f[verts], np.tile(expect / 4, (4, 1)), atol=1e-12)
    iqh = np.zeros(7, np.int32)
    iqh[face] = 3
    p4, _ = element(xl, e, 4, ev, iqh=iqh)
    q = -(p4 - base)[24:]
    np.testing.assert_allclose(q.sum(), 0.07 * 0.25 * area, rtol=1e-13)
    np.testing.assert_allclose(q[verts], 0.07 * 0.25 * area / 4, rtol=1e-13)


def terzaghi_p(depth, t, H, c, p0, nterms=200):
    s = np.zeros_like(depth)
    for m in range(nterms):
        Mm = (2 * m + 1) * np.pi / 2
        s += 2.0 / Mm * np.sin(Mm * depth / H) * np.exp(-Mm * Mm * c * t / (H * H))
    return p0 * s


def test_column_terzaghi():
    """A 1 x 1 x 20 column, E 10 MPa, nu 0.3, M 10 MPa (storage and flow both
    matter), k/mu 0.1 m^2/(MPa s), 1 kPa step load: the first step's pressure
    is the undrained alpha M q / (Kv + alpha^2 M) more than 5 m below the
    drained top (equal-order pressure oscillates within the first elements
    under a drained face at small dt: 1.7 % at 2 m, 0.03 % at 5 m); at
    T = c t / H^2 = 0.05 and 0.2 (backward-Euler steps of T 0.0025) the
    pressure profile follows Terzaghi's series (c = k/mu Kv M / (Kv + M)) to
    1 % and 0.3 % of p0 (measured 0.66 % and 0.14 %); the settlement tends to
    q H / Kv."""
    H, q, E, nu, M, kmu = 20.0, 1.0e-3, 10.0, 0.3, 10.0, 0.1
    Kv = E * (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    c = kmu * Kv * M / (Kv + M)
    p0 = M * q / (Kv + M)
    d = P.column_deck(1, 1, int(H), E=E, nu=nu, k=kmu * 1e-9, mu=1e-9, load_mpa=q, dt=0.5,
                      nsteps=0, nen=8)
    d.e[:, 3] = M
    m = O.OracleModel(d, solvtol=1e-12)
    z = d.x[:, 2]
    depth = H - z
    nv = d.nnodev
    dt_small = 1e-4 * H * H / c
    m.env.dt = dt_small
    m.step()
    deep = depth > 5.0
    assert np.abs(m.u[:nv, 3][deep] / p0 - 1.0).max() < 1e-3
    # the undrained strain -q / (Kv + alpha^2 M) over the 15 m below that layer
    w15 = m.u[np.argmin(np.abs(z - 15.0)), 2]
    assert abs(w15 / (-q * 15.0 / (Kv + M)) - 1.0) < 1e-3
    t = dt_small
    dt = 0.0025 * H * H / c
    m.env.dt = dt
    for T, bar in ((0.05, 0.01), (0.2, 0.003)):
        while t < T * H * H / c - 1e-9:
            m.step()
            t += dt
        ref = terzaghi_p(depth, t, H, c, p0)
        err = np.abs(m.u[:nv, 3] - ref).max() / p0
        assert err < bar, (T, err)
    m.env.dt = 2.0 * H * H / c  # backward Euler leaves ~0.17 of the slowest mode per step
    for _ in range(8):
        m.step()
    wf = m.u[np.argmax(z), 2]
    assert abs(wf / (-q * H / Kv) - 1.0) < 1e-3
