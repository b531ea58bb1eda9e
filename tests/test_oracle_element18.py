"""Pins the oracle's variable-property element elmt18_pm (Elmlib03.f:2640-3640)
and the property laws of Poromec.f.

* bit for bit against the reference's own Fortran, built from its sources
  into oracle/_ref: elmt18_pm isw 2, 3, 4, 7 and 10 with gravity and face
  loads, konzey_carman, hashin_shtrikman, ks_and_mus and
  extrapol_hexa20_v2;
* closed forms that go through neither code: with every point's properties
  equal to a material's drained moduli elmt18 isw 2 is elmt17's element;
  ks_and_mus inverts hashin_shtrikman; Kozeny-Carman returns perm0 at the
  initial porosity; the extrapolation reproduces a trilinear field at the
  vertices; isw 10 on a linear field adds alpha tr(eps) + p / M at every
  point.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

from test_oracle_element import CORNERS, distorted, env, node_coords


def _material(rng):
    e = np.zeros(15)
    E, nu = rng.uniform(5, 50), rng.uniform(0.2, 0.35)
    e[0] = E / (3 * (1 - 2 * nu))        # K (solid, after initial_prop)
    e[1] = E / (2 * (1 + nu))            # mu
    e[2] = 0.1 * rng.uniform(0.5, 2)     # permeability coefficient
    e[3] = rng.uniform(500, 5000)        # Biot modulus
    e[4] = rng.uniform(0.7, 1.0)         # Biot coefficient
    e[5] = 2000.0
    e[6] = 1000.0
    e[7] = rng.uniform(0.2, 0.4)         # initial porosity
    return e


def _vprop(rng, e):
    vp = np.empty((64, 7))
    vp[:, 0] = e[7] * (1 + 0.1 * rng.standard_normal(64))
    vp[:, 1] = e[2] * (1 + 0.2 * rng.standard_normal(64))
    vp[:, 2] = 1.8e-3 * (1 + 0.05 * rng.standard_normal(64))
    vp[:, 3] = e[0] * (1 + 0.1 * rng.standard_normal(64))
    vp[:, 4] = e[1] * (1 + 0.1 * rng.standard_normal(64))
    vp[:, 5] = (1 / e[3]) * (1 + 0.1 * rng.standard_normal(64))
    vp[:, 6] = e[4] * (1 + 0.05 * rng.standard_normal(64))
    return vp


@pytest.mark.parametrize("isw", [2, 3, 4, 7, 10])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_elmt18_bitwise_vs_reference(ref, isw, seed):
    rng = np.random.default_rng(100 + seed)
    xl = distorted(rng)
    e = _material(rng)
    vp = _vprop(rng, e)
    ul = rng.standard_normal(68) * 1e-3
    dpl = rng.standard_normal(8) * 1e-3
    load, fload = P.empty_load_table()
    P.set_traction_load(load, fload, 1, (0.1, -0.2, -1e-3))
    load[1] = (43, 2, 2)
    fload[1, 0, :2] = (0.0, 1e30)
    fload[1, 1, :2] = (0.5, 0.5)
    iqm = np.array([1, 0, 0, 0, 1, 0, 0], np.int32)
    iqh = np.array([0, 0, 1, 0, 0, 0, 0], np.int32)
    grav = (0.1, -0.3, -9.81)
    rc, p, s, vpo = O.elmt18(e, iqm, iqh, xl, ul, dpl, vp, isw,
                             env(0.07, 0.3, grav, 9.81, load, fload))
    assert rc == 0
    pr, sr, vpr = O.ref_elmt18(e, iqm, iqh, xl, ul, dpl, vp, isw, 0.07, 0.3, grav, 9.81, load,
                               fload)
    n = {2: 68, 3: 120, 4: 68, 7: 8, 10: 0}[isw]
    assert np.array_equal(p[:n], pr[:n])
    if isw == 2:
        assert np.array_equal(s, sr)
    assert np.array_equal(vpo, vpr)
    if isw == 10:
        assert not np.array_equal(vpo, vp)


def test_property_laws_bitwise_vs_reference(ref):
    rng = np.random.default_rng(4)
    R = O.load_ref()
    d = C.c_double
    for _ in range(20):
        perm0, poro, poro0 = rng.uniform(1e-4, 1), rng.uniform(0.05, 0.5), rng.uniform(0.05, 0.5)
        assert O.lib.orc_konzey_carman(perm0, poro, poro0) == O.ref_scalar(
            "konzey_carman", perm0, poro, poro0)
        ks, mus = rng.uniform(10, 100), rng.uniform(5, 50)
        mine = [d() for _ in range(4)]
        O.lib.orc_hashin_shtrikman(poro, ks, mus, *[C.byref(v) for v in mine])
        theirs = [d() for _ in range(4)]
        args = [d(poro), d(ks), d(mus)]
        R.hashin_shtrikman_(*[C.byref(v) for v in args + theirs])
        assert [v.value for v in mine] == [v.value for v in theirs]
        # ks_and_mus of the drained moduli that hashin_shtrikman gives back
        k, mu = mine[0].value, mine[1].value
        km, mum = d(k), d(mu)
        assert O.lib.orc_ks_and_mus(0.0, 0.0, poro, C.byref(km), C.byref(mum)) == 0
        kr, mur = d(k), d(mu)
        R.ks_and_mus_(C.byref(d(0.0)), C.byref(d(0.0)), C.byref(d(poro)), C.byref(kr),
                      C.byref(mur))
        assert (km.value, mum.value) == (kr.value, mur.value)
        # ks_and_mus inverts hashin_shtrikman: the solid moduli come back
        assert abs(km.value - ks) <= 1e-10 * ks and abs(mum.value - mus) <= 1e-10 * mus


def test_konzey_carman_at_initial_porosity():
    for perm0, poro0 in ((1e-3, 0.3), (2.5, 0.12)):
        assert abs(O.lib.orc_konzey_carman(perm0, poro0, poro0) - perm0) <= 1e-15 * perm0


def _gauss_points():
    g = np.array([-.861136311594053, -.339981043584856, .339981043584856, .861136311594053])
    pts = []
    for tz in g:
        for sy in g:
            for rx in g:
                pts.append((rx, sy, tz))
    return np.array(pts)


def test_extrapolation_bitwise_and_trilinear(ref):
    R = O.load_ref()
    rng = np.random.default_rng(9)
    rst = _gauss_points()
    coef = rng.standard_normal((8, 3))

    def tri(r, s, t):
        basis = np.array([np.ones_like(r), r, s, t, r * s, s * t, r * t, r * s * t])
        return coef.T @ basis

    up = tri(rst[:, 0], rst[:, 1], rst[:, 2]).T.copy()  # 64 x 3
    ue = np.zeros((8, 3))
    O.lib.orc_extrapol_hexa20(np.ascontiguousarray(up), ue, 3)
    rn = np.array([1, -1, -1, 1, 1, -1, -1, 1, 0, -1, 0, 1, 0, -1, 0, 1, 1, -1, -1, 1], float)
    sn = np.array([1, 1, -1, -1, 1, 1, -1, -1, 1, 0, -1, 0, 1, 0, -1, 0, 1, 1, -1, -1], float)
    tn = np.array([1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0], float)
    uer = np.zeros((8, 3))
    dptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    R.extrapol_hexa20_v2_(dptr(up), dptr(uer), dptr(rn), dptr(sn), dptr(tn),
                          C.byref(C.c_int(3)))
    assert np.array_equal(ue, uer)
    # c2 = 1/(8g) is exact to its 15 digits, but the tabulated c3 and c4
    # differ from 1/(8g^2), 1/(8g^3) (g = 0.861136311594053) by 7e-11 and
    # 1e-10 relative, so the fit reproduces a trilinear field to ~1e-10
    want = tri(rn[:8], sn[:8], tn[:8]).T
    assert np.allclose(ue, want, rtol=0, atol=3e-10 * np.abs(want).max())


def test_elmt18_uniform_properties_is_elmt17():
    """Every point at a material's drained moduli: elmt18 isw 2 and 4 give
    elmt17's element matrix and residual (to rounding: a = K + 4 mu / 3
    there, E (1 - nu) / ((1 + nu)(1 - 2 nu)) here)."""
    rng = np.random.default_rng(12)
    xl = distorted(rng)
    E, nu, kw, M, alpha = 20.0, 0.3, 0.1, 800.0, 0.9
    e17 = P.material_vector([E], nu, kw, M=M, alpha=alpha)[0]
    e17[5] = 2000.0
    e18 = e17.copy()
    e18[0], e18[1] = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    vp = np.zeros((64, 7))
    vp[:, 1] = e17[2]
    vp[:, 2] = e17[5] * 1e-6
    vp[:, 3], vp[:, 4], vp[:, 5], vp[:, 6] = e18[0], e18[1], 1 / M, alpha
    ul = rng.standard_normal(68) * 1e-3
    z7 = np.zeros(7, np.int32)
    ev = env(0.05, 0.2, (0.0, 0.0, -9.81), 9.81)
    for isw in (2, 4):
        _, p17, s17 = O.elmt17(e17, z7, z7, xl, ul, np.zeros(8), isw, ev)
        _, p18, s18, _ = O.elmt18(e18, z7, z7, xl, ul, np.zeros(8), vp, isw, ev)
        assert np.allclose(s18, s17, rtol=0, atol=1e-12 * np.abs(s17).max())
        assert np.allclose(p18[:68], p17[:68], rtol=0, atol=1e-12 * np.abs(p17[:68]).max())


def test_elmt18_isw10_linear_field():
    rng = np.random.default_rng(13)
    xl = distorted(rng)
    e = _material(rng)
    vp = _vprop(rng, e)
    G = rng.standard_normal((3, 3)) * 1e-3
    X = node_coords(xl)
    u = np.zeros(68)
    u[:60] = (X @ G.T).ravel()
    u[60:] = 0.01
    z7 = np.zeros(7, np.int32)
    _, _, _, vpo = O.elmt18(e, z7, z7, xl, u, np.zeros(8), vp, 10, env())
    dphi = e[4] * np.trace(G) + 0.01 / e[3]
    assert np.allclose(vpo[:, 0], vp[:, 0] + dphi, rtol=1e-12, atol=1e-15)
    poro = vpo[:, 0]
    kc = e[2] * poro ** 3 / (1 - poro) ** 2 * (1 - e[7]) ** 2 / e[7] ** 3
    assert np.allclose(vpo[:, 1], kc, rtol=1e-13)
    assert np.allclose(vpo[:, 2], (1 - poro) * e[5] * 1e-6 + poro * e[6] * 1e-6, rtol=1e-14)


def _vdeck(n=2, nz=4, poro=0.3):
    d = P.column_deck(n, n, nz, dt=0.05)
    d.e[:, 7] = poro
    d.e[:, 5] = 2000.0
    return d


@pytest.mark.parametrize("flags", [(1, 0, 0, 0), (1, 1, 1, 1)])
def test_initial_prop_identities(flags):
    """initial_prop: without Hashin-Shtrikman every point carries the
    drained K = E / (3 (1 - 2 nu)), mu = E / (2 (1 + nu)), 1/M and alpha;
    with it, ks_and_mus finds the solid moduli whose bound at the porosity
    is that drained pair.  With the density flag the mixture density of the
    solid density it derives is the input density again."""
    d = _vdeck()
    E, nu = d.e[0, 0], d.e[0, 1]
    m = O.OracleModel(d, ie=[17 + 1] * d.e.shape[0], vprop=flags)
    vp = m.vpropel
    K, mu = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    assert np.allclose(vp[..., 0], 0.3) and np.allclose(vp[..., 1], d.e[0, 2])
    assert np.allclose(vp[..., 2], 2000.0 * 1e-6, rtol=1e-14)
    if flags[3]:
        assert np.allclose(vp[..., 3], K, rtol=1e-10) and np.allclose(vp[..., 4], mu, rtol=1e-10)
        assert m.e[0, 0] > K and m.e[0, 1] > mu  # the solid is stiffer than the drained frame
        assert np.allclose(vp[..., 6], 1 - vp[..., 3] / m.e[0, 0], rtol=1e-14)
    else:
        assert np.allclose(vp[..., 3], K, rtol=1e-14) and np.allclose(vp[..., 4], mu, rtol=1e-14)
        assert np.allclose(vp[..., 5], 1 / d.e[0, 3]) and np.allclose(vp[..., 6], d.e[0, 4])


def test_march_constant_properties_matches_elmt17():
    """vprop alone (no Kozeny-Carman, density or Hashin-Shtrikman updates):
    the points keep the material's moduli and permeability, so the elmt18
    march is the elmt17 march to rounding; only the porosity is carried at
    the integration points."""
    d17 = _vdeck()
    d18 = _vdeck()
    m17 = O.OracleModel(d17, stop_crit=1)
    m18 = O.OracleModel(d18, stop_crit=1, ie=[18], vprop=(1, 0, 0, 0))
    for _ in range(4):
        i17, i18 = m17.step(), m18.step()
        assert i17["nlit"] == i18["nlit"]
        du, dp = np.abs(m17.u[:, :3]).max(), np.abs(m17.u[:, 3]).max()
        assert np.abs(m18.u[:, :3] - m17.u[:, :3]).max() <= 1e-10 * du
        assert np.abs(m18.u[:, 3] - m17.u[:, 3]).max() <= 1e-10 * dp
    assert not np.array_equal(m18.vpropel[..., 0], np.full_like(m18.vpropel[..., 0], 0.3))
    assert np.abs(m18.porosity[: d18.nnodev] - 0.3).max() < 0.05


@pytest.mark.parametrize("flags", [(1, 0, 0, 0), (1, 1, 0, 0), (1, 1, 1, 1)])
def test_update_prop_flags(flags):
    """update_prop keeps every point's porosity and, as the conseq flags
    ask, its Kozeny-Carman permeability, density and Hashin-Shtrikman moduli
    of that porosity; the rest stay at their initial values."""
    d = P.column_deck(2, 2, 4, dt=0.05, load_mpa=0.5)
    d.e[:, 7], d.e[:, 5] = 0.3, 2000.0
    m = O.OracleModel(d, stop_crit=1, ie=[18], vprop=flags)
    vp0 = m.vpropel.copy()
    for _ in range(3):
        m.step()
    vp = m.vpropel
    poro = vp[..., 0]
    assert np.abs(poro - 0.3).max() > 1e-3
    e = m.e[0]
    kc = np.vectorize(O.lib.orc_konzey_carman)(e[2], poro, e[7])
    assert np.array_equal(vp[..., 1], kc) if flags[1] else np.array_equal(vp[..., 1], vp0[..., 1])
    if flags[2]:
        assert np.array_equal(vp[..., 2], (1.0 - poro) * (e[5] * 1e-6) + poro * (e[6] * 1e-6))
    else:
        assert np.array_equal(vp[..., 2], vp0[..., 2])
    if flags[3]:
        assert not np.array_equal(vp[..., 3:], vp0[..., 3:])
    else:
        assert np.array_equal(vp[..., 3:], vp0[..., 3:])
    # the nodal porosity is the points' porosity extrapolated and averaged
    assert np.abs(m.porosity[: d.nnodev] - 0.3).max() > 1e-3
