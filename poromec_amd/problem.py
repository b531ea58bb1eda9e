"""Synthetic input decks for the mef Biot u-p path: hex8 boxes, boundary
conditions, materials and loads, laid out exactly as the reference's deck
reader leaves them in memory.

The reference reads these from a deck (``Rdata.f``: ``elconn`` hexa8 block
``Rdata.f:479-521``, ``coord``, ``constrainpmec`` ``Rdata.f:575-600``,
``materials`` ``Rdata.f:1132-1180``, ``elmtloads``/``loads``
``Rdata.f:1487``) and a gravity file (``Rdata.f:3543``).  This module builds
the same arrays for the box problems the benchmarks use:

* vertices numbered x-fastest, then y, then z; elements likewise;
* local vertex order of ``sfhexa8_m`` (``Elmlib.f:1536-1544``): nodes 1-4 on
  the face t = +1 (top, z = z1), node 1 at (+r, +s), 2 at (-r, +s),
  3 at (-r, -s), 4 at (+r, -s); nodes 5-8 below them.  With r, s, t along
  x, y, z the Jacobian is positive and element face 1 is the top face;
* material vector ``e(1..7)`` of ``elmt17_pm`` (``Elmlib03.f:1833-1840``):
  E [MPa], nu, hydraulic conductivity ``k/mu * rho_f * 1e-6 * g_mod``,
  Biot modulus M [MPa], Biot alpha, rho, rho_f; the element type id (17) is
  carried by ``ie``;
* BCs: roller walls (ux on the x faces, uy on the y faces), fixed base (uz),
  drained top (p = 0);
* top load: a type-40 face load (``tload``, ``Pload.f:330-336``) on face 1;
* element family: ``nen = 20`` reads the mesh as the reference does with
  ``maxno 20`` (promoted to the hexa20/8 element ``elmt17_pm``); ``nen = 8``
  keeps it as read with ``maxno 8``, for the equal-order Q1/Q1 element 7
  (the ``ELMT07_PM`` the reference reserves, ``Elmlib03.f:22-23``), whose
  vertex-only dofs give the benchmark configurations' nominal counts.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

NEN = 20
NENV = 8
NDF = 4
NST = 68
PROP = 15
NUMLOAD = 200
NTERMS = 5
NPAR = 1000
ELEMENT_TYPE = 17


@dataclass
class Deck:
    """One problem: mesh, restraints, materials, loads, time stepping."""

    nx: int
    ny: int
    nz: int
    ix: np.ndarray          # (numel, 21) int32: 1-based vertex ids, zeros 8..19, material id
    x: np.ndarray           # (nnodev, 3) float64
    idv: np.ndarray         # (nnodev, 4) int32 restraint flags (1 = fixed)
    e: np.ndarray           # (numat, 15) float64
    eload: np.ndarray       # (numel, 7) int32 mechanical face-load ids per face
    eloadp: np.ndarray      # (numel, 7) int32 hydraulic face-load ids per face
    load: np.ndarray        # (200, 3) int32   = Fortran load(3,200)
    fload: np.ndarray       # (200, 5, 1000) float64 = Fortran fload(1000,5,200)
    gravity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    gravity_mod: float = 10.0
    dt: float = 0.01
    nsteps: int = 20
    nen: int = 20           # 20: hexa20/8 (elmt17_pm); 8: Q1/Q1 (element type 7)

    @property
    def numel(self) -> int:
        return int(self.ix.shape[0])

    @property
    def nnodev(self) -> int:
        return int(self.x.shape[0])


def box_mesh(nx: int, ny: int, nz: int, h=(1.0, 1.0, 1.0)):
    """hex8 box: returns ix (numel, 21) int32 with the material column set to
    the element number (one material per element) and x (nnodev, 3)."""
    nvx, nvy, nvz = nx + 1, ny + 1, nz + 1
    kx, ky, kz = np.meshgrid(np.arange(nvx), np.arange(nvy), np.arange(nvz), indexing="ij")
    x = np.empty((nvx * nvy * nvz, 3))
    vid = (kx + nvx * (ky + nvy * kz)).ravel(order="F")
    x[vid, 0] = kx.ravel(order="F") * h[0]
    x[vid, 1] = ky.ravel(order="F") * h[1]
    x[vid, 2] = kz.ravel(order="F") * h[2]

    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex = ex.ravel(order="F")
    ey = ey.ravel(order="F")
    ez = ez.ravel(order="F")
    eid = ex + nx * (ey + ny * ez)

    def v(dx, dy, dz):
        return 1 + (ex + dx) + nvx * ((ey + dy) + nvy * (ez + dz))

    ix = np.zeros((nx * ny * nz, NEN + 1), dtype=np.int32)
    corners = [(1, 1, 1), (0, 1, 1), (0, 0, 1), (1, 0, 1),
               (1, 1, 0), (0, 1, 0), (0, 0, 0), (1, 0, 0)]
    for k, (dx, dy, dz) in enumerate(corners):
        ix[eid, k] = v(dx, dy, dz)
    ix[eid, NEN] = eid + 1
    return ix, x


def box_restraints(nx: int, ny: int, nz: int) -> np.ndarray:
    """Vertex restraint flags: rollers on the side walls, fixed base, drained top."""
    nvx, nvy, nvz = nx + 1, ny + 1, nz + 1
    kx, ky, kz = np.meshgrid(np.arange(nvx), np.arange(nvy), np.arange(nvz), indexing="ij")
    vid = (kx + nvx * (ky + nvy * kz)).ravel(order="F")
    kx, ky, kz = kx.ravel(order="F"), ky.ravel(order="F"), kz.ravel(order="F")
    idv = np.zeros((nvx * nvy * nvz, NDF), dtype=np.int32)
    idv[vid, 0] = ((kx == 0) | (kx == nx)).astype(np.int32)
    idv[vid, 1] = ((ky == 0) | (ky == ny)).astype(np.int32)
    idv[vid, 2] = (kz == 0).astype(np.int32)
    idv[vid, 3] = (kz == nz).astype(np.int32)
    return idv


def material_vector(E, nu, k_over_mu, M=1000.0, alpha=1.0, rho=0.0, rho_f=1000.0,
                    gravity_mod=10.0) -> np.ndarray:
    """Rows of e(1..15) for elmt17_pm; k_over_mu in m^2/(MPa s).

    ``perm = e(3) / (rho_f * 1e-6 * g_mod)`` (Elmlib03.f:1961-1969) so
    ``e(3) = k/mu * rho_f * 1e-6 * g_mod`` gives perm = k/mu."""
    E = np.atleast_1d(np.asarray(E, dtype=np.float64))
    n = E.shape[0]
    e = np.zeros((n, PROP))
    e[:, 0] = E
    e[:, 1] = np.broadcast_to(nu, (n,))
    e[:, 2] = np.broadcast_to(k_over_mu, (n,)) * rho_f * 1.0e-6 * gravity_mod
    e[:, 3] = M
    e[:, 4] = alpha
    e[:, 5] = rho
    e[:, 6] = rho_f
    return e


def empty_load_table():
    return np.zeros((NUMLOAD, 3), dtype=np.int32), np.zeros((NUMLOAD, NTERMS, NPAR))


def set_traction_load(load, fload, nc: int, traction, times=(0.0, 1.0e30)):
    """Type-40 load ``nc`` (1-based): a constant traction vector [MPa] in time."""
    load[nc - 1, 0] = 40
    load[nc - 1, 1] = len(times)
    load[nc - 1, 2] = len(times)
    for i, t in enumerate(times):
        fload[nc - 1, 0, i] = t
        for d in range(3):
            fload[nc - 1, 1 + d, i] = traction[d]


def gaussian_field(shape, corr_len: float, rng: np.random.Generator) -> np.ndarray:
    """Zero-mean, unit-variance Gaussian random field on a cell grid with a
    squared-exponential covariance exp(-r^2 / corr_len^2): white noise on a
    grid padded by 3 correlation lengths, smoothed by the matching Gaussian
    kernel (std corr_len / sqrt(2)) through FFTs, cropped and normalised."""
    pad = int(np.ceil(3 * corr_len))
    big = tuple(s + 2 * pad for s in shape)
    noise = rng.standard_normal(big)
    sig = corr_len / np.sqrt(2.0)
    kern = np.ones(big)
    for ax, n in enumerate(big):
        f = np.fft.fftfreq(n)
        g = np.exp(-2.0 * (np.pi * f * sig) ** 2)
        sh = [1] * len(big)
        sh[ax] = n
        kern = kern * g.reshape(sh)
    field_ = np.real(np.fft.ifftn(np.fft.fftn(noise) * kern))
    sl = tuple(slice(pad, pad + s) for s in shape)
    field_ = field_[sl]
    field_ = field_ - field_.mean()
    sd = field_.std()
    return field_ / sd if sd > 0 else field_


def column_deck(nx=4, ny=4, nz=40, E=10.0, nu=0.3, k=1.0e-10, mu=1.0e-9, load_mpa=1.0e-3,
                dt=0.01, nsteps=20, nen=20) -> Deck:
    """1-D consolidation column: homogeneous material, 1 kPa step load on the
    top, drained top, roller walls, fixed base (one material, ``numat = 1``)."""
    ix, x = box_mesh(nx, ny, nz)
    ix[:, NEN] = 1
    idv = box_restraints(nx, ny, nz)
    e = material_vector([E], nu, k / mu)
    load, fload = empty_load_table()
    set_traction_load(load, fload, 1, (0.0, 0.0, -load_mpa))
    eload = np.zeros((ix.shape[0], 7), dtype=np.int32)
    top = np.arange(ix.shape[0]) >= nx * ny * (nz - 1)
    eload[top, 0] = 1
    eloadp = np.zeros_like(eload)
    return Deck(nx, ny, nz, ix, x, idv, e, eload, eloadp, load, fload, dt=dt, nsteps=nsteps,
                nen=nen)


def heterogeneous_box_deck(n, seed=0, sigma_lnk=1.0, mean_lnk=np.log(1.0e-12), corr_len=8.0,
                           E_range=(5.0, 50.0), nu_range=(0.25, 0.35), mu=1.0e-9,
                           load_mpa=1.0e-3, patch_fraction=None, dt=0.1, nsteps=100,
                           shape=None, nen=20) -> Deck:
    """Box of n^3 (or ``shape``) hex8 cells, one material per element: log-k
    from a seeded Gaussian random field, E and nu uniform per element.  With
    ``patch_fraction`` the top load covers a Bernoulli mask of the top faces."""
    nx, ny, nz = shape if shape is not None else (n, n, n)
    rng = np.random.default_rng(seed)
    ix, x = box_mesh(nx, ny, nz)
    idv = box_restraints(nx, ny, nz)
    lnk = mean_lnk + sigma_lnk * gaussian_field((nx, ny, nz), corr_len, rng)
    kcell = np.exp(lnk).ravel(order="F")
    E = rng.uniform(*E_range, size=nx * ny * nz)
    nu = rng.uniform(*nu_range, size=nx * ny * nz)
    e = material_vector(E, nu, kcell / mu)
    load, fload = empty_load_table()
    set_traction_load(load, fload, 1, (0.0, 0.0, -load_mpa))
    eload = np.zeros((ix.shape[0], 7), dtype=np.int32)
    top = np.arange(ix.shape[0]) >= nx * ny * (nz - 1)
    if patch_fraction is not None:
        mask = rng.random(nx * ny) < patch_fraction
        top_ids = np.nonzero(top)[0]
        top = np.zeros_like(top)
        top[top_ids[mask]] = True
    eload[top, 0] = 1
    eloadp = np.zeros_like(eload)
    return Deck(nx, ny, nz, ix, x, idv, e, eload, eloadp, load, fload, dt=dt, nsteps=nsteps,
                nen=nen)


def box_counts(nx, ny, nz, nen=20):
    """Structural counts of a box before boundary conditions: promoted
    hexa20/8 (nen 20) or Q1/Q1 (nen 8, 4 dofs at every vertex)."""
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    if nen == 8:
        return {"numel": nx * ny * nz, "nnodev": nv, "nnode": nv, "dofs": 4 * nv}
    nn = nv + nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    return {"numel": nx * ny * nz, "nnodev": nv, "nnode": nn, "dofs": 3 * nn + nv}


def concat_decks(decks):
    """One deck holding independent boxes side by side (node ids, material ids
    offset per box), for the batching form.  All decks must share dt, gravity
    and the load table.  Returns (deck, elem_off) with box b owning elements
    elem_off[b]:elem_off[b+1]."""
    first = decks[0]
    ix, x, idv, e, el, elp = [], [], [], [], [], []
    node_off = mat_off = 0
    elem_off = [0]
    for d in decks:
        assert d.dt == first.dt and np.array_equal(d.load, first.load) and d.nen == first.nen
        a = d.ix.copy()
        a[:, :8] += node_off
        a[:, NEN] += mat_off
        ix.append(a)
        x.append(d.x)
        idv.append(d.idv)
        e.append(d.e)
        el.append(d.eload)
        elp.append(d.eloadp)
        node_off += d.nnodev
        mat_off += d.e.shape[0]
        elem_off.append(elem_off[-1] + d.numel)
    deck = Deck(0, 0, 0, np.vstack(ix), np.vstack(x), np.vstack(idv), np.vstack(e),
                np.vstack(el), np.vstack(elp), first.load, first.fload, first.gravity,
                first.gravity_mod, first.dt, first.nsteps, first.nen)
    return deck, np.array(elem_off, dtype=np.int32)


def batch_decks(nbox=256, seed=0, nmin=12, nmax=32, dt=0.1, nsteps=50, nen=20):
    """Independent boxes of n^3 cells, n ~ U[nmin, nmax] per box, each with its
    own log-normal permeability (sigma ~ U[0.5, 1.5]) and E ~ U[5, 50] MPa."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=nbox)
    sigmas = rng.uniform(0.5, 1.5, size=nbox)
    seeds = rng.integers(0, 2**31 - 1, size=nbox)
    decks = [heterogeneous_box_deck(int(n), seed=int(sd), sigma_lnk=float(sg), dt=dt,
                                    nsteps=nsteps, corr_len=min(8.0, n / 2), nen=nen)
             for n, sg, sd in zip(ns, sigmas, seeds)]
    return decks
