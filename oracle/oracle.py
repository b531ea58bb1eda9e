"""ctypes wrapper of libpmoracle.so (the C restatement) and of the
reference-subset library _ref/libpmref.so (the reference's own Fortran
element, promotion, numbering and assembly routines, built from the
reference sources when they are present).

TEST INFRASTRUCTURE ONLY -- see the package docstring.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
NEN, NENV, NDF, NST, PROP = 20, 8, 4, 68, 15

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OracleEnv(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("t", C.c_double),
        ("gravity", C.c_double * 3),
        ("gravity_mod", C.c_double),
        ("load", C.c_void_p),
        ("fload", C.c_void_p),
    ]


class _Model(C.Structure):
    _fields_ = [
        ("numel", C.c_int), ("nnode", C.c_int), ("nnodev", C.c_int), ("neq", C.c_int),
        ("nad", C.c_int64),
        ("ix", C.c_void_p), ("x", C.c_void_p), ("id", C.c_void_p), ("fnno", C.c_void_p),
        ("e", C.c_void_p), ("eload", C.c_void_p), ("eloadp", C.c_void_p),
        ("nload", C.c_void_p), ("f", C.c_void_p),
        ("ia", C.c_void_p), ("ja", C.c_void_p), ("ad", C.c_void_p), ("al", C.c_void_p),
        ("u", C.c_void_p), ("u0", C.c_void_p), ("dp", C.c_void_p), ("porosity", C.c_void_p),
        ("b0", C.c_void_p), ("b", C.c_void_p), ("x0", C.c_void_p), ("m", C.c_void_p),
        ("ie", C.c_void_p), ("vprop", C.c_int * 4), ("vpropel", C.c_void_p),
        ("nen", C.c_int),
    ]


class _Cfg(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("precond", C.c_int), ("maxit", C.c_int), ("ngram", C.c_int),
        ("maxnlit", C.c_int), ("stop_crit", C.c_int), ("maxnlit_accept", C.c_int),
        ("solvtol", C.c_double),
        ("tol", C.c_double),
    ]


class _StepInfo(C.Structure):
    _fields_ = [
        ("nlit", C.c_int), ("nsolve", C.c_int), ("iters", C.c_int * 16),
        ("used", C.c_int * 16), ("resid", C.c_double * 16),
    ]


def _load_lib():
    path = HERE / "libpmoracle.so"
    if not path.exists():
        raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
    L = C.CDLL(str(path))
    L.orc_elmt17.restype = C.c_int
    L.orc_elmt17.argtypes = [_f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int,
                             C.POINTER(OracleEnv)]
    L.orc_elmt07.restype = C.c_int
    L.orc_elmt07.argtypes = [_f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int,
                             C.POINTER(OracleEnv)]
    L.orc_promote_hexa20.restype = C.c_int
    L.orc_promote_hexa20.argtypes = [_i32p, C.c_int, C.c_int]
    L.orc_bound_quad.argtypes = [_i32p, _i32p, C.c_int]
    L.orc_graph_count.restype = C.c_int64
    L.orc_graph_count.argtypes = [_i32p, C.c_int, C.c_int]
    L.orc_graph.argtypes = [_i32p, C.c_int, C.c_int, _i64p, _i32p, _i32p]
    L.orc_genrcm.argtypes = [C.c_int, _i64p, _i32p, _i32p]
    L.orc_numeq.restype = C.c_int
    L.orc_numeq.argtypes = [_i32p, _i32p, _i32p, C.c_int]
    L.orc_csrc_count.restype = C.c_int64
    L.orc_csrc_count.argtypes = [_i32p, _i32p, _i64p, _i32p, C.c_int, C.c_int, _i64p]
    L.orc_csrc_fill.argtypes = [_i32p, _i32p, _i64p, _i32p, C.c_int, C.c_int, _i64p, _i32p]
    L.orc_matvec_csrc.argtypes = [C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p]
    L.orc_dot.restype = C.c_double
    L.orc_dot.argtypes = [_f64p, _f64p, C.c_int]
    L.orc_precond.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
    sig = [C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_pcg.restype = C.c_int
    L.orc_pcg.argtypes = sig + [C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.orc_rpsqmr.restype = C.c_int
    L.orc_rpsqmr.argtypes = sig + [C.c_double, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]
    L.orc_pbicgstab.restype = C.c_int
    L.orc_pbicgstab.argtypes = sig + [C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.orc_pminres.restype = C.c_int
    L.orc_pminres.argtypes = sig + [C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.orc_gmres2.restype = C.c_int
    L.orc_gmres2.argtypes = sig + [C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double)]
    L.orc_pform.restype = C.c_int
    L.orc_pform.argtypes = [C.POINTER(_Model), C.c_int, C.c_int, C.c_int, _f64p,
                            C.POINTER(OracleEnv)]
    L.orc_step.restype = C.c_int
    L.orc_step.argtypes = [C.POINTER(_Model), C.POINTER(_Cfg), C.POINTER(OracleEnv),
                           C.POINTER(_StepInfo)]
    L.orc_solv.restype = C.c_int
    L.orc_solv.argtypes = [C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p,
                           C.POINTER(_Cfg), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.omp_colour.restype = C.c_int
    L.omp_colour.argtypes = [C.POINTER(_Model), _i32p, C.c_int]
    L.omp_pform_sample.restype = C.c_double
    L.omp_pform_sample.argtypes = [C.POINTER(_Model), C.c_int, C.c_int, C.c_int, _f64p,
                                   C.POINTER(OracleEnv), _i32p, C.c_int, C.c_int64, C.c_int,
                                   C.POINTER(C.c_int64)]
    L.omp_rpsqmr.restype = C.c_int
    L.omp_rpsqmr.argtypes = sig + [C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]
    L.orc_residuo.restype = C.c_int
    L.orc_residuo.argtypes = [C.POINTER(_Model), _f64p, C.c_int, C.c_double, _f64p, C.c_int]
    L.orc_initial_porosity.argtypes = [C.POINTER(_Model)]
    L.orc_initial_prop.restype = C.c_int
    L.orc_initial_prop.argtypes = [C.POINTER(_Model), _f64p, C.c_int]
    L.orc_update_prop.restype = C.c_int
    L.orc_update_prop.argtypes = [C.POINTER(_Model), C.POINTER(OracleEnv)]
    L.orc_elmt18.restype = C.c_int
    L.orc_elmt18.argtypes = [_f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                             C.c_int, C.POINTER(OracleEnv)]
    L.orc_konzey_carman.restype = C.c_double
    L.orc_konzey_carman.argtypes = [C.c_double, C.c_double, C.c_double]
    dp_ = C.POINTER(C.c_double)
    L.orc_hashin_shtrikman.argtypes = [C.c_double, C.c_double, C.c_double, dp_, dp_, dp_, dp_]
    L.orc_ks_and_mus.restype = C.c_int
    L.orc_ks_and_mus.argtypes = [C.c_double, C.c_double, C.c_double, dp_, dp_]
    L.orc_extrapol_hexa20.argtypes = [_f64p, _f64p, C.c_int]
    L.orc_tform.restype = C.c_int
    L.orc_tform.argtypes = [C.POINTER(_Model), _f64p, _f64p, _f64p, _f64p, C.POINTER(OracleEnv)]
    L.orc_tload.restype = C.c_double
    L.orc_tload.argtypes = [C.c_int, C.c_double, _f64p, C.c_double, _f64p,
                            C.POINTER(OracleEnv)]
    return L


lib = _load_lib()


def make_env(load, fload, dt, t=0.0, gravity=(0.0, 0.0, 0.0), gravity_mod=10.0):
    load = np.ascontiguousarray(load, dtype=np.int32)
    fload = np.ascontiguousarray(fload, dtype=np.float64)
    env = OracleEnv()
    env.dt = dt
    env.t = t
    env.gravity[:] = list(gravity)
    env.gravity_mod = gravity_mod
    env.load = load.ctypes.data
    env.fload = fload.ctypes.data
    env._keep = (load, fload)
    return env


def elmt17(e, iqm, iqh, xl, ul, dpl, isw, env, p=None):
    """One elmt17_pm call; returns (rc, p[68], s[68,68] as s[i, j])."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    iqm = np.ascontiguousarray(iqm, dtype=np.int32)
    iqh = np.ascontiguousarray(iqh, dtype=np.int32)
    xl = np.ascontiguousarray(xl, dtype=np.float64).reshape(-1)
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    dpl = np.ascontiguousarray(dpl, dtype=np.float64)
    p = np.zeros(NST) if p is None else np.ascontiguousarray(p, dtype=np.float64).copy()
    s = np.zeros(NST * NST)
    rc = lib.orc_elmt17(e, iqm, iqh, xl, ul, dpl, p, s, isw, C.byref(env))
    return rc, p, s.reshape(NST, NST).T  # s stored column-major


NST7 = 32


def elmt07(e, iqm, iqh, xl, ul, dpl, isw, env, p=None):
    """One call of the Q1/Q1 element (element type 7, nst 32); returns
    (rc, p[120 for isw 3, else 32], s[32,32] as s[i, j])."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    iqm = np.ascontiguousarray(iqm, dtype=np.int32)
    iqh = np.ascontiguousarray(iqh, dtype=np.int32)
    xl = np.ascontiguousarray(xl, dtype=np.float64).reshape(-1)
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    dpl = np.ascontiguousarray(dpl, dtype=np.float64)
    n = 120 if isw == 3 else NST7
    p = np.zeros(n) if p is None else np.ascontiguousarray(p, dtype=np.float64).copy()
    s = np.zeros(NST7 * NST7)
    rc = lib.orc_elmt07(e, iqm, iqh, xl, ul, dpl, p, s, isw, C.byref(env))
    return rc, p, s.reshape(NST7, NST7).T  # s stored column-major


def elmt18(e, iqm, iqh, xl, ul, dpl, vp, isw, env, p=None):
    """One elmt18_pm call; returns (rc, p, s[68,68] as s[i, j], vp after)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    iqm = np.ascontiguousarray(iqm, dtype=np.int32)
    iqh = np.ascontiguousarray(iqh, dtype=np.int32)
    xl = np.ascontiguousarray(xl, dtype=np.float64).reshape(-1)
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    dpl = np.ascontiguousarray(dpl, dtype=np.float64)
    vp = np.ascontiguousarray(vp, dtype=np.float64).reshape(-1).copy()
    p = np.zeros(max(NST, 120)) if p is None else np.ascontiguousarray(p, np.float64).copy()
    s = np.zeros(NST * NST)
    rc = lib.orc_elmt18(e, iqm, iqh, xl, ul, dpl, p, s, vp, isw, C.byref(env))
    return rc, p, s.reshape(NST, NST).T, vp.reshape(64, 7)


def build_structure(deck, reorder=True):
    """Runs promotion, mk_bound_quad, graph, RCM, numbering and the CSRC
    pattern exactly as the mesh macro does (a deck with nen = 8 -- maxno 8,
    the Q1/Q1 element -- skips promotion and mid-edge restraints: every node
    is a vertex).  Returns a dict of arrays."""
    numel, nnodev = deck.numel, deck.nnodev
    nen = getattr(deck, "nen", NEN)
    ix = np.ascontiguousarray(deck.ix, dtype=np.int32).copy()
    if nen == 8:
        ix[:, 8:NEN] = 0
        nnode = nnodev
    else:
        nnode = lib.orc_promote_hexa20(ix, numel, nnodev)
    idf = np.zeros((nnode, NDF), dtype=np.int32)
    idf[:nnodev] = deck.idv
    if nen != 8:
        lib.orc_bound_quad(idf, ix, numel)
    fnno = np.zeros(nnode, dtype=np.int32)
    fnno[:nnodev] = 1
    ipos = lib.orc_graph_count(ix, numel, nnode)
    ip = np.zeros(nnode + 1, dtype=np.int64)
    ips = np.zeros(max(ipos, 1), dtype=np.int32)
    work = np.zeros(max(ipos, 1), dtype=np.int32)
    lib.orc_graph(ix, numel, nnode, ip, ips, work)
    ips = ips[: ip[-1] - 1].copy()
    if reorder:
        perm = np.zeros(nnode, dtype=np.int32)
        lib.orc_genrcm(nnode, ip, ips, perm)
    else:
        perm = np.arange(1, nnode + 1, dtype=np.int32)
    idflags = idf.copy()
    neq = lib.orc_numeq(idf, perm, fnno, nnode)
    ia = np.zeros(neq + 1, dtype=np.int64)
    nad = lib.orc_csrc_count(idf, perm, ip, ips, nnode, neq, ia)
    ja = np.zeros(max(nad, 1), dtype=np.int32)
    lib.orc_csrc_fill(idf, perm, ip, ips, nnode, neq, ia, ja)
    return dict(ix=ix, nnode=nnode, nnodev=nnodev, numel=numel, idflags=idflags, id=idf,
                fnno=fnno, ip=ip, ips=ips, perm=perm, neq=neq, nad=int(nad), ia=ia,
                ja=ja[:nad], nen=nen)


class OracleModel:
    """All state of the oracle's time march for one deck."""

    def __init__(self, deck, st=None, solver=11, precond=2, solvtol=1e-10, maxit=50000,
                 ngram=150, maxnlit=500, tol=1e-4, stop_crit=1, maxnlit_accept=False,
                 ie=None, vprop=None):
        self.deck = deck
        self.st = st if st is not None else build_structure(deck)
        st = self.st
        nnode, neq, nad = st["nnode"], st["neq"], st["nad"]
        self.e = np.array(deck.e, dtype=np.float64)  # initial_prop rewrites it
        self.eload = np.ascontiguousarray(deck.eload, dtype=np.int32)
        self.eloadp = np.ascontiguousarray(deck.eloadp, dtype=np.int32)
        self.x = np.zeros((nnode, 3))
        self.x[: deck.nnodev] = deck.x
        self.nload = np.zeros((nnode, NDF), dtype=np.int32)
        self.f = np.zeros((nnode, NDF))
        self.ad = np.zeros(neq)
        self.al = np.zeros(max(nad, 1))
        self.u = np.zeros((nnode, NDF))
        self.u0 = np.zeros((nnode, NDF))
        self.dp = np.zeros(nnode)
        self.porosity = np.zeros(nnode)
        self.b0 = np.zeros(neq)
        self.b = np.zeros(neq)
        self.x0 = np.zeros(neq)
        self.m = np.zeros(neq)
        self.env = make_env(deck.load, deck.fload, deck.dt, 0.0, deck.gravity, deck.gravity_mod)
        self.cfg = _Cfg(solver, precond, maxit, ngram, maxnlit, stop_crit, int(maxnlit_accept),
                        solvtol, tol)
        self._md = _Model()
        md = self._md
        md.numel, md.nnode, md.nnodev, md.neq, md.nad = (st["numel"], nnode, deck.nnodev, neq,
                                                         nad)
        md.nen = st.get("nen", NEN)
        for name, arr in (("ix", st["ix"]), ("x", self.x), ("id", st["id"]), ("fnno", st["fnno"]),
                          ("e", self.e), ("eload", self.eload), ("eloadp", self.eloadp),
                          ("nload", self.nload), ("f", self.f), ("ia", st["ia"]),
                          ("ja", st["ja"]), ("ad", self.ad), ("al", self.al), ("u", self.u),
                          ("u0", self.u0), ("dp", self.dp), ("porosity", self.porosity),
                          ("b0", self.b0), ("b", self.b), ("x0", self.x0), ("m", self.m)):
            setattr(md, name, arr.ctypes.data)
        lib.orc_initial_porosity(C.byref(md))  # as the mesh macro does
        # elmt18_pm materials (ie = 18) and the conseq flags: initial_prop
        # rewrites e and sets every integration point's properties
        self.vpropel = np.zeros((st["numel"], 64, 7))
        md.vpropel = self.vpropel.ctypes.data
        if ie is not None:
            self.ie = np.ascontiguousarray(ie, dtype=np.int32)
            md.ie = self.ie.ctypes.data
        if vprop is not None:
            vp = [int(bool(v)) for v in vprop]
            if any(vp[1:]):
                vp[0] = 1
            md.vprop[:] = vp
            if vp[0] and lib.orc_initial_prop(C.byref(md), self.e, self.e.shape[0]):
                raise RuntimeError("initial_prop: ks_and_mus has no solution")
        self.t = 0.0

    def assemble(self, isw=2, lhs=True, rhs=True, bvec=None):
        bvec = self.b if bvec is None else bvec
        rc = lib.orc_pform(C.byref(self._md), isw, int(lhs), int(rhs), bvec, C.byref(self.env))
        if rc:
            raise RuntimeError(f"orc_pform failed: {rc}")
        return bvec

    def tform(self):
        """Vertex stresses and flux of the current state (the pnsf macro):
        total, Terzaghi (total + p), Biot (total + alpha p) as (nnode, 6) and
        the Darcy flux (nnode, 3)."""
        n = self.st["nnode"]
        tx, txe, txb, flux = np.zeros((n, 6)), np.zeros((n, 6)), np.zeros((n, 6)), np.zeros((n, 3))
        rc = lib.orc_tform(C.byref(self._md), tx, txe, txb, flux, C.byref(self.env))
        if rc:
            raise RuntimeError(f"orc_tform failed: {rc}")
        return tx, txe, txb, flux

    def step(self, check=True):
        """One time step (orc_step).  check=False returns the step's return
        code in the result (0, or -10 when maxnlit is reached) instead of
        raising on a nonzero one."""
        info = _StepInfo()
        rc = lib.orc_step(C.byref(self._md), C.byref(self.cfg), C.byref(self.env), C.byref(info))
        if rc and check:
            raise RuntimeError(f"orc_step failed: {rc}")
        self.t = self.env.t
        n = min(info.nsolve, 16)
        return dict(nlit=info.nlit, nsolve=info.nsolve, iters=list(info.iters[:n]),
                    used=list(info.used[:n]), resid=list(info.resid[:n]), rc=rc)

    def colour(self, maxcolors=200):
        """Greedy element colouring (omp_colour): (ncolors, colour of each
        element, 1-based)."""
        col = np.zeros(self.st["numel"], np.int32)
        nc = lib.omp_colour(C.byref(self._md), col, maxcolors)
        return nc, col

    def pform_colour(self, col, colour, isw=2, lhs=True, rhs=True, bvec=None, max_elems=None,
                     nthreads=1):
        """The OpenMP coloured element loop over (up to max_elems) elements of
        one colour, adding into ad/al/bvec without zeroing them first.
        Returns (wall seconds, elements done)."""
        bvec = self.b if bvec is None else bvec
        n = C.c_int64()
        lim = self.st["numel"] if max_elems is None else max_elems
        t = lib.omp_pform_sample(C.byref(self._md), isw, int(lhs), int(rhs), bvec,
                                 C.byref(self.env), col, colour, lim, nthreads, C.byref(n))
        return t, n.value

    def residuo(self, b, nli, tol, state, crit):
        """One cal_residuo_pm call (orc_residuo) on the vector b; state is the
        float64[4] carried between the calls of one step.  Returns istop."""
        return lib.orc_residuo(C.byref(self._md), np.ascontiguousarray(b, np.float64), nli, tol,
                               state, crit)


def omp_sqmr(st, ad, al, m, b, tol=1e-10, maxit=50000, nthreads=1):
    """rpsqmr_omp (the reference's OpenMP SQMR) on nthreads threads.
    Returns (return value, x, fail flag); the return value is the iteration
    count, or -2 at maxit (fail = 1, the GMRES fallback's trigger)."""
    neq = st["neq"]
    x = np.zeros(neq)
    res, fail = C.c_double(), C.c_int()
    it = lib.omp_rpsqmr(neq, st["ia"], st["ja"], np.ascontiguousarray(ad, np.float64),
                        np.ascontiguousarray(al, np.float64), np.ascontiguousarray(m, np.float64),
                        np.ascontiguousarray(b, np.float64), x, tol, maxit, nthreads,
                        C.byref(res), C.byref(fail))
    return it, x, fail.value


def solv(st, ad, al, b, solver=11, precond=2, tol=1e-10, maxit=50000, ngram=150):
    """solv_pm's dispatch (orc_solv) on a CSRC system: the preconditioner, the
    solver and, for SQMR, the GMRES fallback.  Returns (return value, x,
    solver that finished)."""
    neq = st["neq"]
    m = np.zeros(neq)
    x = np.zeros(neq)
    cfg = _Cfg(solver, precond, maxit, ngram, 500, 1, 0, tol, 1e-4)
    used, res = C.c_int(), C.c_double()
    it = lib.orc_solv(neq, st["ia"], st["ja"], np.ascontiguousarray(ad, np.float64),
                      np.ascontiguousarray(al, np.float64), m,
                      np.ascontiguousarray(b, np.float64), x, C.byref(cfg), C.byref(used),
                      C.byref(res))
    return it, x, used.value


# ---------------------------------------------------------------------------
# reference-subset library (built from the reference's own sources)
# ---------------------------------------------------------------------------
_ref = None


def load_ref():
    """The reference's element / promotion / numbering / assembly routines
    compiled from /root/reference by ``make -C oracle ref``; None if absent."""
    global _ref
    if _ref is not None:
        return _ref
    path = HERE / "_ref" / "libpmref.so"
    if not path.exists():
        return None
    # lazy binding: the library leaves symbols of files that need MPI
    # unresolved (stop_mef, vet, ...); none is on the paths called here
    _ref = C.CDLL(str(path), mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _ref


def _common(R, name, ctype):
    return ctype.in_dll(R, name)


def ref_set_common(R, dt, t, gravity, gravity_mod, load, fload):
    transi = (C.c_double * 5).in_dll(R, "transi_")   # alfa, beta, dt, t, dt_mec
    transi[2] = dt
    transi[3] = t
    mec = (C.c_double * 4).in_dll(R, "mecprop_")     # gravity(3), gravity_mod
    mec[0], mec[1], mec[2] = gravity
    mec[3] = gravity_mod
    nload = 3 * 200
    ld = (C.c_int * (nload + 1000 * 5 * 200 * 2)).in_dll(R, "load_")
    buf = np.frombuffer(ld, dtype=np.uint8)
    buf[: nload * 4] = np.ascontiguousarray(load, dtype=np.int32).view(np.uint8).ravel()
    buf[nload * 4: nload * 4 + 1000 * 5 * 200 * 8] = np.ascontiguousarray(
        fload, dtype=np.float64).view(np.uint8).ravel()


def ref_elmt18(e, iqm, iqh, xl, ul, dpl, vp, isw, dt, t=0.0, gravity=(0.0, 0.0, 0.0),
               gravity_mod=10.0, load=None, fload=None):
    """Calls the reference's elmt18_pm directly.  Returns (p, s[i, j], vp)."""
    R = load_ref()
    if R is None:
        raise RuntimeError("reference-subset library not built")
    if load is None:
        load = np.zeros((200, 3), dtype=np.int32)
        fload = np.zeros((200, 5, 1000))
    ref_set_common(R, dt, t, gravity, gravity_mod, load, fload)
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    iqm = np.ascontiguousarray(iqm, dtype=np.int32).copy()
    iqh = np.ascontiguousarray(iqh, dtype=np.int32).copy()
    xl = np.ascontiguousarray(xl, dtype=np.float64).reshape(-1).copy()
    ul = np.ascontiguousarray(ul, dtype=np.float64).copy()
    dpl = np.ascontiguousarray(dpl, dtype=np.float64).copy()
    vp = np.ascontiguousarray(vp, dtype=np.float64).reshape(-1).copy()
    p = np.zeros(max(NST, 120))
    s = np.zeros(NST * NST)
    txn = np.zeros(6 * NEN)
    ndm, nst, nel, isw_, block_pu = C.c_int(3), C.c_int(NST), C.c_int(1), C.c_int(isw), C.c_int(0)
    dptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    R.elmt18_pm_(dptr(e), dptr(iqm), dptr(iqh), dptr(xl), dptr(ul), dptr(dpl), dptr(p), dptr(s),
                 dptr(txn), dptr(vp), C.byref(ndm), C.byref(nst), C.byref(nel), C.byref(isw_),
                 C.byref(block_pu))
    return p, s.reshape(NST, NST).T, vp.reshape(64, 7)


def ref_scalar(name, *args):
    """Calls a reference real*8 function of real*8 arguments."""
    R = load_ref()
    f = getattr(R, name + "_")
    f.restype = C.c_double
    vals = [C.c_double(a) for a in args]
    return f(*[C.byref(v) for v in vals])


def ref_elmt17(e, iqm, iqh, xl, ul, dpl, isw, dt, t=0.0, gravity=(0.0, 0.0, 0.0),
               gravity_mod=10.0, load=None, fload=None, p=None):
    """Calls the reference's elmt17_pm directly.  Returns (p, s[i, j])."""
    R = load_ref()
    if R is None:
        raise RuntimeError("reference-subset library not built")
    if load is None:
        load = np.zeros((200, 3), dtype=np.int32)
        fload = np.zeros((200, 5, 1000))
    ref_set_common(R, dt, t, gravity, gravity_mod, load, fload)
    e = np.ascontiguousarray(e, dtype=np.float64).copy()
    iqm = np.ascontiguousarray(iqm, dtype=np.int32).copy()
    iqh = np.ascontiguousarray(iqh, dtype=np.int32).copy()
    xl = np.ascontiguousarray(xl, dtype=np.float64).reshape(-1).copy()
    ul = np.ascontiguousarray(ul, dtype=np.float64).copy()
    dpl = np.ascontiguousarray(dpl, dtype=np.float64).copy()
    p = np.zeros(NST) if p is None else np.ascontiguousarray(p, dtype=np.float64).copy()
    s = np.zeros(NST * NST)
    txn = np.zeros(6 * NEN)
    ndm, nst, nel, isw_, block_pu = C.c_int(3), C.c_int(NST), C.c_int(1), C.c_int(isw), C.c_int(0)
    dptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    R.elmt17_pm_(dptr(e), dptr(iqm), dptr(iqh), dptr(xl), dptr(ul), dptr(dpl), dptr(p), dptr(s),
                 dptr(txn), C.byref(ndm), C.byref(nst), C.byref(nel), C.byref(isw_),
                 C.byref(block_pu))
    return p, s.reshape(NST, NST).T
