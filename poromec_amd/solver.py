"""Host side of the mef Biot u-p path on the MI355X: a thin Python mirror of
the reference's driver around the libpmhip C ABI.

The reference's operator interface for this path is the ``mesh`` macro
(structure), ``pform_pm`` (element loop + assembly, ``Pform.f:1``),
``solv_pm`` (``Solv.f:27``) and the ``solv`` macro time step
(``Main.f:876-1123``).  :class:`MefSolver` exposes the same steps with the
same argument meaning (``isw``, ``lhs``, ``rhs``, ``solver``, ``precond``,
``tol``, ``maxit``, ``ngram``) and raises :class:`PmError` where the
reference calls ``stop_mef``.  All compute runs in the HIP library; the
arrays here are inputs and outputs only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .problem import NDF, NEN, Deck

SOLVERS = {"pcg": 1, "gmres": 2, "bicgstab": 4, "minres": 7, "sqmr": 11}
SOLVER_NAMES = {v: k for k, v in SOLVERS.items()}
SPMV_SHAPES = 16  # PM_SPMV_SHAPES: (lanes per node block, columns in flight per lane) choices


@dataclass
class SolverConfig:
    """The ``solver`` config file fields (``read_solver_config_pm``,
    ``Rdata.f:3141``) plus the ``nl`` macro (``Main.f:2100-2121``)."""

    solver: str = "sqmr"        # reference default solver 11 (Main.f:289)
    precond: int = 2            # 2 diag, 5 |diag|
    tol: float = 1.0e-10
    maxit: int = 50000
    ngram: int = 150
    precision: str = "fp64"     # "fp64" or "fp32ir"
    ir_inner_tol: float = 1.0e-5
    ir_maxouter: int = 10
    maxnlit: int = 500
    nltol: float = 1.0e-4
    # The reference's nl macro defaults to criterion 2 (Main.f:273), and
    # read_deck keeps that default for decks.  Criterion 2 freezes the
    # displacement scale at nli 2 (Residuo.f:117-119); after a load step the
    # displacement residual is round-off from nli 1 on, so such a step never
    # converges and the reference stops at maxnlit.  Marches built in code
    # therefore default to criterion 1 (one norm over all equations).
    stop_crit: int = 1
    maxnlit_accept: bool = False  # keep the step at maxnlit instead of PM_ERR_NLIT

    def c(self) -> L.SolverCfg:
        return L.SolverCfg(SOLVERS[self.solver], self.precond, self.maxit, self.ngram, self.tol,
                           0 if self.precision == "fp64" else 1, self.ir_inner_tol,
                           self.ir_maxouter, self.maxnlit, self.nltol, self.stop_crit,
                           int(self.maxnlit_accept))


def _setup_ix(deck: Deck):
    """The connectivity as the mesh entry points take it: (numel, 21) for a
    maxno-20 deck, (numel, 9) -- 8 vertices + material -- for maxno 8."""
    if deck.nen == 8:
        return L.as_i32(np.concatenate([deck.ix[:, :8], deck.ix[:, NEN:]], axis=1))
    return L.as_i32(deck.ix)


def mesh_numbering(deck: Deck, reorder: bool = True):
    """Promotion (maxno 20), restraints, RCM and numbering on the host (no
    device)."""
    lib = L.lib()
    ix = _setup_ix(deck)
    idv = L.as_i32(deck.idv)
    cap = deck.nnodev + (12 * deck.numel if deck.nen == 20 else 0)
    ix20 = np.zeros((deck.numel, NEN + 1), np.int32)
    idm = np.zeros((cap, NDF), np.int32)
    perm = np.zeros(cap, np.int32)
    nnode, neq, nad = C.c_int(), C.c_int(), C.c_int64()
    L.check(lib.pm_mesh_numbering_nen(deck.nen, deck.numel, deck.nnodev, L.ptr(ix), L.ptr(idv),
                                      int(reorder), cap, L.ptr(ix20), L.ptr(idm), L.ptr(perm),
                                      C.byref(nnode), C.byref(neq), C.byref(nad)),
            "pm_mesh_numbering_nen")
    n = nnode.value
    return dict(ix=ix20, id=idm[:n].copy(), perm=perm[:n].copy(), nnode=n, neq=neq.value,
                nad=int(nad.value))


class MefSolver:
    """One problem resident on one MI355X."""

    def __init__(self, deck: Deck, device: int = 0, reorder: bool = True):
        self.lib = L.lib()
        self.deck = deck
        self.ctx = C.c_void_p()
        L.check(self.lib.pm_ctx_create(C.byref(self.ctx), device), "pm_ctx_create")
        x = L.as_f64(deck.x)
        idv = L.as_i32(deck.idv)
        ix = _setup_ix(deck)
        nnode, neq, nad = C.c_int(), C.c_int(), C.c_int64()
        if deck.nen == 8:  # maxno 8: the Q1/Q1 element on the mesh as read
            L.check(self.lib.pm_mesh_setup_nen(self.ctx, 8, deck.numel, deck.nnodev, L.ptr(ix),
                                               L.ptr(x), L.ptr(idv), int(reorder),
                                               C.byref(nnode), C.byref(neq), C.byref(nad)),
                    "pm_mesh_setup_nen")
        else:
            L.check(self.lib.pm_mesh_setup(self.ctx, deck.numel, deck.nnodev, L.ptr(ix),
                                           L.ptr(x), L.ptr(idv), int(reorder), C.byref(nnode),
                                           C.byref(neq), C.byref(nad)), "pm_mesh_setup")
        self.nnode, self.neq, self.nad = nnode.value, neq.value, int(nad.value)
        self.set_materials(deck)
        self.t = 0.0

    # -- set-up ------------------------------------------------------------
    def set_materials(self, deck: Deck):
        e = L.as_f64(deck.e)
        L.check(self.lib.pm_set_materials(self.ctx, e.shape[0], L.ptr(e), L.ptr(L.as_i32(deck.eload)),
                                          L.ptr(L.as_i32(deck.eloadp)), L.ptr(L.as_i32(deck.load)),
                                          L.ptr(L.as_f64(deck.fload)),
                                          L.ptr(L.as_f64(deck.gravity)), float(deck.gravity_mod)),
                "pm_set_materials")

    def set_vprop(self, ie, vprop):
        """Element type per material (17 / 18) and the conseq flags
        (vprop, kozeny_carman, density, hashin_shtrikman): properties at the
        integration points (elmt18_pm), set up as initial_prop does."""
        self.ie = L.as_i32(ie)
        fl = L.as_i32([int(bool(v)) for v in vprop])
        L.check(self.lib.pm_set_vprop(self.ctx, L.ptr(self.ie), L.ptr(fl)), "pm_set_vprop")

    def vprop(self):
        """The integration points' properties, (numel, 64, 7)."""
        v = np.zeros((self.deck.numel, 64, 7))
        L.check(self.lib.pm_get_vprop(self.ctx, L.ptr(v)), "pm_get_vprop")
        return v

    def numbering(self):
        ix20 = np.zeros((self.deck.numel, NEN + 1), np.int32)
        idm = np.zeros((self.nnode, NDF), np.int32)
        perm = np.zeros(self.nnode, np.int32)
        L.check(self.lib.pm_get_numbering(self.ctx, L.ptr(ix20), L.ptr(idm), L.ptr(perm)),
                "pm_get_numbering")
        return ix20, idm, perm

    # -- pform_pm / solv_pm --------------------------------------------------
    def pform(self, isw: int, lhs: bool, rhs: bool, dt: float, t: float, target: str = "b"):
        L.check(self.lib.pm_pform(self.ctx, isw, int(lhs), int(rhs), dt, t,
                                  1 if target == "b" else 0), "pm_pform")

    def solve(self, b=None, cfg: SolverConfig | None = None, return_x: bool = True):
        cfg = cfg or SolverConfig()
        info = L.SolveInfo()
        bb = None if b is None else L.as_f64(b)
        x = np.zeros(self.neq) if return_x else None
        L.check(self.lib.pm_solve(self.ctx, C.byref(cfg.c()), L.ptr(bb), L.ptr(x),
                                  C.byref(info)), "pm_solve")
        return x, _solve_info(info)

    def spmv(self, x):
        y = np.zeros(self.neq)
        L.check(self.lib.pm_spmv(self.ctx, L.ptr(L.as_f64(x)), L.ptr(y)), "pm_spmv")
        return y

    def export_csrc(self):
        ia = np.zeros(self.neq + 1, np.int64)
        L.check(self.lib.pm_export_csrc(self.ctx, L.ptr(ia), None, None, None), "pm_export_csrc")
        nad = int(ia[-1] - 1)
        ja = np.zeros(max(nad, 1), np.int32)
        al = np.zeros(max(nad, 1))
        ad = np.zeros(self.neq)
        L.check(self.lib.pm_export_csrc(self.ctx, L.ptr(ia), L.ptr(ja), L.ptr(ad), L.ptr(al)),
                "pm_export_csrc")
        return ia, ja[:nad], ad, al[:nad]

    # -- state and the time step ------------------------------------------
    def set_state(self, u=None, u0=None, dp=None, porosity=None, t=0.0):
        L.check(self.lib.pm_set_state(self.ctx, L.ptr(None if u is None else L.as_f64(u)),
                                      L.ptr(None if u0 is None else L.as_f64(u0)),
                                      L.ptr(None if dp is None else L.as_f64(dp)),
                                      L.ptr(None if porosity is None else L.as_f64(porosity)),
                                      float(t)), "pm_set_state")
        self.t = t

    def state(self):
        u = np.zeros((self.nnode, NDF))
        u0 = np.zeros((self.nnode, NDF))
        dp = np.zeros(self.nnode)
        por = np.zeros(self.nnode)
        t = C.c_double()
        L.check(self.lib.pm_get_state(self.ctx, L.ptr(u), L.ptr(u0), L.ptr(dp), L.ptr(por),
                                      C.byref(t)), "pm_get_state")
        return dict(u=u, u0=u0, dp=dp, porosity=por, t=t.value)

    def vectors(self):
        b0, b, x0 = np.zeros(self.neq), np.zeros(self.neq), np.zeros(self.neq)
        L.check(self.lib.pm_get_vectors(self.ctx, L.ptr(b0), L.ptr(b), L.ptr(x0)), "pm_get_vectors")
        return b0, b, x0

    def set_vectors(self, b0=None, b=None, x0=None):
        f = lambda a: None if a is None else L.ptr(L.as_f64(a))  # noqa: E731
        L.check(self.lib.pm_set_vectors(self.ctx, f(b0), f(b), f(x0)), "pm_set_vectors")

    def step(self, cfg: SolverConfig | None = None, dt: float | None = None):
        cfg = cfg or SolverConfig()
        info = L.StepInfo()
        dt = self.deck.dt if dt is None else dt
        L.check(self.lib.pm_step(self.ctx, dt, C.byref(cfg.c()), C.byref(info)), "pm_step")
        self.t += dt
        n = min(info.nsolve, 16)
        return dict(nlit=info.nlit, nsolve=info.nsolve, iters=list(info.iters[:n]),
                    used=list(info.used[:n]), true_rel=list(info.true_rel[:n]),
                    ms_total=info.ms_total, ms_solve=info.ms_solve, ms_element=info.ms_element)

    def tform(self):
        """Vertex stresses and Darcy flux of the current state (the pnsf
        macro's tform_pm + effective_stress): total, Terzaghi and Biot stress
        as (nnode, 6) arrays (xx yy zz xy yz xz) and the flux as (nnode, 3)."""
        n = self.nnode
        tx, txe, txb, flux = np.zeros((n, 6)), np.zeros((n, 6)), np.zeros((n, 6)), np.zeros((n, 3))
        L.check(self.lib.pm_tform(self.ctx, L.ptr(tx), L.ptr(txe), L.ptr(txb), L.ptr(flux)),
                "pm_tform")
        return tx, txe, txb, flux

    def set_ordered(self, on: bool = True):
        """Sum every reduction and SpMV row in the reference's sequential
        order (slow; for bit-level parity checks)."""
        L.check(self.lib.pm_set_ordered(self.ctx, int(on)), "pm_set_ordered")

    # -- adjoint form -----------------------------------------------------------
    def adjoint(self, nsteps: int, sensors, target, cfg: SolverConfig | None = None,
                dt: float | None = None, ckpt_every: int = 8):
        """March ``nsteps`` steps from the current state, measure
        ``J = sum_n sum_i (u[node_i, dof_i] after step n - target[n, i])**2``
        and return ``(J, dJ/dlog E, dJ/dlog k)`` per material, the gradients
        with respect to ``log e[:, 0]`` and ``log e[:, 2]``.  ``sensors`` is a
        list of ``(node, dof)`` pairs of free dofs (dof 3 = pressure).  The
        state is left at the end of the march."""
        cfg = cfg or SolverConfig()
        dt = self.deck.dt if dt is None else dt
        sens = np.asarray(sensors, dtype=np.int64).reshape(-1, 2)
        sdof = L.as_i32(4 * sens[:, 0] + sens[:, 1])
        tgt = L.as_f64(np.asarray(target, dtype=np.float64).reshape(nsteps, len(sdof)))
        numat = self.deck.e.shape[0]
        gE, gk = np.zeros(numat), np.zeros(numat)
        J = C.c_double()
        L.check(self.lib.pm_adjoint(self.ctx, int(nsteps), float(dt), C.byref(cfg.c()), len(sdof),
                                    L.ptr(sdof), L.ptr(tgt), int(ckpt_every), C.byref(J),
                                    L.ptr(gE), L.ptr(gk)), "pm_adjoint")
        self.t += nsteps * dt
        return J.value, gE, gk

    # -- batching form --------------------------------------------------------
    def batch_define(self, elem_off):
        self.elem_off = L.as_i32(elem_off)
        self.nbox = len(self.elem_off) - 1
        L.check(self.lib.pm_batch_define(self.ctx, self.nbox, L.ptr(self.elem_off)),
                "pm_batch_define")

    def batch_iters(self):
        it = np.zeros(self.nbox, np.int32)
        L.check(self.lib.pm_batch_iters(self.ctx, L.ptr(it)), "pm_batch_iters")
        return it

    def batch_graphs(self, on: bool):
        L.check(self.lib.pm_batch_graphs(self.ctx, int(on)), "pm_batch_graphs")

    def matrix_info(self):
        neq, nb, nnz, ncols = C.c_int(), C.c_int(), C.c_int64(), C.c_int64()
        L.check(self.lib.pm_matrix_info(self.ctx, C.byref(neq), C.byref(nb), C.byref(nnz),
                                        C.byref(ncols)), "pm_matrix_info")
        nruns = C.c_int64()
        L.check(self.lib.pm_matrix_runs(self.ctx, C.byref(nruns)), "pm_matrix_runs")
        return dict(neq=neq.value, nblocks=nb.value, nnz=int(nnz.value), ncols=int(ncols.value),
                    nruns=int(nruns.value))

    def bench_spmv(self, use32: bool, reps: int = 20) -> float:
        ms = C.c_double()
        L.check(self.lib.pm_bench_spmv(self.ctx, int(use32), reps, C.byref(ms)), "pm_bench_spmv")
        return ms.value

    def tune_spmv(self, use32: bool, reps: int = 10):
        """Time every SpMV shape on the current matrix and keep the fastest;
        returns (best shape, ms per launch of each shape)."""
        ms = np.zeros(SPMV_SHAPES)
        best = C.c_int()
        L.check(self.lib.pm_tune_spmv(self.ctx, int(use32), int(reps), C.byref(best), L.ptr(ms)),
                "pm_tune_spmv")
        return best.value, ms

    def set_spmv_shape(self, use32: bool, shape: int):
        L.check(self.lib.pm_set_spmv_shape(self.ctx, int(use32), int(shape)), "pm_set_spmv_shape")

    def spmv_timing(self, every: int) -> dict:
        """Time every `every`-th solver SpMV with HIP events (0: off); returns
        what was accumulated since the previous call (ms and timed launches
        per value precision)."""
        st = np.zeros(4)
        L.check(self.lib.pm_spmv_timing(self.ctx, int(every), L.ptr(st)), "pm_spmv_timing")
        return dict(fp64_ms=st[0], fp64_n=int(st[1]), fp32_ms=st[2], fp32_n=int(st[3]))

    def bench_copy(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double()
        L.check(self.lib.pm_bench_copy(self.ctx, int(nbytes), reps, C.byref(ms)), "pm_bench_copy")
        return ms.value

    def bench_stream(self, kind: int, nbytes: int, reps: int = 20) -> float:
        """ms per launch of streaming kernel `kind` over nbytes (pm_bench_stream)."""
        ms = C.c_double()
        L.check(self.lib.pm_bench_stream(self.ctx, int(kind), int(nbytes), reps, C.byref(ms)),
                "pm_bench_stream")
        return ms.value

    def close(self):
        if self.ctx:
            self.lib.pm_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CsrcSolver:
    """Drop-in ``solv_pm`` for a caller that assembles its own CSRC matrix."""

    def __init__(self, ia, ja, device: int = 0):
        self.lib = L.lib()
        self.ctx = C.c_void_p()
        L.check(self.lib.pm_ctx_create(C.byref(self.ctx), device), "pm_ctx_create")
        self.ia = np.ascontiguousarray(ia, np.int64)
        self.ja = L.as_i32(ja)
        self.neq = len(self.ia) - 1
        L.check(self.lib.pm_struct_upload(self.ctx, self.neq, int(self.ia[-1] - 1),
                                          L.ptr(self.ia), L.ptr(self.ja)), "pm_struct_upload")

    def set_ordered(self, on: bool = True):
        """Reference summation order (see MefSolver.set_ordered)."""
        L.check(self.lib.pm_set_ordered(self.ctx, int(on)), "pm_set_ordered")

    def set_values(self, ad, al):
        L.check(self.lib.pm_values_upload(self.ctx, L.ptr(L.as_f64(ad)), L.ptr(L.as_f64(al))),
                "pm_values_upload")

    def solve(self, b, cfg: SolverConfig | None = None):
        cfg = cfg or SolverConfig()
        info = L.SolveInfo()
        x = np.zeros(self.neq)
        L.check(self.lib.pm_solve(self.ctx, C.byref(cfg.c()), L.ptr(L.as_f64(b)), L.ptr(x),
                                  C.byref(info)), "pm_solve")
        return x, _solve_info(info)

    def spmv(self, x):
        y = np.zeros(self.neq)
        L.check(self.lib.pm_spmv(self.ctx, L.ptr(L.as_f64(x)), L.ptr(y)), "pm_spmv")
        return y

    def close(self):
        if self.ctx:
            self.lib.pm_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _solve_info(info: L.SolveInfo) -> dict:
    return dict(iters=info.iters, outer=info.outer, used=info.used_solver, resid=info.resid,
                true_rel=info.true_rel, ms=info.ms, spmv64=info.spmv64, spmv32=info.spmv32)
