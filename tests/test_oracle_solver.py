"""Pins the oracle's SpMV, Jacobi and Krylov solvers (Matvec.f, Precond.f,
Iters.f) -- none of which can be built from the reference sources here
(they need MPI) -- by what the mathematics fixes:

* matvec_csrc_sym_pm: y = (D + L + L^T) x against a dense product;
* pre_diag: m = 1/ad and |1/ad|;
* every solver: the solution of small SPD and symmetric-indefinite systems
  against a dense LU solve, and the residual it claims;
* PCG on A = a I + b 11^T (two distinct eigenvalues, constant diagonal so
  Jacobi keeps them) converges in 2 iterations, as CG does in exact
  arithmetic; GMRES(k >= n) converges within n iterations;
* the stop rules: the solver returns the first iteration that passes its
  test, so a run with maxit one below that count reports no convergence.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P

lib = O.lib


def dense_to_csrc(A):
    n = A.shape[0]
    ia = [1]
    ja, al = [], []
    for i in range(n):
        for j in range(i):
            if A[i, j] != 0.0:
                ja.append(j + 1)
                al.append(A[i, j])
        ia.append(len(ja) + 1)
    return (np.array(ia, np.int64), np.array(ja or [0], np.int32),
            np.ascontiguousarray(np.diag(A)), np.array(al or [0.0]))


def csrc_to_dense(ia, ja, ad, al):
    n = len(ad)
    A = np.diag(ad).astype(float)
    for i in range(n):
        for k in range(ia[i] - 1, ia[i + 1] - 1):
            A[i, ja[k] - 1] = al[k]
            A[ja[k] - 1, i] = al[k]
    return A


def run(solver, A, b, precond=2, tol=1e-12, maxit=2000, k=150):
    ia, ja, ad, al = dense_to_csrc(A)
    n = len(b)
    m = np.zeros(n)
    lib.orc_precond(precond, n, ad, m)
    x = np.zeros(n)
    res = C.c_double()
    args = [n, ia, ja, ad, al, m, np.ascontiguousarray(b, dtype=float), x]
    if solver == "pcg":
        it = lib.orc_pcg(*args, tol, maxit, C.byref(res))
    elif solver == "sqmr":
        fail = C.c_int()
        it = lib.orc_rpsqmr(*args, tol, maxit, C.byref(res), C.byref(fail))
    elif solver == "bicgstab":
        it = lib.orc_pbicgstab(*args, tol, maxit, C.byref(res))
    elif solver == "minres":
        it = lib.orc_pminres(*args, tol, maxit, 0, C.byref(res))
    else:
        it = lib.orc_gmres2(*args, k, tol, maxit, C.byref(res))
    return it, x


def spd(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return Q @ np.diag(rng.uniform(1, 10, n)) @ Q.T


def biot_small():
    d = P.column_deck(1, 1, 3)
    m = O.OracleModel(d)
    m.assemble(isw=2, lhs=True, rhs=False)
    st = m.st
    return csrc_to_dense(st["ia"], st["ja"], m.ad, m.al[: st["nad"]])


def test_matvec_against_dense():
    A = biot_small()
    ia, ja, ad, al = dense_to_csrc(A)
    x = np.random.default_rng(0).standard_normal(len(ad))
    y = np.zeros(len(ad))
    lib.orc_matvec_csrc(len(ad), ia, ja, ad, al, x, y)
    assert np.allclose(y, A @ x, rtol=1e-13, atol=1e-13 * np.abs(A @ x).max())


def test_jacobi():
    ad = np.array([2.0, -4.0, 0.5])
    m = np.zeros(3)
    lib.orc_precond(2, 3, ad, m)
    assert np.array_equal(m, 1.0 / ad)
    lib.orc_precond(5, 3, ad, m)
    assert np.array_equal(m, np.abs(1.0 / ad))


@pytest.mark.parametrize("solver", ["pcg", "sqmr", "bicgstab", "minres", "gmres"])
def test_spd_solution(solver):
    A = spd(30, 1)
    b = np.random.default_rng(2).standard_normal(30)
    it, x = run(solver, A, b, precond=5 if solver == "minres" else 2)
    assert it > 0
    assert np.allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-9 * np.abs(x).max())


@pytest.mark.parametrize("solver", ["sqmr", "bicgstab", "gmres"])
def test_biot_indefinite_solution(solver):
    A = biot_small()
    ev = np.linalg.eigvalsh(A)
    assert ev.min() < 0 < ev.max()  # saddle point: Kpp is negative
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    # BiCGSTAB on this saddle-point system stagnates below ~1e-8 and then
    # trips the reference's |w| < 1e-32 ||b|| breakdown test (Iters.f:2985)
    tol = 1e-8 if solver == "bicgstab" else 1e-13
    it, x = run(solver, A, b, precond=5 if solver == "minres" else 2, tol=tol, maxit=5000)
    assert it > 0
    xd = np.linalg.solve(A, b)
    assert np.abs(x - xd).max() <= (1e-6 if solver == "bicgstab" else 1e-7) * np.abs(xd).max()


def test_minres_abs_alpha_quirk():
    """pminres takes alpha = |z0 . A z0| / beta^2 (Iters.f:4146).  On the
    indefinite Biot matrix some Lanczos alphas are negative, so the
    reference's MINRES (and its restatement) does not converge there, while a
    plain numpy MINRES with signed alpha does; on SPD systems it converges."""
    A = biot_small()
    b = np.random.default_rng(3).standard_normal(A.shape[0])
    it, _ = run("minres", A, b, precond=5, tol=1e-10, maxit=400)
    assert it == -2
    m = np.abs(1.0 / np.diag(A))
    x, conv = np.zeros_like(b), 1e-10 * np.sqrt(np.sum((b * np.sqrt(m)) ** 2))
    # signed-alpha preconditioned Lanczos + QR (same recurrences, no abs)
    v, v0, w0, w00 = b.copy(), np.zeros_like(b), np.zeros_like(b), np.zeros_like(b)
    z0 = v * m
    beta = np.sqrt(abs(v @ z0))
    normr, neta, beta_old, c_old, c, s_old, s_ = beta, beta, beta, 1.0, 1.0, 0.0, 0.0
    for _ in range(400):
        p = A @ z0
        alpha = (z0 @ p) / beta**2
        p = p / beta - (alpha / beta) * v - (beta / beta_old) * v0
        z = p * m
        beta_old, beta = beta, np.sqrt(abs(z @ p))
        delta = c * alpha - c_old * s_ * beta_old
        ro2, ro3 = s_ * alpha + c_old * c * beta_old, s_old * beta_old
        c_old, s_old = c, s_
        ro1 = np.hypot(delta, beta)
        c, s_ = delta / ro1, beta / ro1
        w = z0 / (beta_old * ro1) - (ro2 / ro1) * w0 - (ro3 / ro1) * w00
        x = x + c * neta * w
        v0, v, z0, w00, w0 = v, p, z, w0, w
        normr = abs(s_) * normr
        if normr < conv:
            break
        neta = -s_ * neta
    assert normr < conv
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-7 * np.abs(x).max())


def test_pcg_two_eigenvalues():
    n = 25
    A = 3.0 * np.eye(n) + 0.5 * np.ones((n, n))
    b = np.random.default_rng(4).standard_normal(n)
    it, x = run("pcg", A, b, tol=1e-12)
    assert it == 2
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-12)


def test_gmres_full_krylov_space():
    A = spd(20, 5) + np.diag(np.linspace(-3, 3, 20))
    b = np.random.default_rng(6).standard_normal(20)
    it, x = run("gmres", A, b, tol=1e-12, k=150)
    assert 0 < it <= 20
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-8 * np.abs(x).max())


@pytest.mark.parametrize("solver", ["pcg", "sqmr", "bicgstab"])
def test_stop_rule_is_first_passing_iteration(solver):
    A = spd(40, 7)
    b = np.random.default_rng(8).standard_normal(40)
    it, _ = run(solver, A, b, tol=1e-10)
    assert it > 1
    it2, _ = run(solver, A, b, tol=1e-10, maxit=it - 1)
    assert it2 < 0
    it3, _ = run(solver, A, b, tol=1e-10, maxit=it)
    assert it3 == it
