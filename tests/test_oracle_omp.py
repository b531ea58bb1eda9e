"""The OpenMP restatement of the reference's threaded CPU path (the CPU
baseline bench.py times): rpsqmr_omp with matvec_csrc_sym_pm_omp and
dot_par_omp (Iters_omp.f:349-700, Matvec_omp.f:461-587,845-873,
Matrix_partition.f:130-236), and the coloured element loop of pform_pm
(Pform.f:172-303).

On one thread every sum runs in the sequential order, so the OpenMP SQMR is
the sequential oracle's SQMR bit for bit.  On several threads the transposed
terms meet in per-thread buffers, which moves the last bits only: the solve
still meets its stop rule and agrees with the dense solution.  The coloured
element loop assembles the same matrix as the sequential loop to rounding,
and no colour holds two elements that share a vertex."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from poromec_amd import problem as P


@pytest.fixture(scope="module")
def system():
    d = P.heterogeneous_box_deck(3, seed=3, sigma_lnk=0.8, corr_len=2.0, dt=0.05, nsteps=4)
    st = O.build_structure(d)
    m = O.OracleModel(d, st)
    m.assemble(isw=2, lhs=True, rhs=False)
    mv = np.zeros(st["neq"])
    O.lib.orc_precond(2, st["neq"], m.ad, mv)
    b = np.random.default_rng(1).standard_normal(st["neq"])
    return d, st, m, mv, b


def test_one_thread_is_the_sequential_sqmr(system):
    d, st, m, mv, b = system
    x = np.zeros(st["neq"])
    res, fail = C.c_double(), C.c_int()
    it_s = O.lib.orc_rpsqmr(st["neq"], st["ia"], st["ja"], m.ad, m.al, mv, b, x, 1e-10, 50000,
                            C.byref(res), C.byref(fail))
    it_o, x_o, fail_o = O.omp_sqmr(st, m.ad, m.al, mv, b, nthreads=1)
    assert it_s > 0 and fail_o == 0
    assert it_o == it_s
    assert np.array_equal(x_o, x)


@pytest.mark.parametrize("nthreads", [2, 3, 4])
def test_threads_solve_the_system(system, nthreads):
    d, st, m, mv, b = system
    it_s, x_s, _ = O.omp_sqmr(st, m.ad, m.al, mv, b, nthreads=1)
    it, x, fail = O.omp_sqmr(st, m.ad, m.al, mv, b, nthreads=nthreads)
    assert fail == 0 and it > 0
    # SQMR's counts on this saddle-point system move by a few percent under
    # rounding-level changes (the oracle's own 1-ulp spread)
    assert abs(it - it_s) <= 2 + 0.1 * it_s, (it, it_s)
    assert np.abs(x - x_s).max() <= 1e-6 * np.abs(x_s).max()


def test_maxit_sets_the_fallback_flag(system):
    d, st, m, mv, b = system
    it, x, fail = O.omp_sqmr(st, m.ad, m.al, mv, b, maxit=3, nthreads=2)
    assert (it, fail) == (-2, 1)


def test_colouring_separates_elements_sharing_a_vertex(system):
    d, st, m, mv, b = system
    nc, col = m.colour()
    assert 8 <= nc <= 27 and col.min() >= 1 and col.max() == nc
    ix = st["ix"][:, :8]
    for c in range(1, nc + 1):
        verts = ix[col == c].ravel()
        assert len(np.unique(verts)) == len(verts), c


def test_coloured_loop_assembles_the_matrix(system):
    d, st, m, mv, b = system
    ad_ref, al_ref = m.ad.copy(), m.al.copy()
    nc, col = m.colour()
    m.ad[:] = 0.0
    m.al[:] = 0.0
    done = 0
    for c in range(1, nc + 1):
        _, n = m.pform_colour(col, c, isw=2, lhs=True, rhs=False, nthreads=3)
        done += n
    assert done == st["numel"]
    scale = np.abs(ad_ref).max()
    assert np.abs(m.ad - ad_ref).max() <= 1e-13 * scale
    assert np.abs(m.al - al_ref).max() <= 1e-13 * scale
    m.ad[:], m.al[:] = ad_ref, al_ref
