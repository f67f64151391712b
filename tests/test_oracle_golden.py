"""The CPU oracle against the reference's golden fixture.

test/test_tmatrixcalculator.txt holds 101 rows of outputs of Mishchenko's
T-matrix code for spheroids (axi = 10 um equal-surface radius, lambda ~
6.28 um, m = 1.5 + 0.02i, d in [0.47, 3.99]); the reference checks
Q_ext, Q_sca and the albedo to 1e-5 (testTMatrixCalculator.cpp:64-136) with
the relative difference |a-b|/|a+b| (test/Assert.hpp:67-72).
"""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROWS = np.loadtxt(os.path.join(HERE, "golden",
                               "mishchenko_tmatrixcalculator.txt"),
                  comments="#")


def rel(a, b):
    return abs(a - b) / abs(a + b)


@pytest.mark.parametrize("row", list(range(0, 101, 4)))
def test_calculator_path_matches_fixture(row):
    r = ROWS[row]
    axi, rat, lam, mrr, mri, eps, ddelt, ndgs = r[:8]
    st, n, g, rv = O.converge_calculator(rat, eps, axi, lam, complex(mrr, mri),
                                         200, ddelt * 0.1, int(ndgs), 500)
    assert st == 0
    T = O.tmatrix_all(rv, eps, lam, complex(mrr, mri), n, g)
    qs, qe = O.qcoeffs(n, T)
    assert rel(qe, r[15]) < 1e-5
    assert rel(qs, r[14]) < 1e-5
    assert rel(-qs / qe, r[16]) < 1e-5


def test_task_path_row1_matches_fixture():
    """Task-path convergence (orders from 4, ngauss = 2 order) on fixture
    row 1 (testTaskBasedTMatrixCalculation.cpp:191-323: Q to 1e-5)."""
    r = ROWS[0]
    axi, rat, lam, mrr, mri, eps = r[:6]
    rv = O.rv_over_rs(eps) * axi
    st, n, g, dqs, dqe = O.converge_task(rv, eps, lam, complex(mrr, mri), 4,
                                         40, 2, 1e-4)
    assert st == 0 and n < 39
    T = O.tmatrix_all(rv, eps, lam, complex(mrr, mri), n, g)
    qs, qe = O.qcoeffs(n, T)
    assert rel(qe, r[15]) < 1e-5
    assert rel(qs, r[14]) < 1e-5


def test_planted_mistake_is_caught():
    """A dropped factor in Q_ext (here: the f_m = 2 weight of m > 0) moves
    the fixture row far outside the tolerance, so the fixture pins that
    part of the sum."""
    r = ROWS[0]
    axi, rat, lam, mrr, mri, eps, ddelt, ndgs = r[:8]
    st, n, g, rv = O.converge_calculator(rat, eps, axi, lam, complex(mrr, mri),
                                         200, ddelt * 0.1, int(ndgs), 500)
    T = O.tmatrix_all(rv, eps, lam, complex(mrr, mri), n, g)
    blocks = O.tmatrix_blocks(T, n)
    qe_wrong = sum(np.trace(b).real for b in blocks)  # no f_m, no 12/21
    assert rel(qe_wrong, r[15]) > 1e-2


@pytest.mark.parametrize("row", list(range(0, 101, 10)))
def test_scattering_matrix_matches_fixture(row):
    """Z at the fixture's orientation and angles, in um^2, to the
    reference's 2e-2 relative bar (testTMatrixCalculator.cpp:115-133)."""
    r = ROWS[row]
    axi, rat, lam, mrr, mri, eps, ddelt, ndgs = r[:8]
    alpha, beta, th0, th, ph0, ph = np.radians(r[8:14])
    mr = complex(mrr, mri)
    st, n, g, rv = O.converge_calculator(rat, eps, axi, lam, mr, 200,
                                         ddelt * 0.1, int(ndgs), 500)
    T = O.tmatrix_all(rv, eps, lam, mr, n, g)
    S, Z, K = O.scattering_query(n, T, lam, alpha, beta, th0, ph0, th, ph)
    Zref = r[17:33].reshape(4, 4)
    err = np.abs(Z - Zref) / np.abs(Z + Zref)
    assert err.max() < 2e-2, err.max()
