"""Special functions and linear algebra of the oracle against closed forms,
scipy/sympy and the reference's own known answers
(testSpecialFunctions.cpp:190-490, test_matrix.py:34-58)."""
import numpy as np
import pytest
from scipy.special import lpmv, spherical_jn, spherical_yn, gammaln

import oracle as O


def test_rv_over_rs_known_answers():
    assert abs(O.rv_over_rs(0.5) - 0.9637112829756893) < 1e-10
    assert abs(O.rv_over_rs(2.0) - 0.955443263377723) < 1e-10
    assert O.rv_over_rs(1.0) == 1.0


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 16, 41, 200])
def test_gauss_legendre_vs_numpy(order):
    x, w = O.gauss_legendre(order)
    xr, wr = np.polynomial.legendre.leggauss(order)
    assert np.allclose(x, xr, atol=1e-14, rtol=0)
    assert np.allclose(w, wr, atol=1e-14, rtol=0)


def test_gauss_legendre_closed_forms():
    x, w = O.gauss_legendre(2)
    assert np.allclose(x, [-1 / np.sqrt(3), 1 / np.sqrt(3)], atol=1e-15)
    x, w = O.gauss_legendre(3)
    assert np.allclose(x, [-np.sqrt(0.6), 0., np.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], atol=1e-15)
    x, w = O.gauss_legendre(10, 0., 1.)
    assert abs(w.sum() - 1.) < 1e-14 and x.min() > 0 and x.max() < 1


@pytest.mark.parametrize("z", [0.01, 0.7, 3.3, 17.2, 63.0])
def test_bessel_real_vs_scipy(z):
    n = np.arange(1, 41)
    j, dj = O.sph_j(40, z)
    y, dy = O.sph_y(40, z)
    jr = spherical_jn(n, z)
    assert np.allclose(j, jr, rtol=1e-11, atol=1e-300)
    assert np.allclose(dj, (jr + z * spherical_jn(n, z, True)) / z,
                       rtol=1e-9, atol=1e-13 * np.abs(jr).max())
    yr = spherical_yn(n, z)
    big = np.abs(yr) < 1e250
    assert np.allclose(y[big], yr[big], rtol=1e-10)


@pytest.mark.parametrize("z", [0.3 + 0.01j, 2.5 + 0.4j, 20.0 + 1.0j,
                               45.0 + 9.0j])
def test_bessel_complex_vs_scipy(z):
    n = np.arange(1, 31)
    j, dj = O.sph_j(30, z)
    jr = spherical_jn(n, z)
    assert np.allclose(j, jr, rtol=1e-10, atol=1e-14 * np.abs(jr).max())
    djr = (jr + z * spherical_jn(n, z, True)) / z
    assert np.allclose(dj, djr, rtol=1e-9, atol=1e-13 * np.abs(djr).max())


def test_bessel_j_near_zero_of_j0_loses_accuracy():
    """The ratio recurrence multiplies rho_0 = j_1/j_0 by j_0; at a zero of
    j_0 (z = pi) that is 0 x inf and the reference algorithm
    (SpecialFunctions.hpp:202-240) returns O(1)-wrong values.  The oracle
    keeps the reference's behaviour; callers avoid exact zeros."""
    j, _ = O.sph_j(3, np.pi)
    assert abs(j[0] - spherical_jn(1, np.pi)) > 1e-3


@pytest.mark.parametrize("m", [0, 1, 2, 5, 9])
def test_wigner_vs_associated_legendre(m):
    nmax = 12
    for x in (-0.93, -0.31, 0.2, 0.77):
        y, dy = O.wigner_dn_0m(x, nmax, m)
        for n in range(max(m, 1), nmax + 1):
            # d^n_{0m} = (-1)^m sqrt((n-m)!/(n+m)!) P^m_n
            # (SpecialFunctions.hpp:246-249); scipy's P^m_n carries the
            # Condon-Shortley phase like the reference's
            f = np.exp(0.5 * (gammaln(n - m + 1) - gammaln(n + m + 1)))
            ref = (-1) ** m * f * lpmv(m, n, x)
            assert abs(y[n - 1] - ref) < 1e-12, (m, n, x)
        # derivative against a central difference in theta
        th = np.arccos(x)
        h = 1e-6
        yp, _ = O.wigner_dn_0m(np.cos(th + h), nmax, m)
        ym, _ = O.wigner_dn_0m(np.cos(th - h), nmax, m)
        assert np.allclose(dy[m:], ((yp - ym) / (2 * h))[m:], atol=1e-7)


def test_wigner_sinx_special_case():
    y, dy = O.wigner_dn_0m_sinx(1.0, 0.0, 9000.0, 6, 1)
    assert np.allclose(y, 0.5 * np.sqrt(np.arange(1, 7) * np.arange(2, 8)))
    y, dy = O.wigner_dn_0m_sinx(0.3, np.sqrt(0.91), 1 / np.sqrt(0.91), 6, 2)
    yr, _ = O.wigner_dn_0m(0.3, 6, 2)
    assert np.allclose(y, yr / np.sqrt(0.91))


def test_plu_inverse_vs_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 3, 7, 20):
        A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        Ai = O.plu_inverse(A)
        assert np.allclose(Ai, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
    # the reference's 3x3 case (testMatrix.cpp) shape and a pivoting case
    A = np.array([[0, 1, 2], [3, 0, 1], [1, 2, 0]], dtype=complex)
    assert np.allclose(O.plu_inverse(A) @ A, np.eye(3), atol=1e-14)


def test_clebsch_gordan_vs_sympy():
    from sympy import S
    from sympy.physics.quantum.cg import CG
    cases = [(1, 0, 1, 0, 2, 0), (2, 1, 1, -1, 2, 0), (3, 2, 2, -1, 4, 1),
             (2, -1, 2, 1, 3, 0), (4, 0, 3, 0, 5, 0), (3, -2, 3, 1, 4, -1),
             (5, 3, 4, -3, 6, 0)]
    for (n1, m1, n2, m2, N, M) in cases:
        ref = float(CG(S(n1), S(m1), S(n2), S(m2), S(N), S(M)).doit())
        assert abs(O.clebsch_gordan(n1, m1, n2, m2, N, M) - ref) < 1e-10


def test_r_dr_spheroid_mirror():
    x, _ = O.gauss_legendre(16)
    r2, dr = O.r_dr_spheroid(x, 0.5)
    assert np.allclose(r2, r2[::-1]) and np.allclose(dr, -dr[::-1])
    # r^2 = d^(2/3)/(sin^2 + d^2 cos^2): volume of the spheroid = sphere of 1
    d = 0.5
    assert np.allclose(r2, d ** (2 / 3) / ((1 - x ** 2) + d * d * x ** 2))
