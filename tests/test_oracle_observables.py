"""Observable and averaging restatements of the oracle, pinned by facts
independent of their formulas: Mie theory for spheres (scipy Bessel
functions), rotational invariance of isotropic ensembles, a brute-force
orientation quadrature (the reference's own check,
testTMatrixCalculator.cpp:139-195) and the backscattering identity
Z11 - Z22 + Z33 - Z44 = 0 (testTMatrixCalculator.cpp:198-212)."""
import numpy as np
import pytest

import oracle as O
from test_oracle_mie import mie_ab


def sphere_case():
    rv, lam, mr = 0.3, 0.63, 1.7 + 0.1j
    x = 2 * np.pi * rv / lam
    nmax = int(x + 4.05 * x ** (1 / 3)) + 6
    return rv, lam, mr, x, nmax


def test_sphere_extinction_absorption_vs_mie():
    rv, lam, mr, x, nmax = sphere_case()
    a, b = mie_ab(x, mr, nmax)
    n = np.arange(1, nmax + 1)
    qext = 2 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    qsca = 2 / x ** 2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
    T = O.tmatrix_all(rv, 1.0, lam, mr, nmax, 2 * nmax)
    th = np.array([0.0, 0.4, 1.2, np.pi / 2, 2.5, np.pi])
    qe, qep, qec = O.extinction(nmax, T, lam, rv, th)
    assert np.allclose(qe, qext, rtol=1e-9)
    assert np.allclose(qep, 0, atol=1e-12) and np.allclose(qec, 0, atol=1e-12)
    # with a fine quadrature of the scattered light Q_abs -> Q_ext - Q_sca
    qa, qap = O.absorption(nmax, T, lam, rv, th[:3], ngauss=40)
    assert np.allclose(qa, qext - qsca, rtol=1e-7)
    qa0, _ = O.absorption(nmax, T, lam, rv, th[:3], account_for_scattering=0)
    assert np.allclose(qa0, qext, rtol=1e-9)


def test_sphere_scattering_matrix_vs_mie():
    rv, lam, mr, x, nmax = sphere_case()
    k = 2 * np.pi / lam
    T = O.tmatrix_all(rv, 1.0, lam, mr, nmax, 2 * nmax)
    nth = 6
    Z = O.scattering_grid(nmax, T, lam, nth, use_gl=True)
    a, b = mie_ab(x, mr, nmax)
    cth, _ = np.polynomial.legendre.leggauss(nth)
    # incidence theta = pi/2, phi = 0; scattering angle of (theta, phi)
    phis, _ = np.polynomial.legendre.leggauss(2 * nth)
    phis = np.pi * (phis + 1)
    from scipy.special import lpmv
    n = np.arange(1, nmax + 1)
    for it, ct in enumerate(cth):
        st = np.sqrt(1 - ct * ct)
        for ip, ph in enumerate(phis):
            cs = st * np.cos(ph)  # cos of the scattering angle
            mu = np.clip(cs, -1, 1)
            s = np.sqrt(1 - mu * mu)
            p1 = np.array([-lpmv(1, nn, mu) for nn in n])
            pi_n = p1 / s
            tau_n = np.array([nn * mu * p1[i] / s - (nn + 1) * (
                -lpmv(1, nn - 1, mu) if nn > 1 else 0) / s if True else 0
                for i, nn in enumerate(n)])
            # tau_n = dP_n^1/dtheta via the recurrence n mu pi_n - (n+1) pi_{n-1}
            pim1 = np.concatenate([[0.], pi_n[:-1]])
            tau_n = n * mu * pi_n - (n + 1) * pim1
            fac = (2 * n + 1) / (n * (n + 1))
            S1 = np.sum(fac * (a * pi_n + b * tau_n))
            S2 = np.sum(fac * (a * tau_n + b * pi_n))
            z11 = (abs(S1) ** 2 + abs(S2) ** 2) / (2 * k * k)
            assert abs(Z[it, ip, 0, 0] - z11) < 1e-9 * z11, (it, ip)


def test_backscatter_identity():
    """Uniform grid with 4 polar angles: theta_out = pi/2, phi_out = pi is
    exactly backwards of the incidence (theta = pi/2, phi = 0)."""
    T = O.tmatrix_all(0.4, 0.6, 1.1, 1.5 + 0.02j, 12, 24)
    Z = O.scattering_grid(12, T, 1.1, 4, use_gl=False)
    Zb = Z[2, 4]
    assert abs(Zb[0, 0] - Zb[1, 1] + Zb[2, 2] - Zb[3, 3]) < 1e-10 * Zb[0, 0]
    # a planted sign error in Z22 breaks it
    assert abs(Zb[0, 0] + Zb[1, 1] + Zb[2, 2] - Zb[3, 3]) > 1e-3 * Zb[0, 0]


def test_orientation_coefficients():
    p = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0., 12)
    assert p[0] == 1. and np.abs(p[1:]).max() < 1e-9
    p = O.orientation_coefficients(O.ORIENT_MISHCHENKO, 0.4, 6)
    assert np.allclose(p, [1, 0, 0.1, 0, 0, 0, 0])
    p = O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, 0.5, 6)
    # prolate: p_2n = (-1)^n (2n-1)!!/(2n)!! = P_2n(0)
    from scipy.special import eval_legendre
    assert np.allclose(p, [eval_legendre(i, 0.) for i in range(7)])
    p = O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, 2.0, 6)
    assert np.allclose(p, 1.)


def test_isotropic_average_invariants():
    nmax = 8
    rv, d, lam, mr = 0.25, 0.55, 0.6, 1.5 + 0.02j
    T = O.tmatrix_all(rv, d, lam, mr, nmax, 2 * nmax)
    p = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0., 2 * nmax)
    Te = O.orientation_average(nmax, T, p)
    th = np.linspace(0.05, 3.0, 7)
    qe, qep, _ = O.extinction(nmax, Te, lam, rv, th)
    # isotropic ensemble: extinction independent of direction, no
    # dichroism, and equal to -(2 pi / k^2) times the T-matrix trace sum
    assert np.ptp(qe) < 1e-9 * abs(qe[0])
    assert np.abs(qep).max() < 1e-9 * abs(qe[0])
    k = 2 * np.pi / lam
    _, qtr = O.qcoeffs(nmax, T)
    assert abs(qe[0] + 2 * np.pi / k ** 2 * qtr / (np.pi * rv * rv)) < \
        1e-9 * abs(qe[0])
    # brute-force: the average over incidence directions of the single
    # grain's extinction equals the ensemble's (azimuthal symmetry)
    x, w = np.polynomial.legendre.leggauss(40)
    qf, _, _ = O.extinction(nmax, T, lam, rv, np.arccos(x))
    assert abs(0.5 * np.sum(w * qf) - qe[0]) < 1e-9 * abs(qe[0])


def test_sphere_average_is_identity():
    nmax = 6
    T = O.tmatrix_all(0.2, 1.0, 0.7, 1.4 + 0.01j, nmax, 12)
    for kind, par in ((0, 0.), (1, 0.5), (2, 0.7)):
        p = O.orientation_coefficients(kind, par, 2 * nmax)
        Te = O.orientation_average(nmax, T, p)
        assert np.allclose(Te, T, atol=1e-12 * np.abs(T).max())
