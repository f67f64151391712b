"""The oracle's aligned orientation average and the non-spherical
scattered-light term of Q_abs, pinned by brute force.

Ensemble T-matrix (AlignmentAverageTask.hpp:115-215, Mishchenko 1991 eq.
3.27): the S matrix of the ensemble T equals the orientation average of the
single grain's S over the particle orientations (alpha uniform, beta with
density p(beta)), because S is linear in T.  The reference checks this for
the extinction matrix K with a (alpha, beta) Gauss-Legendre quadrature
(testTMatrixCalculator.cpp:139-195).  Here it is checked for densities with
non-zero p_N at N > 0 (Mishchenko <cos^2 beta> = 0.2 and 0.6, and the two
perfect-alignment limits of Davis-Greenstein), where a wrong sign or weight
of the p_2 term changes the result.  The brute force uses
oracle.scattering_query, which the 101-row Mishchenko fixture pins
(tests/test_oracle_golden.py).  Z is quadratic in S, so only S and K
(linear) are compared.

Scattered light (AbsorptionCoefficientTask.hpp:742-773): Q_abs subtracts
from Q_ext the scattered light summed over incoming directions for the fixed
outgoing direction.  Checked on a prolate d = 0.5 grain against a brute-force
integral of Z11 (from scattering_query) in that geometry at any order, and,
at a converged order, against the physical absorption Q_ext - C_sca /
(pi R_V^2) with C_sca integrated over outgoing directions (equal by
reciprocity and the spheroid's inversion symmetry once the truncated
T-matrix is converged).
"""
import numpy as np
import pytest

import oracle as O

GRAIN = dict(rv=0.25, d=0.5, lam=0.6, mr=1.5 + 0.05j)


def _grain(nmax=8):
    g = GRAIN
    T = O.tmatrix_all(g["rv"], g["d"], g["lam"], g["mr"], nmax, 2 * nmax)
    return nmax, T


def _mishchenko_density(c2):
    """p(beta) of MishchenkoOrientationDistribution.hpp:83-96, normalised
    so that int_0^pi sin(beta) p(beta) d beta = 1."""
    p2 = 1.5 * c2 - 0.5
    return lambda b: 0.5 + 1.25 * p2 * (3. * np.cos(b) ** 2 - 1.)


def _brute_average(nmax, T, lam, density, th_in, ph_in, th_out, ph_out,
                   nb=48, na=36):
    """(S, K) averaged over alpha (uniform, trapezoid: exact for the
    trigonometric polynomial in alpha) and beta (Gauss-Legendre on [0, pi]
    with weight sin(beta) p(beta))."""
    xb, wb = np.polynomial.legendre.leggauss(nb)
    beta = 0.5 * np.pi * (xb + 1.)
    wb = 0.5 * np.pi * wb * np.sin(beta) * np.array([density(b)
                                                     for b in beta])
    alpha = 2. * np.pi * np.arange(na) / na
    S = np.zeros((2, 2), dtype=complex)
    K = np.zeros((4, 4))
    for b, w in zip(beta, wb):
        for a in alpha:
            s, _, k = O.scattering_query(nmax, T, lam, a, b, th_in, ph_in,
                                         th_out, ph_out)
            S += w / na * s
            K += w / na * k
    return S, K


def _ensemble(nmax, T, lam, pN, th_in, ph_in, th_out, ph_out):
    Te = O.orientation_average(nmax, T, pN)
    s, _, _ = O.scattering_query(nmax, Te, lam, 0., 0., th_in, ph_in,
                                 th_out, ph_out)
    _, _, k = O.scattering_query(nmax, Te, lam, 0., 0., th_in, ph_in,
                                 th_in, ph_in)
    return s, k


DIRS = (0.3 * np.pi, 0.2, 1.1, 2.3)  # theta_in, phi_in, theta_out, phi_out


@pytest.mark.parametrize("c2", [0.2, 0.6])
def test_mishchenko_average_matches_brute_force(c2):
    nmax, T = _grain()
    lam = GRAIN["lam"]
    pN = np.zeros(2 * nmax + 1)
    pN[0] = 1.
    pN[2] = 1.5 * c2 - 0.5
    th_in, ph_in, th_out, ph_out = DIRS
    S_e, K_e = _ensemble(nmax, T, lam, pN, *DIRS)
    S_b, _ = _brute_average(nmax, T, lam, _mishchenko_density(c2), *DIRS)
    _, K_b = _brute_average(nmax, T, lam, _mishchenko_density(c2), th_in,
                            ph_in, th_in, ph_in)
    scale_s = np.abs(S_b).max()
    scale_k = np.abs(K_b).max()
    assert np.abs(S_e - S_b).max() < 1e-9 * scale_s
    assert np.abs(K_e - K_b).max() < 1e-9 * scale_k
    # a planted sign error on the N = 2 term of the sum fails the check
    bad = pN.copy()
    bad[2] = -bad[2]
    S_x, K_x = _ensemble(nmax, T, lam, bad, *DIRS)
    assert np.abs(S_x - S_b).max() > 1e-4 * scale_s
    assert np.abs(K_x - K_b).max() > 1e-4 * scale_k
    # so does a planted factor 2 on it
    bad = pN.copy()
    bad[2] *= 2.
    S_x, _ = _ensemble(nmax, T, lam, bad, *DIRS)
    assert np.abs(S_x - S_b).max() > 1e-4 * scale_s


def test_davis_greenstein_prolate_is_perpendicular_alignment():
    """Davis-Greenstein for a prolate grain: p_2n = (-1)^n (2n-1)!!/(2n)!!
    = P_2n(0), the coefficients of a density concentrated at beta = pi/2
    (symmetry axis perpendicular to z), uniform in alpha."""
    nmax, T = _grain()
    lam = GRAIN["lam"]
    pN = O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, 0.5,
                                    2 * nmax)
    S_e, K_e = _ensemble(nmax, T, lam, pN, *DIRS)
    na = 4 * nmax + 8
    S_b = np.zeros((2, 2), dtype=complex)
    K_b = np.zeros((4, 4))
    th_in, ph_in, th_out, ph_out = DIRS
    for a in 2. * np.pi * np.arange(na) / na:
        s, _, _ = O.scattering_query(nmax, T, lam, a, 0.5 * np.pi, *DIRS)
        _, _, k = O.scattering_query(nmax, T, lam, a, 0.5 * np.pi, th_in,
                                     ph_in, th_in, ph_in)
        S_b += s / na
        K_b += k / na
    assert np.abs(S_e - S_b).max() < 1e-9 * np.abs(S_b).max()
    assert np.abs(K_e - K_b).max() < 1e-9 * np.abs(K_b).max()


def test_davis_greenstein_oblate_is_full_alignment():
    """Oblate: every p_N = 1, the density concentrated at beta = 0; the
    ensemble of axially symmetric grains aligned with z is the grain."""
    nmax, T = _grain()
    pN = O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, 2.0, 2 * nmax)
    assert np.all(pN == 1.)
    Te = O.orientation_average(nmax, T, pN)
    assert np.abs(Te - T).max() < 1e-10 * np.abs(T).max()


def test_isotropic_average_matches_brute_force():
    """p(beta) = 1/2: the reference's own check (K at theta = 0.3 pi,
    testTMatrixCalculator.cpp:155-195), here with S as well."""
    nmax, T = _grain()
    lam = GRAIN["lam"]
    pN = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0., 2 * nmax)
    S_e, K_e = _ensemble(nmax, T, lam, pN, *DIRS)
    S_b, _ = _brute_average(nmax, T, lam, lambda b: 0.5, *DIRS)
    th_in, ph_in = DIRS[:2]
    _, K_b = _brute_average(nmax, T, lam, lambda b: 0.5, th_in, ph_in, th_in,
                            ph_in)
    # the isotropic p_N (N > 0) come from an adaptive quadrature with
    # tolerance 1e-10 / 1e-5, as in the reference
    assert np.abs(S_e - S_b).max() < 1e-7 * np.abs(S_b).max()
    assert np.abs(K_e - K_b).max() < 1e-7 * np.abs(K_b).max()


def _csca(nmax, T, th, incoming, nb=40, nphi=40):
    """Z11 integrated over the sphere of directions (GL in cos theta' x
    uniform phi'): over outgoing directions for incidence (th, 0), or, with
    `incoming`, over incoming directions for the outgoing direction (th, 0)
    (the reference's geometry)."""
    xo, wo = np.polynomial.legendre.leggauss(nb)
    lam = GRAIN["lam"]
    c = 0.
    for x, w in zip(xo, wo):
        for ph in 2. * np.pi * np.arange(nphi) / nphi:
            d = (th, 0., np.arccos(x), ph)
            if incoming:
                d = (np.arccos(x), ph, th, 0.)
            _, Z, _ = O.scattering_query(nmax, T, lam, 0., 0., *d)
            c += w * (2. * np.pi / nphi) * Z[0, 0]
    return c


@pytest.mark.parametrize("nmax", [8, 18])
def test_nonspherical_absorption_subtracts_scattered_light(nmax):
    """Q_abs(theta) = Q_ext(theta) minus the scattered light over incoming
    directions for the outgoing direction theta, divided by pi R_V^2: the
    oracle (GL x GL quadrature over incoming directions) against a brute
    force (GL x uniform phi) built on scattering_query, at any order."""
    g = GRAIN
    T = O.tmatrix_all(g["rv"], g["d"], g["lam"], g["mr"], nmax, 2 * nmax)
    thetas = np.array([0.25, 1.0, 0.5 * np.pi])
    qa, _ = O.absorption(nmax, T, g["lam"], g["rv"], thetas, ngauss=40)
    qe, _, _ = O.extinction(nmax, T, g["lam"], g["rv"], thetas)
    area = np.pi * g["rv"] ** 2
    for j, th in enumerate(thetas):
        ci = _csca(nmax, T, th, incoming=True)
        assert abs(qa[j] - (qe[j] - ci / area)) < 1e-12 * abs(qe[j])
        # the scattered light is a sizeable part of the extinction here, so
        # a dropped or doubled term cannot pass
        assert ci / area > 0.05 * qe[j]


def test_absorption_is_extinction_minus_csca_at_convergence():
    """At a converged order the reference's geometry gives the physical
    absorption: by reciprocity and the spheroid's inversion symmetry the
    scattered light over incoming directions equals C_sca for incidence
    along theta (Z11 over outgoing directions).  Reciprocity holds for the
    truncated EBCM T-matrix only to its convergence (6e-5 at order 8 for
    this grain), so the identity is checked at order 18."""
    nmax = 18
    g = GRAIN
    T = O.tmatrix_all(g["rv"], g["d"], g["lam"], g["mr"], nmax, 2 * nmax)
    thetas = np.array([0.25, 1.0, 0.5 * np.pi])
    qa, _ = O.absorption(nmax, T, g["lam"], g["rv"], thetas, ngauss=40)
    qe, _, _ = O.extinction(nmax, T, g["lam"], g["rv"], thetas)
    area = np.pi * g["rv"] ** 2
    for j, th in enumerate(thetas):
        co = _csca(nmax, T, th, incoming=False)
        assert abs(qa[j] - (qe[j] - co / area)) < 1e-10 * abs(qe[j])
