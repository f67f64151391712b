"""Host side of the CosTuuM module surface (no GPU): Gauss-Legendre rules,
R_V/R_S, orientation-distribution coefficients, the Draine-Hensley shape
distribution and the alignment selection, each pinned by closed forms or the
oracle."""
import math
import os

import numpy as np
import pytest
from scipy import special

import oracle as O
from costuum_amd import CosTuuM as C


@pytest.mark.parametrize("order", [1, 2, 3, 5, 8, 20, 41, 100, 400])
def test_gauss_legendre_matches_oracle_and_numpy(order):
    x, w = C.gauss_legendre(order)
    xo, wo = O.gauss_legendre(order)
    assert np.array_equal(x, xo) and np.array_equal(w, wo)
    xn, wn = np.polynomial.legendre.leggauss(order)
    assert np.abs(x - xn).max() < 1e-14
    assert np.abs(w - wn).max() < 1e-13
    xa, wa = C.gauss_legendre(order, 0.0, np.pi)
    assert np.allclose(xa, 0.5 * np.pi * (xn + 1.0), atol=1e-14)
    assert abs(wa.sum() - np.pi) < 1e-12


def test_gauss_legendre_closed_form_order3():
    # nodes 0, +-sqrt(3/5), weights 8/9, 5/9 (testSpecialFunctions.cpp:204-304)
    x, w = C.gauss_legendre(3)
    assert np.allclose(x, [-math.sqrt(0.6), 0.0, math.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w, [5 / 9, 8 / 9, 5 / 9], atol=1e-15)


def test_equal_volume_radius_known_values():
    # testSpecialFunctions.cpp:190-200
    assert abs(C.get_equal_volume_radius(1.0, 0.5) - 0.9637112829756893) \
        < 1e-15
    assert abs(C.get_equal_volume_radius(1.0, 2.0) - 0.955443263377723) \
        < 1e-15
    assert C.get_equal_volume_radius(2.0, 1.0) == 2.0
    assert C.get_equal_volume_radius(3.0, 0.7) == 3.0 * O.rv_over_rs(0.7)


def test_isotropic_coefficients_vanish_and_match_oracle():
    od = C.RandomOrientationDistribution(40)
    p = od.get_coefficients()
    po = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0.0, 40)
    assert p[0] == 1.0
    assert np.abs(p[1:]).max() < 1e-10
    assert np.abs(p - po).max() < 1e-15
    # p(theta) = 1/2 from the expansion
    assert np.allclose(od.evaluate(np.linspace(0.1, 3.0, 7)), 0.5,
                       atol=1e-8)


def test_mishchenko_closed_form_and_evaluate():
    od = C.MishchenkoOrientationDistribution(0.4, 10)
    p2 = 1.5 * 0.4 - 0.5
    assert od.get_coefficient(0) == 1.0 and od.get_coefficient(2) == p2
    assert od.get_coefficient(1) == 0.0 and od.get_coefficient(4) == 0.0
    th = np.linspace(0.05, 3.1, 9)
    ref = 0.5 + 1.25 * p2 * (3 * np.cos(th) ** 2 - 1)
    assert np.allclose(od.evaluate(th), ref, atol=1e-14)
    with pytest.raises(ValueError):
        C.MishchenkoOrientationDistribution(0.7)


def test_davis_greenstein_is_legendre_at_zero():
    prolate = C.DavisGreensteinOrientationDistribution(0.5, 12).\
        get_coefficients()
    ref = np.array([special.eval_legendre(n, 0.0) for n in range(13)])
    assert np.abs(prolate - ref).max() < 1e-15
    oblate = C.DavisGreensteinOrientationDistribution(2.0, 12)
    assert np.all(oblate.get_coefficients() == 1.0)
    assert np.array_equal(
        prolate, O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, 0.5,
                                            12))


def test_custom_distribution_recovers_mishchenko():
    p2 = 1.5 * 0.3 - 0.5

    def density(beta, scale):
        return scale * (0.5 + 1.25 * p2 * (3 * math.cos(beta) ** 2 - 1))

    od = C.CustomOrientationDistribution(density, 8,
                                         additional_arguments=3.0)
    p = od.get_coefficients()
    # the normalisation removes the scale; the expansion gives p_2 back to
    # the quadrature tolerance of custom distributions (1e-4)
    assert p[0] == 1.0
    assert abs(p[2] - p2) < 1e-5 + 1e-3 * abs(p2)
    assert np.abs(np.delete(p, [0, 2])).max() < 1e-4


def test_custom_distribution_error_propagates():
    def bad(beta):
        raise RuntimeError("boom")
    with pytest.raises(RuntimeError):
        C.CustomOrientationDistribution(bad, 4)


def test_draine_hensley_pdf_normalised_and_jacobian():
    # the CDE2 density is normalised over all shapes: int_0^inf p(d) dd = 1
    tot = C.adaptive_quadrature(C._dh_cde2, 1e-9, 1.0) + \
        C.adaptive_quadrature(C._dh_cde2, 1.0, 200.0, 10, 4000)
    assert abs(tot - 1.0) < 1e-3
    # the Jacobian is |dL/dd| of the depolarisation factor
    for d in (0.3, 0.6, 0.95, 1.05, 1.7, 3.0):
        h = 1e-6
        num = abs(C._dh_shape_factor(d + h) - C._dh_shape_factor(d - h)) / \
            (2 * h)
        assert abs(C._dh_jacobian(d) - num) < 1e-7 * max(num, 1.0)
    assert C._dh_shape_factor(1.0) == 1.0 / 3.0
    # sphere limits of L and the series branches of the Jacobian
    assert abs(C._dh_shape_factor(0.999999) - 1 / 3) < 1e-6
    assert abs(C._dh_jacobian(0.999999) - 4 / 15) < 1e-5


def test_draine_hensley_cutoff_limits():
    dh = C.DraineHensleyShapeDistribution(20, 0.15)
    dmin, dmax = dh.get_limits()
    assert dmin < 1.0 < dmax
    # bisection to 0.5 % of the bracket: the density is near the cutoff
    assert abs(C._dh_cde2(dmin) - 0.15) < 0.02
    assert abs(C._dh_cde2(dmax) - 0.15) < 0.02
    assert C._dh_cde2(0.99 * dmin) < 0.15 + 0.02
    assert dh.get_number_of_points() == 20
    x, w = C.gauss_legendre(20, dmin, dmax)
    assert np.array_equal([dh.get_shape(i) for i in range(20)], x)
    assert np.allclose([dh.get_weight(i) for i in range(20)],
                       w * np.array([C._dh_cde2(d) for d in x]), rtol=0,
                       atol=0)


def test_draine_hensley_fraction_limits():
    dh = C.DraineHensleyShapeDistribution(10, 0.0, 0.9)
    dmin, dmax = dh.get_limits()
    prolate = C.adaptive_quadrature(C._dh_cde2, 0.0, 1.0)
    got = C.adaptive_quadrature(C._dh_cde2, dmin, 1.0)
    assert abs(got - 0.9 * prolate) <= 0.005 * (got + 0.9 * prolate)
    oblate = C.adaptive_quadrature(C._dh_cde2, 1.0, dmax)
    target = 0.9 * (1.0 - prolate)
    assert abs(oblate - target) <= 1e-5 * (oblate + target)


def test_single_shape_and_base_distribution():
    s = C.SingleShapeShapeDistribution(1.7)
    assert s.get_limits() == (1.7, 1.7)
    assert s.get_number_of_points() == 1 and s.get_weight(0) == 1.0
    b = C.ShapeDistribution()
    b.evaluate(5)
    x, w = C.gauss_legendre(5, 0.5, 1.5)
    assert np.allclose([b.get_weight(i) for i in range(5)],
                       w * np.exp(-0.5 * (x - 1) ** 2 / 0.027777778))


def test_size_based_alignment_selection():
    a = C.SizeBasedAlignmentDistribution(0.1, C.MISHCHENKO_ALIGNMENT, 5)
    big_oblate = a.get_distribution(0.2, 2.0)
    big_prolate = a.get_distribution(0.2, 0.5)
    small = a.get_distribution(0.05, 2.0)
    assert big_oblate.get_coefficient(2) == 1.5 * 0.6 - 0.5
    assert big_prolate.get_coefficient(2) == 1.5 * 0.2 - 0.5
    assert small.get_maximum_order() == 10
    assert abs(small.get_coefficient(2)) < 1e-10
    d = C.SizeBasedAlignmentDistribution(0.1, C.DISABLE_ALIGNMENT)
    assert not d.get_distribution(1.0, 2.0).compute_alignment
    assert d.get_distribution(0.01, 2.0).compute_alignment
    with pytest.raises(ValueError):
        C.SizeBasedAlignmentDistribution(0.1, C.CUSTOM_ALIGNMENT)
    with pytest.raises(ValueError):
        C.SizeBasedAlignmentDistribution(0.1, 7)


def test_custom_dust_properties():
    dust = C.CustomDustProperties(lambda lam, a, t, k: k * (1.5 + 0.01j * t),
                                  2.0)
    assert dust.get_refractive_index(C.SILICON, 1e-6, 5e-7) == 3.0 + 0.04j
    bad = C.CustomDustProperties(lambda lam, a, t: 1.5)
    with pytest.raises(ValueError):
        bad.get_refractive_index(0, 1.0, 1.0)
    arr = C.CustomDustProperties(lambda lam, a, t: np.array([1.5 + 1j,
                                                             2.0 + 0j]))
    with pytest.raises(ValueError):
        arr.get_refractive_index(0, 1.0, 1.0)
    import os
    if not os.environ.get("COSTUUM_DRAINE_DATA"):
        with pytest.raises(NotImplementedError):
            C.DraineDustProperties()


def test_module_constants():
    assert (C.DAVIS_GREENSTEIN_ALIGNMENT, C.MISHCHENKO_ALIGNMENT,
            C.DISABLE_ALIGNMENT, C.CUSTOM_ALIGNMENT) == (0, 1, 2, 3)
    assert (C.CARBON_PARALLEL, C.CARBON_PERPENDICULAR, C.SILICON) == (0, 1, 2)


DRAINE_DATA = "/root/reference/data"


@pytest.mark.skipif(not os.path.isdir(DRAINE_DATA),
                    reason="Draine's callindex tables are not available")
def test_draine_dust_properties_against_tables():
    d = C.DraineDustProperties(20.0, DRAINE_DATA)
    sil = np.array([[float(v) for v in l.split()[:5]]
                    for l in open(os.path.join(DRAINE_DATA,
                                               "callindex.out_silD03"))
                    if l.split() and not l.startswith("#")])
    # at a table node the silicate index is the tabulated one
    i = 300
    lam = sil[i, 0] * 1e-6
    m = d.get_refractive_index(C.SILICON, 1e-7, lam)
    assert m == complex(sil[i, 3] + 1.0, sil[i, 4])
    # halfway between two nodes: the average (linear interpolation)
    mid = 0.5 * (sil[i, 0] + sil[i + 1, 0]) * 1e-6
    mm = d.get_refractive_index(C.SILICON, 1e-7, mid)
    assert abs(mm - 0.5 * (complex(sil[i, 3] + 1, sil[i, 4]) +
                           complex(sil[i + 1, 3] + 1, sil[i + 1, 4]))) < 1e-12
    # graphite at the tables' own temperature and size: the correction
    # cancels, and m = sqrt(eps) (principal branch)
    cpa = np.array([[float(v) for v in l.split()[:5]]
                    for l in open(os.path.join(DRAINE_DATA,
                                               "callindex.out_CpaD03_0.10"))
                    if l.split() and not l.startswith("#")])
    j = 200
    lam = cpa[j, 0] * 1e-6
    m = d.get_refractive_index(C.CARBON_PARALLEL, 1e-7, lam)
    eps = complex(cpa[j, 1] + 1.0, cpa[j, 2])
    assert abs(m - np.sqrt(eps)) < 1e-9 * abs(m)
    # the table's own n agrees with sqrt(eps) to its printed precision
    assert abs(m - complex(cpa[j, 3] + 1, cpa[j, 4])) < 2e-3 * abs(m)
    with pytest.raises(ValueError):
        d.get_refractive_index(7, 1e-7, lam)
    with pytest.raises(ValueError):
        d.get_refractive_index(C.SILICON, 1e-7, 10.0)


def test_drude_correction_closed_form():
    # -(w_p tau)^2 / ((w tau)^2 + i w tau) = -w_p^2 / (w^2 + i w / tau)
    for par in (True, False):
        for T, lam, a in ((20.0, 1e-6, 1e-7), (100.0, 3e-5, 5e-9)):
            c = C._carbon_free_electrons(T, lam, a, par)
            if par:
                wp, tb, TF, vF = 1.53e14, 1.4e-14, 255.0, 3.7e4
            else:
                wp = 4.33e14 * math.sqrt(1 - 6.24e-3 * T + 3.66e-5 * T * T)
                tb = 4.2e-11 / (1 + 0.322 * T + 0.0013 * T * T)
                TF, vF = 255.0, 4.5e5
            tau = 1 / (1 / tb + vF * math.sqrt(1 + T / TF) / a)
            w = 2 * math.pi * 299792458.0 / lam
            ref = -wp * wp / complex(w * w, w / tau)
            assert abs(c - ref) < 1e-12 * abs(ref)
