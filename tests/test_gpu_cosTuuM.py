"""The CosTuuM module surface on the GPU: TMatrix (two-phase convergence,
orientation averaging, cross sections, extinction and scattering matrices)
and get_table (orientation, shape and size axes), against the reference's
golden fixture and the CPU oracle, FP64."""
import math
import os

import numpy as np
import torch
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROWS = np.loadtxt(os.path.join(HERE, "golden",
                               "mishchenko_tmatrixcalculator.txt"),
                  comments="#")


def rel(a, b):
    return abs(a - b) / abs(a + b)


def relmax(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("row", list(range(0, 101, 10)))
def test_tmatrix_matches_fixture_and_oracle(row):
    """TMatrix with alignment off reproduces the fixture's Q_ext, Q_sca
    (1e-5) and Z (2e-2) (testTMatrixCalculator.cpp:64-136) and the oracle's
    orders and T-matrix (1e-8)."""
    from costuum_amd import CosTuuM as C
    r = ROWS[row]
    axi, rat, lam, mrr, mri, eps, ddelt, ndgs = r[:8]
    alpha, beta, th0, th, ph0, ph = np.radians(r[8:14])
    mr = complex(mrr, mri)
    print("fixture row %d" % row, flush=True)
    tm = C.TMatrix(axi, eps, lam, mr, cos2beta=-1.0, tolerance=ddelt * 0.1,
                   maximum_order=200, gauss_legendre_factor=int(ndgs),
                   maximum_ngauss=500,
                   is_equal_volume_radius=abs(rat - 1.0) <= 1e-8)
    st, n, g, rv = O.converge_calculator(rat, eps, axi, lam, mr, 200,
                                         ddelt * 0.1, int(ndgs), 500)
    assert st == 0
    assert (tm.get_nmax(), tm.get_ngauss()) == (n, g)
    T = tm.get_tmatrix()
    Tref = O.tmatrix_all(rv, eps, lam, mr, n, g)
    # Strongly oblate rows are ill-conditioned: the oracle itself moves by
    # up to 1e-5 when R_V changes by one rounding, so two FP64 solves that
    # round differently differ by ~cond(Q) eps.  The bar is 1e-8 or 50 times
    # that measured sensitivity, whichever is larger.
    T2 = O.tmatrix_all(rv * (1 + 4e-16), eps, lam, mr, n, g)
    bar = max(1e-8, 50 * relmax(T2, Tref))
    assert relmax(T, Tref) < bar, bar
    qs, qe = O.qcoeffs(n, T)
    assert rel(qe, r[15]) < 1e-5 and rel(qs, r[14]) < 1e-5
    Z = tm.get_scattering_matrix(th0, ph0, th, ph, alpha, beta)
    Zref = r[17:33].reshape(4, 4)
    assert (np.abs(Z - Zref) / np.abs(Z + Zref)).max() < 2e-2
    _, Zo, _ = O.scattering_query(n, Tref, lam, alpha, beta, th0, ph0, th,
                                  ph)
    _, Z2, _ = O.scattering_query(n, T2, lam, alpha, beta, th0, ph0, th, ph)
    assert relmax(Z, Zo) < max(1e-9, 50 * relmax(Z2, Zo))
    K = tm.get_extinction_matrix(th0, ph0, alpha, beta)
    _, _, Ko = O.scattering_query(n, Tref, lam, alpha, beta, th0, ph0, th0,
                                  ph0)
    _, _, K2 = O.scattering_query(n, T2, lam, alpha, beta, th0, ph0, th0,
                                  ph0)
    assert relmax(K, Ko) < max(1e-9, 50 * relmax(K2, Ko))
    tm.close()


@pytest.mark.parametrize("cos2beta", [1.0 / 3.0, 0.4, 0.0])
def test_tmatrix_orientation_average_and_cross_sections(cos2beta):
    from costuum_amd import CosTuuM as C
    rv, d, lam, mr = 0.2, 0.5, 0.55, 1.5 + 0.02j
    print("TMatrix cos2beta %.3f" % cos2beta, flush=True)
    tm = C.TMatrix(rv, d, lam, mr, cos2beta=cos2beta)
    n, g = tm.get_nmax(), tm.get_ngauss()
    st, n_o, g_o, rv_o = O.converge_calculator(1.0, d, rv, lam, mr, 200,
                                               1e-4, 2, 500)
    assert (n, g) == (n_o, g_o)
    T = O.tmatrix_all(rv, d, lam, mr, n, g)
    if cos2beta == 1.0 / 3.0:
        pN = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0.0, 2 * n)
    elif cos2beta == 0.0:
        pN = O.orientation_coefficients(O.ORIENT_DAVIS_GREENSTEIN, d, 2 * n)
    else:
        pN = O.orientation_coefficients(O.ORIENT_MISHCHENKO, cos2beta, 2 * n)
    Tavg = O.orientation_average(n, T, pN)
    assert relmax(tm.get_tmatrix(), Tavg) < 1e-9
    theta = np.array([0.0, 0.5, math.pi / 6, 1.3])
    qa, qap = O.absorption(n, Tavg, lam, rv, theta, ngauss=30)
    area = math.pi * rv * rv
    c = tm.get_absorption_cross_section(theta, ngauss=30)
    assert relmax(c, qa * area) < 1e-9
    cs = tm.get_absorption_cross_sections(theta, ngauss=30)
    assert cs.shape == (4, 2)
    # C_abspol vanishes for random orientation: scale by C_abs
    assert np.abs(cs[:, 1] - qap * area).max() < \
        1e-9 * np.abs(qa * area).max()
    assert tm.get_absorption_cross_sections(0.5, ngauss=30).shape == (2,)
    x, w = np.polynomial.legendre.leggauss(20)
    qa20, _ = O.absorption(n, Tavg, lam, rv, np.arccos(x), ngauss=20)
    avg = 0.5 * np.dot(qa20 * area, w)
    assert abs(tm.get_average_absorption_cross_section() - avg) < \
        1e-9 * abs(avg)
    # shapes of the matrix queries (squeezed like the reference)
    assert tm.get_extinction_matrix([0.1, 0.2, 0.3], 0.0).shape == (3, 4, 4)
    assert tm.get_scattering_matrix(0.3, [0.0, 1.0], [0.5, 1.0, 2.0],
                                    0.2).shape == (2, 3, 4, 4)
    K = tm.get_extinction_matrix(0.7, 0.4, 0.3, 0.2)
    _, _, Ko = O.scattering_query(n, Tavg, lam, 0.3, 0.2, 0.7, 0.4, 0.7, 0.4)
    assert relmax(K, Ko) < 1e-9
    if cos2beta == 1.0 / 3.0:
        # random orientation: K is diagonal and independent of direction
        # (testTMatrixCalculator.cpp:215-226)
        K2 = tm.get_extinction_matrix(2.0, 1.0)
        assert abs(K[0, 0] - K2[0, 0]) < 1e-7 * abs(K[0, 0])
        assert np.abs(K - np.diag(np.diag(K))).max() < 1e-7 * abs(K[0, 0])
    tm.close()


def _oracle_table(types, sizes, lams, thetas, shape, align, dust, min_order,
                  max_order, do_abs, do_ext, do_sca, afs, ngauss):
    """get_table restated on the oracle: per grain task-path convergence,
    orientation average, observables, shape average."""
    from costuum_amd import CosTuuM as C
    sizes = np.sort(sizes)
    lams = np.sort(lams)
    ns = shape.get_number_of_points()
    out_q, out_z = [], []
    for t in types:
        for a in sizes:
            for lam in lams:
                mr = dust.get_refractive_index(t, a, lam)
                qs, zs, ws = [], [], []
                for i in range(ns):
                    d = shape.get_shape(i)
                    st, n, g, _, _ = O.converge_task(a, d, lam, mr, min_order,
                                                     max_order, 2, 1e-4)
                    T = O.tmatrix_all(a, d, lam, mr, n, g)
                    od = align.get_distribution(a, d)
                    if od.compute_alignment:
                        T = O.orientation_average(n, T,
                                                  od.get_coefficients())
                    cols = []
                    if do_abs:
                        cols += list(O.absorption(n, T, lam, a, thetas,
                                                  ngauss=ngauss,
                                                  account_for_scattering=afs))
                    if do_ext:
                        cols += list(O.extinction(n, T, lam, a, thetas))
                    qs.append(np.stack(cols, axis=1) if cols else None)
                    if do_sca:
                        zs.append(O.scattering_grid(n, T, lam, len(thetas)))
                    ws.append(shape.get_weight(i))
                ws = np.array(ws)
                if qs[0] is not None:
                    out_q.append(sum(w * q for w, q in zip(ws, qs)) *
                                 (1.0 / ws.sum()))
                if do_sca:
                    out_z.append(sum(w * z for w, z in zip(ws, zs)) *
                                 (1.0 / ws.sum()))
    return out_q, out_z


def test_get_table_matches_oracle():
    from costuum_amd import CosTuuM as C
    types = [C.SILICON, C.CARBON_PARALLEL]
    sizes = [0.25, 0.08]          # unsorted on purpose
    lams = [1.1, 0.55]
    thetas = np.array([0.3, 1.2, 2.0])
    shape = C.DraineHensleyShapeDistribution(3)
    align = C.SizeBasedAlignmentDistribution(0.1, C.MISHCHENKO_ALIGNMENT, 50)
    dust = C.CustomDustProperties(
        lambda lam, a, t: complex(1.4 + 0.1 * t + 0.05 * lam, 0.02 + 0.01 * t))
    print("get_table 2 x 2 x 2 x 3 shapes", flush=True)
    q, z = C.get_table(types, sizes, lams, thetas, shape, align, dust,
                       minimum_order=4, maximum_order=60,
                       do_absorption=True, do_extinction=True,
                       do_scattering=True, account_for_scattering=True,
                       number_of_quadrature_angles=12)
    assert q.shape == (2, 2, 2, 3, 5)
    assert z.shape == (2, 2, 2, 3, 6, 4, 4)
    rq, rz = _oracle_table(types, sizes, lams, thetas, shape, align, dust, 4,
                           60, True, True, True, True, 12)
    rq = np.array(rq).reshape(q.shape)
    rz = np.array(rz).reshape(z.shape)
    for k in range(5):
        assert relmax(q[..., k], rq[..., k]) < 1e-8, k
    assert relmax(z, rz) < 1e-8


def test_get_table_memory_budget_streams_batches():
    """maximum_memory_size bounds each batch: with a budget below one
    interaction's estimate every (composition, size, wavelength) becomes its
    own batch, and the table is the same as in one batch."""
    from costuum_amd import CosTuuM as C
    shape = C.DraineHensleyShapeDistribution(3)
    align = C.SizeBasedAlignmentDistribution(0.1, C.MISHCHENKO_ALIGNMENT, 50)
    dust = C.CustomDustProperties(lambda lam, a, t: complex(1.5, 0.03))
    args = ([C.SILICON], [0.25, 0.08, 0.15], [1.1, 0.55], [0.3, 1.2], shape,
            align, dust)
    kw = dict(minimum_order=4, maximum_order=60, do_absorption=True,
              do_extinction=True, account_for_scattering=True,
              number_of_quadrature_angles=8)
    inter = [(a, lam, 1.5 + 0.03j) for a in (0.08, 0.15, 0.25)
             for lam in (0.55, 1.1)]
    shapes = [shape.get_shape(i) for i in range(3)]
    dev = torch.device("cuda", 0)
    assert len(C._memory_batches(inter, shapes, 4, 60, 18, 1.0, dev)) == 6
    assert len(C._memory_batches(inter, shapes, 4, 60, 18, 1e10, dev)) == 1
    print("get_table with a one-byte budget", flush=True)
    small = C.get_table(*args, maximum_memory_size=1, **kw)
    big = C.get_table(*args, **kw)
    assert small.shape == big.shape == (3, 2, 2, 5)
    assert np.abs(small - big).max() <= 1e-13 * np.abs(big).max()


def test_get_table_disabled_alignment_and_squeeze():
    from costuum_amd import CosTuuM as C
    shape = C.SingleShapeShapeDistribution(2.0)
    align = C.SizeBasedAlignmentDistribution(0.01, C.DISABLE_ALIGNMENT)
    dust = C.CustomDustProperties(lambda lam, a, t: 1.5 + 0.02j)
    thetas = np.array([0.4, 1.0])
    q = C.get_table([C.SILICON], [0.2], [0.55], thetas, shape, align, dust,
                    minimum_order=4, maximum_order=60)
    # default: absorption only, no scattered-light correction; squeezed to
    # [ntheta, 2]
    assert q.shape == (2, 2)
    st, n, g, _, _ = O.converge_task(0.2, 2.0, 0.55, 1.5 + 0.02j, 4, 60, 2,
                                     1e-4)
    T = O.tmatrix_all(0.2, 2.0, 0.55, 1.5 + 0.02j, n, g)
    qa, qap = O.absorption(n, T, 0.55, 0.2, thetas, ngauss=20,
                           account_for_scattering=False)
    assert relmax(q[:, 0], qa) < 1e-9 and relmax(q[:, 1], qap) < 1e-9


@pytest.mark.parametrize("d", [0.5, 2.0])
def test_single_grain_workload(d):
    """The single-grain case: R_V = 0.2 um, lambda = 0.55 um, m = 1.5 +
    0.02i, isotropic ensemble; Q_ext, Q_abs at theta = 30 deg and Z at 91
    scattering angles 0, 2, ..., 180 deg (incidence along z, phi = 0)."""
    from costuum_amd import CosTuuM as C
    rv, lam, mr = 0.2, 0.55, 1.5 + 0.02j
    tm = C.TMatrix(rv, d, lam, mr, cos2beta=1.0 / 3.0)
    n, g = tm.get_nmax(), tm.get_ngauss()
    T = O.tmatrix_all(rv, d, lam, mr, n, g)
    pN = O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0.0, 2 * n)
    Tavg = O.orientation_average(n, T, pN)
    th = np.radians(30.0)
    area = math.pi * rv * rv
    qa, _ = O.absorption(n, Tavg, lam, rv, [th], ngauss=100)
    assert abs(tm.get_absorption_cross_section(th)[0] - qa[0] * area) < \
        1e-9 * abs(qa[0] * area)
    K = tm.get_extinction_matrix(th, 0.0)
    qe, _, _ = O.extinction(n, Tavg, lam, rv, [th])
    assert abs(K[0, 0] - qe[0] * area) < 1e-9 * abs(qe[0] * area)
    Theta = np.radians(np.arange(0, 181, 2.0))
    Z = tm.get_scattering_matrix(0.0, 0.0, Theta, 0.0)
    assert Z.shape == (91, 4, 4)
    for i in (0, 1, 45, 89, 90):
        _, Zo, _ = O.scattering_query(n, Tavg, lam, 0.0, 0.0, 0.0, 0.0,
                                      Theta[i], 0.0)
        assert relmax(Z[i], Zo) < 1e-9, i
    # backscattering invariant Z11 - Z22 + Z33 - Z44 = 0
    # (testTMatrixCalculator.cpp:198-212)
    zb = Z[90]
    assert abs(zb[0, 0] - zb[1, 1] + zb[2, 2] - zb[3, 3]) < 1e-9 * zb[0, 0]
    tm.close()
