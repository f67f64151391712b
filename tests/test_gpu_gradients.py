"""Derivatives with respect to (R_V, d, Re m, Im m) at frozen orders:
the library's forward-mode tangents against central differences of the GPU
forward path and of the CPU oracle (relative step 1e-5), FP64."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# radius, d, wavelength, m, nmax, ngauss
CASES = [
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 10, 20),
    (0.3, 1.6, 0.7, 1.7 + 0.1j, 10, 20),
    (0.05, 0.8, 2.0, 1.9 + 0.3j, 5, 10),
    (0.5, 1.2, 0.9, 1.33 + 0.001j, 12, 24),
]
H = 1e-5


def _params(c):
    rv, d, lam, mr, n, g = c
    return np.array([rv, d, mr.real, mr.imag])


def _shift(cases, j, sign):
    out = []
    for c in cases:
        p = _params(c)
        p[j] += sign * H * max(abs(p[j]), 1e-3)
        out.append((p[0], p[1], c[2], complex(p[2], p[3]), c[4], c[5]))
    return out


def _step(c, j):
    return 2 * H * max(abs(_params(c)[j]), 1e-3)


@pytest.fixture(scope="module")
def ctx():
    from costuum_amd.tmatrix import Context
    c = Context(0)
    yield c
    c.close()


def _fixed(ctx, cases):
    from costuum_amd.tmatrix import tmatrix_fixed
    return tmatrix_fixed([c[0] for c in cases], [c[1] for c in cases],
                         [c[2] for c in cases], [c[3].real for c in cases],
                         [c[3].imag for c in cases], [c[4] for c in cases],
                         [c[5] for c in cases], ctx=ctx)


def test_tangent_tmatrix_matches_central_differences(ctx):
    from costuum_amd.tmatrix import tangent_tmatrix, tangents
    print("tangent T-matrices, %d grains" % len(CASES), flush=True)
    fd = {}
    for j in range(4):
        rp = _fixed(ctx, _shift(CASES, j, +1))
        Tp = [rp.tmatrix(g) for g in range(len(CASES))]
        rm = _fixed(ctx, _shift(CASES, j, -1))
        Tm = [rm.tmatrix(g) for g in range(len(CASES))]
        fd[j] = [(a - b) / _step(CASES[g], j)
                 for g, (a, b) in enumerate(zip(Tp, Tm))]
    res = _fixed(ctx, CASES)
    tangents(res)
    for j in range(4):
        for g in range(len(CASES)):
            dT = tangent_tmatrix(res, g, j)
            err = np.abs(dT - fd[j][g]).max() / np.abs(fd[j][g]).max()
            assert err < 1e-5, (j, g, err)


# grains near their natural convergence at orders 45, 68 and 110 (size
# parameters 32, 50 and 90, near-spherical so that differences are well
# conditioned): the blocked elimination in LDS, with two rows per lane, and
# in global scratch (blocks above order 99)
HIGH = [
    (2.8, 1.15, 0.55, 1.33 + 0.005j, 45, 90),
    (4.4, 0.9, 0.55, 1.4 + 0.01j, 68, 136),
    (7.9, 1.05, 0.55, 1.3 + 0.002j, 110, 220),
]


def test_tangent_tmatrix_high_orders(ctx):
    """Tangents at orders 45, 68 and 110 against central differences of the
    GPU forward path at relative step 1e-6 (measured: at most 1.4e-7 of the
    largest derivative element, 100 times less than at step 1e-5, so the
    difference is truncation, not a wrong tangent): bar 5e-7."""
    from costuum_amd.tmatrix import tangent_tmatrix, tangents
    h = 1e-6

    def shift(j, sign):
        out = []
        for c in HIGH:
            p = _params(c)
            p[j] += sign * h * max(abs(p[j]), 1e-3)
            out.append((p[0], p[1], c[2], complex(p[2], p[3]), c[4], c[5]))
        return out

    res = _fixed(ctx, HIGH)
    tangents(res)
    dT = {(g, j): tangent_tmatrix(res, g, j)
          for g in range(len(HIGH)) for j in range(4)}
    for j in range(4):
        print("orders 45/68/110: direction %d" % j, flush=True)
        rp = _fixed(ctx, shift(j, +1))
        Tp = [rp.tmatrix(g) for g in range(len(HIGH))]
        rm = _fixed(ctx, shift(j, -1))
        Tm = [rm.tmatrix(g) for g in range(len(HIGH))]
        for g in range(len(HIGH)):
            step = 2 * h * max(abs(_params(HIGH[g])[j]), 1e-3)
            fd = (Tp[g] - Tm[g]) / step
            err = np.abs(dT[(g, j)] - fd).max() / np.abs(fd).max()
            print("  order %d: %.2e" % (HIGH[g][4], err), flush=True)
            assert err < 5e-7, (j, g, err)


def test_tangent_tmatrix_matches_dual_oracle(ctx):
    """Tangent T-matrices against the dual-number oracle (oracle/dual.py:
    the oracle's chain carried as value and exact derivative, itself pinned
    to the C oracle's values and differences by tests/test_oracle_dual.py)
    at the CASES' forced orders, with no difference step in between:
    measured at most 9e-13 of the largest derivative element, bar 1e-11."""
    from oracle import dual
    from costuum_amd.tmatrix import tangent_tmatrix, tangents
    res = _fixed(ctx, CASES)
    tangents(res)
    for g, c in enumerate(CASES):
        for j in range(4):
            _, ref = dual.tmatrix_all_dual(*c, direction=j)
            dT = tangent_tmatrix(res, g, j)
            err = np.abs(dT - ref).max() / np.abs(ref).max()
            print("grain %d (order %d) direction %d: %.2e" % (g, c[4], j, err),
                  flush=True)
            assert err < 1e-11, (g, j, err)


def test_sum_tangents_match_oracle_differences(ctx):
    """d(Q_sca, Q_ext)/dp against central differences of the oracle's
    sums: pins the derivative to the reference algorithm."""
    from costuum_amd.tmatrix import tangents, vjp
    res = _fixed(ctx, CASES)
    dqs, dqe = tangents(res)
    for g, c in enumerate(CASES):
        for j in range(4):
            vals = []
            for sign in (+1, -1):
                rv, d, lam, mr, n, ng = _shift([c], j, sign)[0]
                vals.append(O.qcoeffs(n, O.tmatrix_all(rv, d, lam, mr, n,
                                                        ng)))
            h = _step(c, j)
            fs = (vals[0][0] - vals[1][0]) / h
            fe = (vals[0][1] - vals[1][1]) / h
            scale_s = max(abs(fs), 1e-6 * abs(vals[0][0]))
            scale_e = max(abs(fe), 1e-6 * abs(vals[0][1]))
            assert abs(dqs[g, j] - fs) < 1e-6 * scale_s + 1e-9, (g, j)
            assert abs(dqe[g, j] - fe) < 1e-6 * scale_e + 1e-9, (g, j)
    cot_s = np.array([1.0, -2.0, 0.5, 3.0])
    cot_e = np.array([0.3, 0.0, -1.0, 2.0])
    grad = vjp(res, cot_s, cot_e)
    assert np.allclose(grad, cot_s[:, None] * dqs + cot_e[:, None] * dqe,
                       rtol=0, atol=0)


def test_extinction_tangents_match_differences(ctx):
    from costuum_amd.tmatrix import observables, observables_tangent, \
        tangents
    theta = np.array([0.2, 0.9, 1.6, 2.5])
    n = len(CASES)
    fd = {}
    for j in range(4):
        out = []
        for sign in (+1, -1):
            _fixed(ctx, _shift(CASES, j, sign))
            q = observables(theta, ctx=ctx, account_for_scattering=False,
                            n_grains=n)
            out.append({k: v.cpu().numpy() for k, v in q.items()})
        steps = np.array([_step(c, j) for c in CASES])[:, None]
        fd[j] = {k: (out[0][k] - out[1][k]) / steps for k in out[0]}
    res = _fixed(ctx, CASES)
    tangents(res)
    for j in range(4):
        d = observables_tangent(theta, j, ctx, n,
                                radius=[c[0] for c in CASES])
        for k in ("qext", "qextpol", "qextcpol", "qabs", "qabspol"):
            got = d[k].cpu().numpy()
            ref = fd[j][k]
            scale = np.abs(fd[j]["qext"]).max()
            assert np.abs(got - ref).max() < 1e-5 * scale, (j, k)


def test_autograd_of_sums(ctx):
    import torch
    from costuum_amd.tmatrix import tmatrix_sums
    dev = ctx.device
    c = CASES[:2]
    r = torch.tensor([x[0] for x in c], dtype=torch.float64, device=dev,
                     requires_grad=True)
    d = torch.tensor([x[1] for x in c], dtype=torch.float64, device=dev,
                     requires_grad=True)
    mre = torch.tensor([x[3].real for x in c], dtype=torch.float64,
                       device=dev, requires_grad=True)
    mim = torch.tensor([x[3].imag for x in c], dtype=torch.float64,
                       device=dev, requires_grad=True)
    qs, qe = tmatrix_sums(r, d, [x[2] for x in c], mre, mim,
                          [x[4] for x in c], [x[5] for x in c], ctx=ctx)
    loss = (qe - qs).sum()
    loss.backward()
    for g, case in enumerate(c):
        rv, dd, lam, mr, n, ng = case
        for j, tensor in enumerate((r, d, mre, mim)):
            vals = []
            for sign in (+1, -1):
                rv2, d2, _, mr2, _, _ = _shift([case], j, sign)[0]
                s, e = O.qcoeffs(n, O.tmatrix_all(rv2, d2, lam, mr2, n, ng))
                vals.append(e - s)
            fd = (vals[0] - vals[1]) / _step(case, j)
            got = tensor.grad[g].item()
            assert abs(got - fd) < 1e-6 * max(abs(fd), 1e-3), (g, j, got, fd)


# seed-1 grains (the 1024-grain differentiation batch): 12 spread over the
# size parameter plus grain 401 (order 47); 785, 906 and 401 (orders 11, 22,
# 47) are checked against the oracle
SEED1 = [458, 825, 805, 765, 811, 79, 302, 679, 463, 785, 602, 590, 906, 401]
SEED1_CHECK = (785, 906, 401)
# tangent T-matrices of the checked grains against the dual-number oracle:
# measured 7e-15 (order 11), 5e-14 (order 22) and 3.6e-9 (order 47, whose
# Q blocks are the worst conditioned) of the largest derivative element
DUAL_BAR_SEED1 = 1e-7


def test_seed1_batch_gradients_match_oracle(ctx):
    """The differentiation workload on a stratified subset of its seed-1
    grains: task-path convergence, then d/d(R_V, d, Re m, Im m) of Q_ext and
    Q_abs (without the scattered light) at 16 Gauss-Legendre incidence
    angles at the frozen orders, against the oracle's central differences
    at relative steps h = 1e-5 and 2h, Richardson-extrapolated ((4 D(h) -
    D(2h)) / 3, truncation O(h^4)), to 1e-5 of the derivative's largest
    value.  The step balances the two grains that limit it, measured: at
    h = 1e-6 the difference's rounding error reaches 2.5e-5 on grain 401
    (order 47) along Im m, at h = 1e-4 the truncation error reaches 5e-3 on
    grain 906 (x = 13.8, Im m = 2e-4, resonant) along R_V; the extrapolated
    difference at 1e-5 is within 2e-6 of the best step on both."""
    import make_inputs
    from costuum_amd.tmatrix import (observables_tangent, tangents,
                                     tmatrix_batch)
    g = make_inputs.load(1, 1024)
    keys = ("radius", "axis_ratio", "wavelength", "mr_re", "mr_im")
    sub = {k: g[k][SEED1] for k in keys}
    theta = np.arccos(np.polynomial.legendre.leggauss(16)[0])
    print("seed-1 subset of %d grains" % len(SEED1), flush=True)
    r = tmatrix_batch(*[sub[k] for k in keys], ctx=ctx, min_order=4,
                      max_order=100, glfac=2, tolerance=1e-4)
    nmax, ngs, status = r.orders()
    tangents(r)
    # the tangent T-matrices of the checked grains against the dual-number
    # oracle (exact derivatives, no difference step)
    from oracle import dual
    from costuum_amd.tmatrix import tangent_tmatrix
    for gi in SEED1_CHECK:
        k = SEED1.index(gi)
        n, ng = int(nmax[k]), int(ngs[k])
        mr = complex(sub["mr_re"][k], sub["mr_im"][k])
        for j in range(4):
            _, ref = dual.tmatrix_all_dual(sub["radius"][k],
                                           sub["axis_ratio"][k],
                                           sub["wavelength"][k], mr, n, ng, j)
            dT = tangent_tmatrix(r, k, j)
            err = np.abs(dT - ref).max() / np.abs(ref).max()
            print("seed-1 grain %d (order %d) direction %d: dT vs dual "
                  "oracle %.2e" % (gi, n, j, err), flush=True)
            assert err < DUAL_BAR_SEED1, (gi, n, j, err)
    d = [observables_tangent(theta, j, ctx, len(SEED1),
                             radius=sub["radius"]) for j in range(4)]
    def central(p0, lam, n, ng, j, h):
        vals = []
        for sign in (+1, -1):
            p = p0.copy()
            p[j] += sign * h * abs(p[j])
            T = O.tmatrix_all(p[0], p[1], lam, complex(p[2], p[3]), n, ng)
            qe, _, _ = O.extinction(n, T, lam, p[0], theta)
            qa, _ = O.absorption(n, T, lam, p[0], theta,
                                 account_for_scattering=False)
            vals.append(np.stack([qe, qa]))
        return (vals[0] - vals[1]) / (2 * h * abs(p0[j]))

    h = 1e-5
    for gi in SEED1_CHECK:
        k = SEED1.index(gi)
        assert status[k] == 0
        n, ng = int(nmax[k]), int(ngs[k])
        p0 = np.array([sub["radius"][k], sub["axis_ratio"][k],
                       sub["mr_re"][k], sub["mr_im"][k]])
        lam = sub["wavelength"][k]
        for j in range(4):
            fd = (4. * central(p0, lam, n, ng, j, h) -
                  central(p0, lam, n, ng, j, 2 * h)) / 3.
            for c, key in enumerate(("qext", "qabs")):
                got = d[j][key][k].cpu().numpy()
                err = np.abs(got - fd[c]).max() / np.abs(fd[c]).max()
                assert err < 1e-5, (gi, n, j, key, err)


def _richardson(f, p0, j, h):
    """(4 D(h) - D(2h)) / 3 with D the central difference of f along
    parameter j at relative step h."""
    def central(hh):
        p1, p2 = p0.copy(), p0.copy()
        p1[j] += hh * abs(p0[j])
        p2[j] -= hh * abs(p0[j])
        return (f(p1) - f(p2)) / (2 * hh * abs(p0[j]))
    return (4. * central(h) - central(2 * h)) / 3.


def test_quadratic_observable_derivatives(ctx):
    """d/dp of Q_abs with the 12 x 12 scattered-light correction and of Z
    (ctm_observables_derivative) against Richardson-extrapolated central
    differences of the GPU forward observables (1e-6 of the largest value);
    the linear outputs equal the tangent-selected path; one grain's dQ_abs
    against the oracle's differences (1e-6); and every grain's dQ_abs and dZ
    along all four parameters against exact oracle derivatives built from
    the dual-number oracle's dT (measured at most 1.2e-10 of the largest
    value, at grain 3 along R_V and d, mostly below 1e-12; bar 1e-9)."""
    from costuum_amd.tmatrix import (observables, observables_derivative,
                                     observables_tangent, tangents)
    theta = np.array([0.3, 1.1, 2.4])
    n = len(CASES)
    kw = dict(ngauss=12, n_scatter_theta=4)

    def forward(p_all):
        cases = [(p[0], p[1], c[2], complex(p[2], p[3]), c[4], c[5])
                 for p, c in zip(p_all, CASES)]
        _fixed(ctx, cases)
        q = observables(theta, ctx=ctx, n_grains=n, **kw)
        return {k: v.cpu().numpy() for k, v in q.items()}

    p0 = np.array([_params(c) for c in CASES])
    res = _fixed(ctx, CASES)
    tangents(res)
    rv = [c[0] for c in CASES]
    got = [observables_derivative(theta, j, ctx, n, radius=rv, **kw)
           for j in range(4)]
    lin = [observables_tangent(theta, j, ctx, n, radius=rv) for j in range(4)]
    for j in range(4):
        for k in ("qext", "qextpol", "qextcpol"):
            a = got[j][k].cpu().numpy()
            b = lin[j][k].cpu().numpy()
            assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (j, k)
    for g in range(n):
        for j in range(4):
            def f(p, key, g=g):
                pa = p0.copy()
                pa[g] = p
                return forward(pa)[key][g]
            for key in ("qabs", "Z"):
                fd = _richardson(lambda p: f(p, key), p0[g].copy(), j, 1e-5)
                d = got[j][key][g].cpu().numpy()
                err = np.abs(d - fd).max() / np.abs(fd).max()
                assert err < 1e-6, (g, j, key, err)
    # one grain against the oracle: dQ_abs with the scattered light
    rv0, d0, lam, mr, nmax, ng = CASES[0]

    def oq(p):
        T = O.tmatrix_all(p[0], p[1], lam, complex(p[2], p[3]), nmax, ng)
        qa, _ = O.absorption(nmax, T, lam, p[0], theta, ngauss=12)
        return qa
    for j in range(4):
        fd = _richardson(oq, p0[0].copy(), j, 1e-5)
        d = got[j]["qabs"][0].cpu().numpy()
        assert np.abs(d - fd).max() < 1e-6 * np.abs(fd).max(), j
    # every grain against exact oracle derivatives: dT from the dual-number
    # oracle, then the observables' own exact rule for a function linear
    # plus quadratic in T, [f(T + s dT) - f(T - s dT)] / 2s (s makes the
    # step as large as T, so the difference loses no digits), and for Q_abs
    # the derivative of its 1 / (pi R_V^2) normalisation along R_V
    from oracle import dual
    for g, c in enumerate(CASES):
        rvg, dg, lamg, mrg, ng_, gg = c
        for j in range(4):
            T, dT = dual.tmatrix_all_dual(*c, direction=j)
            sc = np.abs(T).max() / np.abs(dT).max()
            qa_p, _ = O.absorption(ng_, T + sc * dT, lamg, rvg, theta,
                                   ngauss=12)
            qa_m, _ = O.absorption(ng_, T - sc * dT, lamg, rvg, theta,
                                   ngauss=12)
            ref = (qa_p - qa_m) / (2. * sc)
            if j == 0:
                qa0, _ = O.absorption(ng_, T, lamg, rvg, theta, ngauss=12)
                ref = ref - 2. * qa0 / rvg
            d = got[j]["qabs"][g].cpu().numpy()
            ea = np.abs(d - ref).max() / np.abs(ref).max()
            zp = O.scattering_grid(ng_, T + sc * dT, lamg, 4, use_gl=True)
            zm = O.scattering_grid(ng_, T - sc * dT, lamg, 4, use_gl=True)
            zref = (zp - zm) / (2. * sc)
            dz = got[j]["Z"][g].cpu().numpy()
            ez = np.abs(dz - zref).max() / np.abs(zref).max()
            print("grain %d direction %d: dQ_abs %.2e, dZ %.2e against the "
                  "dual oracle" % (g, j, ea, ez), flush=True)
            assert ea < DUAL_BAR_QUAD and ez < DUAL_BAR_QUAD, (g, j, ea, ez)


DUAL_BAR_QUAD = 1e-9


def test_autograd_of_observables(ctx):
    """torch autograd through Q_ext, Q_abs (scattered light included) and
    Q_sca at orientations against the oracle's Richardson differences of the
    same loss."""
    import torch
    from costuum_amd.tmatrix import observables_autograd
    dev = ctx.device
    c = CASES[:2]
    theta = np.array([0.4, 1.3])
    leaves = [torch.tensor([_params(x)[j] for x in c], dtype=torch.float64,
                           device=dev, requires_grad=True) for j in range(4)]
    qe, qa, qs = observables_autograd(leaves[0], leaves[1],
                                      [x[2] for x in c], leaves[2],
                                      leaves[3], [x[4] for x in c],
                                      [x[5] for x in c], theta,
                                      ngauss_abs=10, ctx=ctx)
    w = torch.tensor([1.0, -0.5], dtype=torch.float64, device=dev)
    loss = ((qe + 2.0 * qa - 0.5 * qs) @ w).sum()
    loss.backward()
    for g, case in enumerate(c):
        rv, dd, lam, mr, n, ng = case

        def fl(p):
            T = O.tmatrix_all(p[0], p[1], lam, complex(p[2], p[3]), n, ng)
            e, _, _ = O.extinction(n, T, lam, p[0], theta)
            a, _ = O.absorption(n, T, lam, p[0], theta, ngauss=10)
            return float(np.dot(e + 2.0 * a - 0.5 * (e - a),
                                [1.0, -0.5]))
        for j in range(4):
            fd = _richardson(fl, _params(case), j, 1e-5)
            got = leaves[j].grad[g].item()
            assert abs(got - fd) < 1e-6 * max(abs(fd), 1e-3), (g, j, got, fd)
