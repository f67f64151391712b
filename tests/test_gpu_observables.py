"""GPU observables (extinction, absorption with the scattered-light
quadrature, scattering-matrix grid) against the CPU oracle on the same
T-matrices (forced orders), FP64: relative error <= 1e-9."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 10, 20),
    (0.2, 2.0, 0.55, 1.5 + 0.02j, 10, 20),
    (0.3, 1.0, 0.63, 1.7 + 0.1j, 12, 24),
    (0.05, 0.7, 2.0, 1.9 + 0.3j, 5, 10),
    (0.8, 1.3, 1.0, 1.33 + 0.001j, 16, 32),
]


@pytest.fixture(scope="module")
def batch():
    from costuum_amd.tmatrix import Context, tmatrix_fixed
    ctx = Context(0)
    c = CASES
    r = tmatrix_fixed([x[0] for x in c], [x[1] for x in c],
                      [x[2] for x in c], [x[3].real for x in c],
                      [x[3].imag for x in c], [x[4] for x in c],
                      [x[5] for x in c], ctx=ctx)
    yield ctx, r
    ctx.close()


def test_extinction_absorption_match_oracle(batch):
    from costuum_amd.tmatrix import observables
    ctx, r = batch
    x, _ = np.polynomial.legendre.leggauss(8)
    theta = np.concatenate([np.arccos(x), [0.0, np.pi / 6, np.pi]])
    print("observables: %d grains x %d angles" % (len(CASES), len(theta)),
          flush=True)
    out = observables(theta, ctx=ctx, ngauss=20, n_grains=len(CASES))
    got = {k: v.cpu().numpy() for k, v in out.items()}
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        qe, qep, qec = O.extinction(nmax, T, lam, rv, theta)
        qa, qap = O.absorption(nmax, T, lam, rv, theta, ngauss=20)
        scale = np.abs(qe).max()
        for name, ref in (("qext", qe), ("qextpol", qep), ("qextcpol", qec),
                          ("qabs", qa), ("qabspol", qap)):
            err = np.abs(got[name][g] - ref).max() / scale
            assert err < 1e-9, (g, name, err)


def test_scattering_grid_matches_oracle(batch):
    from costuum_amd.tmatrix import observables
    ctx, r = batch
    for use_gl in (True, False):
        out = observables([0.3], ctx=ctx, do_absorption=False,
                          account_for_scattering=False, n_scatter_theta=6,
                          scatter_gl=use_gl, n_grains=len(CASES))
        Z = out["Z"].cpu().numpy()
        for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
            T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
            Zr = O.scattering_grid(nmax, T, lam, 6, use_gl=use_gl)
            err = np.abs(Z[g] - Zr).max() / np.abs(Zr).max()
            assert err < 1e-9, (g, use_gl, err)


def test_scattering_grid_many_directions_high_order():
    """120 scattering angles at order 70: the outgoing pi/tau of the blocks
    above order 32 no longer fit the workgroup's LDS and are read from the
    table.  Z at 48 grid points against oracle queries in the same
    directions on the GPU's own T-matrix (observables path only; the
    oracle's grid and its queries agree to 4e-16), <= 1e-9."""
    from costuum_amd.tmatrix import Context, observables, tmatrix_fixed
    ctx = Context(0)
    try:
        rv, d, lam, mr, nmax, ng = 1.8, 1.2, 0.25, 1.4 + 0.01j, 70, 140
        nst = 120
        r = tmatrix_fixed([rv], [d], [lam], [mr.real], [mr.imag], [nmax],
                          [ng], ctx=ctx)
        print("Z grid: order %d, %d angles" % (nmax, nst), flush=True)
        out = observables([0.3], ctx=ctx, do_absorption=False,
                          account_for_scattering=False, n_scatter_theta=nst,
                          scatter_gl=True, n_grains=1)
        Z = out["Z"].cpu().numpy()[0]
        T = r.tmatrix(0)
        xt = O.gauss_legendre(nst)[0]
        xp = O.gauss_legendre(2 * nst)[0]
        rng = np.random.default_rng(3)
        pts = [(int(i), int(j)) for i, j in
               zip(rng.integers(0, nst, 48), rng.integers(0, 2 * nst, 48))]
        scale = np.abs(Z).max()
        for i, j in pts:
            _, Zq, _ = O.scattering_query(nmax, T, lam, 0., 0., np.pi / 2, 0.,
                                          np.arccos(xt[i]),
                                          np.pi * xp[j] + np.pi)
            err = np.abs(Z[i, j] - Zq).max() / scale
            assert err < 1e-9, (i, j, err)
    finally:
        ctx.close()
