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
