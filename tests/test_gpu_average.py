"""GPU orientation averaging (ensemble T-matrix) against the oracle's
restatement of AlignmentAverageTask, then observables of the ensemble."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 10, 20),
    (0.15, 1.6, 0.7, 1.33 + 0.01j, 7, 14),
    (0.05, 0.8, 1.0, 1.7 + 0.2j, 5, 10),
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


@pytest.mark.parametrize("kind,param", [(0, 0.), (1, 0.3), (2, 0.5),
                                        (2, 2.0)])
def test_average_matches_oracle(batch, kind, param):
    ctx, r = batch
    nmax_all = max(c[4] for c in CASES)
    p = O.orientation_coefficients(kind, param, 2 * nmax_all)
    print("average kind", kind, "param", param, flush=True)
    r.average(p)
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        Te = O.orientation_average(nmax, T, p)
        got = r.averaged_tmatrix(g)
        err = np.abs(got - Te).max() / np.abs(Te).max()
        assert err < 1e-10, (g, err)


def test_observables_of_ensemble(batch):
    from costuum_amd.tmatrix import observables
    ctx, r = batch
    nmax_all = max(c[4] for c in CASES)
    p = O.orientation_coefficients(1, 0.5, 2 * nmax_all)
    r.average(p)
    theta = np.array([0.2, np.pi / 6, 1.3, 2.7])
    out = observables(theta, ctx=ctx, ngauss=12, n_grains=len(CASES))
    qe = out["qext"].cpu().numpy()
    qa = out["qabs"].cpu().numpy()
    qep = out["qextpol"].cpu().numpy()
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        Te = O.orientation_average(nmax, T, p)
        e, ep, _ = O.extinction(nmax, Te, lam, rv, theta)
        a, _ = O.absorption(nmax, Te, lam, rv, theta, ngauss=12)
        s = np.abs(e).max()
        assert np.abs(qe[g] - e).max() < 1e-9 * s
        assert np.abs(qep[g] - ep).max() < 1e-9 * s
        assert np.abs(qa[g] - a).max() < 1e-9 * s


def test_weighted_sums_shape_and_size():
    """Shape averages (normalised) and size integrals (plain weighted sums)
    on the GPU against numpy, summed in the same order."""
    import torch
    from costuum_amd.tmatrix import Context, weighted_sums
    ctx = Context(0)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(6 * 7, 3, 5))
    w = rng.uniform(0.1, 2.0, size=6 * 7)
    got = weighted_sums(torch.as_tensor(x, device=ctx.device), w, 7,
                        normalise=True, ctx=ctx).cpu().numpy()
    xs = x.reshape(6, 7, 3, 5)
    ws = w.reshape(6, 7)
    ref = np.einsum("gs,gsij->gij", ws, xs) / ws.sum(axis=1)[:, None, None]
    assert np.abs(got - ref).max() < 1e-14 * np.abs(ref).max()
    got2 = weighted_sums(torch.as_tensor(x, device=ctx.device), w, 42,
                         normalise=False, ctx=ctx).cpu().numpy()
    ref2 = np.einsum("s,sij->ij", w, x)[None]
    assert np.abs(got2 - ref2).max() < 1e-13 * np.abs(ref2).max()
    ctx.close()


def test_sparse_coefficients_after_dense(batch):
    # Davis-Greenstein coefficients reach every (n1, n2) pair; the isotropic
    # ones only n1 = n2: the second average must leave no entry of the first
    ctx, r = batch
    nmax_all = max(c[4] for c in CASES)
    r.average(O.orientation_coefficients(2, 2.0, 2 * nmax_all))
    p = O.orientation_coefficients(0, 0., 2 * nmax_all)
    print("average DG then isotropic", flush=True)
    r.average(p)
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        Te = O.orientation_average(nmax, T, p)
        got = r.averaged_tmatrix(g)
        assert np.abs(got - Te).max() / np.abs(Te).max() < 1e-10, g
