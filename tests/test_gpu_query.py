"""GPU arbitrary-orientation queries (amplitude, phase and extinction
matrices at a particle orientation (alpha, beta) for laboratory directions)
and the m = 0 convergence sums, against the CPU oracle on the same forced
orders, FP64."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    (0.2, 0.5, 0.55, 1.5 + 0.02j, 10, 20),
    (0.3, 1.0, 0.63, 1.7 + 0.1j, 12, 24),
    (0.05, 0.7, 2.0, 1.9 + 0.3j, 5, 10),
    (0.8, 1.3, 1.0, 1.33 + 0.001j, 16, 32),
]


def _fixed(ctx):
    from costuum_amd.tmatrix import tmatrix_fixed
    c = CASES
    return tmatrix_fixed([x[0] for x in c], [x[1] for x in c],
                         [x[2] for x in c], [x[3].real for x in c],
                         [x[3].imag for x in c], [x[4] for x in c],
                         [x[5] for x in c], ctx=ctx)


@pytest.fixture(scope="module")
def ctx():
    from costuum_amd.tmatrix import Context
    c = Context(0)
    yield c
    c.close()


def _queries(seed, n):
    rng = np.random.default_rng(seed)
    q = np.stack([rng.uniform(0, 2 * np.pi, n), rng.uniform(0, np.pi, n),
                  rng.uniform(0.05, np.pi - 0.05, n),
                  rng.uniform(0, 2 * np.pi, n),
                  rng.uniform(0.05, np.pi - 0.05, n),
                  rng.uniform(0, 2 * np.pi, n)], axis=1)
    # special directions: forward scattering, the poles of the particle
    # frame, backscattering
    q[0] = [0.0, 0.0, 0.7, 0.3, 0.7, 0.3]
    q[1] = [0.4, 0.9, 1.1, 2.0, np.pi - 1.1, 2.0 + np.pi]
    q[2] = [1.0, 0.0, 0.0, 0.0, np.pi / 2, 1.0]
    # outgoing along the symmetry axis; incidence along it with
    # backscattering along it (only m = 1 contributes to S there)
    q[3] = [0.3, 0.0, 0.9, 0.4, 0.0, 0.0]
    q[4] = [0.0, 0.0, 0.0, 0.0, np.pi, 0.0]
    return q


def _check(ctx, Ts, tag):
    from costuum_amd.tmatrix import query
    q = _queries(7, 12)
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        print("%s query grain %d (nmax %d)" % (tag, g, nmax), flush=True)
        S, Z, K = query(ctx, g, *q.T)
        for i in range(len(q)):
            Sr, Zr, Kr = O.scattering_query(nmax, Ts[g], lam, *q[i])
            sc = max(np.abs(Sr).max(), 1e-300)
            assert np.abs(S[i] - Sr).max() / sc < 1e-9, (tag, g, i, "S")
            zc = max(np.abs(Zr).max(), 1e-300)
            assert np.abs(Z[i] - Zr).max() / zc < 1e-9, (tag, g, i, "Z")
            kc = max(np.abs(Kr).max(), 1e-300)
            assert np.abs(K[i] - Kr).max() / kc < 1e-9, (tag, g, i, "K")


def test_query_matches_oracle(ctx):
    _fixed(ctx)
    Ts = [O.tmatrix_all(rv, d, lam, mr, nmax, ng)
          for (rv, d, lam, mr, nmax, ng) in CASES]
    _check(ctx, Ts, "plain")


def test_query_on_averaged_tmatrices(ctx):
    r = _fixed(ctx)
    nmax_all = max(c[4] for c in CASES)
    ncoef = 2 * nmax_all + 1
    pN = np.stack([O.orientation_coefficients(O.ORIENT_MISHCHENKO, 0.6,
                                              ncoef - 1),
                   O.orientation_coefficients(O.ORIENT_ISOTROPIC, 0.0,
                                              ncoef - 1)])
    dist = np.array([0, 1, 0, 1], dtype=np.int32)
    r.average(pN, dist)
    Ts = []
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T = O.tmatrix_all(rv, d, lam, mr, nmax, ng)
        Ts.append(O.orientation_average(nmax, T, pN[dist[g]]))
    _check(ctx, Ts, "averaged")


def test_query_forward_extinction_consistent(ctx):
    """Optical theorem inside the query: K[0,0] = (2 pi / k) Im(S11 + S22)
    in the forward direction."""
    from costuum_amd.tmatrix import query
    _fixed(ctx)
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        k = 2 * np.pi / lam
        S, Z, K = query(ctx, g, 0.3, 0.8, 1.0, 0.5, 1.0, 0.5)
        assert abs(K[0, 0] - 2 * np.pi / k * (S[0, 0] + S[1, 1]).imag) \
            <= 1e-12 * abs(K[0, 0])


def test_m0_sums_match_oracle(ctx):
    from costuum_amd.tmatrix import m0_sums
    c = CASES
    qs, qe = m0_sums([x[0] for x in c], [x[1] for x in c],
                     [x[2] for x in c], [x[3].real for x in c],
                     [x[3].imag for x in c], [x[4] for x in c],
                     [x[5] for x in c], ctx=ctx)
    for g, (rv, d, lam, mr, nmax, ng) in enumerate(CASES):
        T0 = O.tmatrix_m(rv, d, lam, mr, nmax, ng, 0)
        s, e = O.m0_qsums(nmax, T0)
        assert abs(qs[g] - s) <= 1e-10 * abs(s), (g, qs[g], s)
        assert abs(qe[g] - e) <= 1e-10 * abs(e), (g, qe[g], e)
