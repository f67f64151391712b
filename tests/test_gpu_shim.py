"""The C++ TMatrixCalculator / TMatrix surface (include/costuum_amd.hpp,
exercised by tests/cpp/shim_check) against the oracle: the calculator's
two-phase converged orders, the T-matrix sums and elements, Z and K at
arbitrary orientations, and the isotropic ensemble."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "shim_check")

# the single-grain workload (prolate and oblate twin) and a larger grain
GRAINS = [(0.2, 0.5, 0.55, 1.5 + 0.02j), (0.2, 2.0, 0.55, 1.5 + 0.02j),
          (0.5, 1.3, 1.2, 1.33 + 0.001j)]


def _close(got, ref, tol=1e-9):
    got = np.asarray(got, dtype=float)
    ref = np.asarray(ref, dtype=float)
    return np.abs(got - ref).max() <= tol * max(np.abs(ref).max(), 1e-300)


def test_shim_matches_oracle():
    assert os.path.exists(BIN), "tests/cpp/shim_check is not built " \
                                "(python __graft_entry__.py)"
    inp = "".join("%r %r %r %r %r\n" % (r, d, lam, m.real, m.imag)
                  for r, d, lam, m in GRAINS)
    print("shim_check on %d grains" % len(GRAINS), flush=True)
    res = subprocess.run([BIN], input=inp, capture_output=True, text=True,
                         timeout=100)
    assert res.returncode == 0, res.stderr
    outs = [json.loads(x) for x in res.stdout.splitlines()
            if x.startswith("{")]
    assert len(outs) == len(GRAINS)
    for (r, d, lam, mr), out in zip(GRAINS, outs):
        st, n, g, rv = O.converge_calculator(1.0, d, r, lam, mr, 200, 1e-4,
                                             2, 500)
        assert st == 0
        assert (out["nmax"], out["ngauss"]) == (n, g), (out["nmax"], n)
        T = O.tmatrix_all(rv, d, lam, mr, n, g)
        qs, qe = O.qcoeffs(n, T)
        assert _close(out["qsca"], qs) and _close(out["qext"], qe)
        B = O.tmatrix_blocks(T, n)

        def el(i1, n1, m, i2, n2):
            nm = n if m == 0 else n + 1 - m
            p = max(m, 1)
            return B[m][i1 * nm + n1 - p, i2 * nm + n2 - p]

        ref = [el(0, 1, 0, 0, 1), el(1, 2, 1, 1, 2), el(0, n, n, 1, n)]
        t = out["t"]
        scale = np.abs(T).max()
        for k, v in enumerate(ref):
            assert abs(complex(t[2 * k], t[2 * k + 1]) - v) <= 1e-9 * scale
        _, Z1, _ = O.scattering_query(n, T, lam, 0., 0., 0., 0., 1.2, 0.)
        _, Z2, _ = O.scattering_query(n, T, lam, 0.4, 0.9, 1.1, 2.0, 2.5, 0.3)
        _, _, K = O.scattering_query(n, T, lam, 0.3, 0.7, 0.9, 0.2, 0.9, 0.2)
        assert _close(out["z1"], Z1.ravel()) and _close(out["z2"], Z2.ravel())
        assert _close(out["k"], K.ravel())
        p = np.zeros(2 * n + 1)
        p[0] = 1.0
        Ta = O.orientation_average(n, T, p)
        qsa, qea = O.qcoeffs(n, Ta)
        assert _close(out["avg_qsca"], qsa) and _close(out["avg_qext"], qea)
        _, Za, _ = O.scattering_query(n, Ta, lam, 0., 0., 0.5, 0., 1.7, 0.)
        assert _close(out["avg_z"], Za.ravel())
