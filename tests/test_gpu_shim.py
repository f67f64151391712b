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
        # forward scattering matrices (TMatrix.hpp:1109,1246,1306)
        S1, _, _ = O.scattering_query(n, T, lam, 0.4, 0.9, 1.1, 2.0, 2.5, 0.3)
        S2, _, _ = O.scattering_query(n, T, lam, 0., 0., 1.1, 2.0, 2.5, 0.3)
        s = np.array(out["s"]).reshape(3, 4, 2)
        for k, Sref in enumerate((S1, S2, S2)):
            got = s[k, :, 0] + 1j * s[k, :, 1]
            assert np.abs(got - Sref.ravel()).max() <= \
                1e-9 * np.abs(Sref).max(), k
        # reduced extinction matrix [K11, K12, K34] (TMatrix.hpp:2034-2048)
        assert _close(out["reduced_k"], [K[0, 0], K[0, 1], K[2, 3]])
        # absorption cross sections: C = Q pi R_V^2
        area = np.pi * rv * rv
        qa, qap = O.absorption(n, T, lam, rv, [0.7, 1.9], ngauss=6)
        assert _close(out["cabs_quad"], [qa[0] * area, qap[0] * area])
        assert _close(out["cabs_pairs"], [qa[0] * area, qap[0] * area,
                                          qa[1] * area, qap[1] * area])
        assert _close(out["cabs"], qa * area)
        x, w = np.polynomial.legendre.leggauss(6)
        qg, _ = O.absorption(n, T, lam, rv, np.arccos(x), ngauss=6)
        assert _close(out["cabs_avg"], 0.5 * np.sum(w * qg) * area)
        # set_element doubles T^(11)_{110}; the queries follow
        Tm = T.copy()
        Tm[0] *= 2.
        assert abs(complex(*out["set_t"]) - Tm[0]) <= 1e-9 * abs(Tm[0])
        _, Zs, _ = O.scattering_query(n, Tm, lam, 0.4, 0.9, 1.1, 2.0, 2.5,
                                      0.3)
        assert _close(out["set_z"], Zs.ravel())
        assert not _close(out["set_z"], Z2.ravel(), 1e-6)


TASKS_BIN = os.path.join(ROOT, "tests", "cpp", "tasks_check")


def test_task_manager_matches_oracle():
    """TaskManager / generate_tasks / QuickSched / results through
    ResultKey (include/costuum_amd_tasks.hpp, tests/cpp/tasks_check.cpp)
    against the oracle restatement of the table: two compositions, unsorted
    sizes and wavelengths, three Draine-Hensley shapes, Mishchenko
    alignment above 0.1, absorption with the 12 x 12 scattered light,
    extinction and the scattering grid."""
    from costuum_amd import CosTuuM as C
    from test_gpu_cosTuuM import _oracle_table
    assert os.path.exists(TASKS_BIN), "tests/cpp/tasks_check is not built"
    print("tasks_check: 2 x 2 x 2 x 3 shapes", flush=True)
    res = subprocess.run([TASKS_BIN], capture_output=True, text=True,
                         timeout=100)
    assert res.returncode == 0, res.stderr
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["sizes"] == [0.08, 0.25]
    assert out["wavelengths"] == [0.55, 1.1]
    assert out["n_results"] == 2 * 2 * 2 * 3
    shape = C.DraineHensleyShapeDistribution(3)
    ref_shapes = [[shape.get_shape(i), shape.get_weight(i)] for i in range(3)]
    assert np.allclose(out["shapes"], ref_shapes, rtol=1e-12, atol=0)
    types = [C.SILICON, C.CARBON_PARALLEL]
    align = C.SizeBasedAlignmentDistribution(0.1, C.MISHCHENKO_ALIGNMENT, 50)
    dust = C.CustomDustProperties(
        lambda lam, a, t: complex(1.4 + 0.1 * t + 0.05 * lam, 0.02 + 0.01 * t))
    thetas = np.array([0.3, 1.2, 2.0])
    rq, rz = _oracle_table(types, [0.25, 0.08], [1.1, 0.55], thetas, shape,
                           align, dust, 4, 60, True, True, True, True, 12)
    q = np.array(out["q"]).reshape(2, 2, 2, 3, 5)
    rq = np.array(rq).reshape(q.shape)
    z = np.array(out["z"]).reshape(2, 2, 2, 18, 16)
    rz = np.array(rz).reshape(z.shape)
    for k in range(5):
        e = np.abs(q[..., k] - rq[..., k]).max() / np.abs(rq[..., k]).max()
        assert e < 1e-8, (k, e)
    assert np.abs(z - rz).max() / np.abs(rz).max() < 1e-8
