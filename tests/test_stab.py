"""SKIRT stored-table writer: byte layout of the reference's writer
(test/testTaskManager.cpp:77-119,236-298) and a round trip."""
import struct

import numpy as np
import pytest

from costuum_amd.stab import read_stab, write_absorption_table, write_stab


def test_layout_matches_reference_writer(tmp_path):
    sizes = [1e-7, 1e-6]
    lams = [1e-5, 1e-4, 1e-3]
    thetas = [0.5, 1.5]
    table = np.arange(2 * 3 * 2 * 2, dtype=np.float64).reshape(2, 3, 2, 2)
    p = tmp_path / "t.stab"
    write_absorption_table(p, sizes, lams, thetas, table)
    b = p.read_bytes()
    # the bytes the reference's helpers produce, built independently
    u64 = lambda x: struct.pack("=Q", x)  # noqa: E731
    f64 = lambda x: struct.pack("=d", x)  # noqa: E731
    exp = b"SKIRT X\n" + u64(0x010203040A0BFEFF) + u64(3)
    for s in ("a", "lambda", "theta", "m", "m", "rad", "log", "log", "lin"):
        exp += s.encode().ljust(8, b" ")
    for arr in (sizes, lams, thetas):
        exp += u64(len(arr)) + b"".join(f64(v) for v in arr)
    exp += u64(2)
    for s in ("Qabs", "Qabspol", "1", "1", "log", "lin"):
        exp += s.encode().ljust(8, b" ")
    for it in range(2):
        for il in range(3):
            for isz in range(2):
                exp += f64(table[isz, il, it, 0]) + f64(table[isz, il, it, 1])
    exp += b"STABEND\n"
    assert b == exp


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    axes = [("x", "m", "lin", np.linspace(0, 1, 4)),
            ("y", "1", "log", np.logspace(-2, 1, 3))]
    q = [("A", "1", "lin"), ("B", "W", "log"), ("C", "1", "lin")]
    vals = rng.normal(size=(3, 4, 3))
    p = tmp_path / "r.stab"
    write_stab(p, axes, q, vals)
    a2, q2, v2 = read_stab(p)
    assert [x[:3] for x in a2] == [x[:3] for x in axes]
    assert all(np.array_equal(x[3], y[3]) for x, y in zip(a2, axes))
    assert q2 == q
    assert np.array_equal(v2, vals)


def test_rejects_bad_shapes_and_long_names(tmp_path):
    with pytest.raises(ValueError):
        write_stab(tmp_path / "x.stab", [("x", "m", "lin", [1, 2])],
                   [("A", "1", "lin")], np.zeros((3, 1)))
    with pytest.raises(ValueError):
        write_stab(tmp_path / "x.stab", [("toolongname", "m", "lin", [1])],
                   [("A", "1", "lin")], np.zeros((1, 1)))
