"""SKIRT stored-table (.stab) files for get_table results.

The binary layout is the one the reference writes for SKIRT
(test/testTaskManager.cpp:236-298, helpers :77-119), all in the host's byte
order:
  * the tag ``"SKIRT X\\n"`` and the 64-bit integer 0x010203040A0BFEFF
    (lets a reader detect the byte order);
  * the number of axes, then the axis names, units and scales ("lin" or
    "log"), each an 8-character field padded with spaces;
  * per axis its number of points and the points as 64-bit floats;
  * the number of quantities, their names, units and scales (8 characters
    each);
  * the values as 64-bit floats, quantities fastest, then the first axis,
    the second, ... (the reference's loop order: theta outside lambda
    outside size);
  * the trailer ``"STABEND\\n"``.
"""
import numpy as np

_TAG = b"SKIRT X\n"
_END = b"STABEND\n"
_ORDER_MARK = 0x010203040A0BFEFF


def _field(s):
    b = s.encode("ascii")
    if len(b) > 8:
        raise ValueError("stored-table names are at most 8 characters: %r"
                         % s)
    return b + b" " * (8 - len(b))


def write_stab(path, axes, quantities, values):
    """Writes a stored table.

    axes: list of (name, unit, scale, points) from the fastest-varying axis
    to the slowest; quantities: list of (name, unit, scale); values: array
    of shape [n_last_axis, ..., n_first_axis, n_quantities] (C order, the
    first axis varying fastest after the quantity index)."""
    shape = tuple(len(a[3]) for a in reversed(axes)) + (len(quantities),)
    v = np.ascontiguousarray(values, dtype="=f8")
    if v.shape != shape:
        raise ValueError("values have shape %s, axes and quantities need %s"
                         % (v.shape, shape))
    u64 = lambda x: np.array([x], dtype="=u8").tobytes()  # noqa: E731
    with open(path, "wb") as f:
        f.write(_TAG)
        f.write(u64(_ORDER_MARK))
        f.write(u64(len(axes)))
        for k in range(3):
            for a in axes:
                f.write(_field(a[k]))
        for a in axes:
            pts = np.ascontiguousarray(a[3], dtype="=f8")
            f.write(u64(len(pts)))
            f.write(pts.tobytes())
        f.write(u64(len(quantities)))
        for k in range(3):
            for q in quantities:
                f.write(_field(q[k]))
        f.write(v.tobytes())
        f.write(_END)


def read_stab(path):
    """Reads a stored table written in this byte order: returns (axes,
    quantities, values) in the form write_stab takes."""
    data = open(path, "rb").read()
    if data[:8] != _TAG or data[-8:] != _END:
        raise ValueError("not a SKIRT stored table")
    pos = 8

    def u64():
        nonlocal pos
        x = int(np.frombuffer(data, dtype="=u8", count=1, offset=pos)[0])
        pos += 8
        return x

    def names(n):
        nonlocal pos
        out = [data[pos + 8 * i:pos + 8 * i + 8].decode("ascii").rstrip(" ")
               for i in range(n)]
        pos += 8 * n
        return out

    if u64() != _ORDER_MARK:
        raise ValueError("stored table in the other byte order")
    na = u64()
    an, au, asc = names(na), names(na), names(na)
    pts = []
    for _ in range(na):
        n = u64()
        pts.append(np.frombuffer(data, dtype="=f8", count=n, offset=pos).copy())
        pos += 8 * n
    nq = u64()
    qn, qu, qs = names(nq), names(nq), names(nq)
    shape = tuple(len(p) for p in reversed(pts)) + (nq,)
    count = int(np.prod(shape))
    vals = np.frombuffer(data, dtype="=f8", count=count,
                         offset=pos).reshape(shape).copy()
    axes = [(an[i], au[i], asc[i], pts[i]) for i in range(na)]
    quants = [(qn[i], qu[i], qs[i]) for i in range(nq)]
    return axes, quants, vals


def write_absorption_table(path, sizes, wavelengths, thetas, table):
    """The reference's SKIRT hand-off of one composition's absorption
    coefficients (test/testTaskManager.cpp:236-298): axes a [m, log],
    lambda [m, log], theta [rad, lin]; quantities Qabs [1, log] and
    Qabspol [1, lin].  table is get_table's [nsize, nlambda, ntheta, 2]
    output (do_absorption only) for sorted sizes and wavelengths in metres."""
    t = np.asarray(table, dtype=np.float64)
    t = t.reshape(len(sizes), len(wavelengths), len(thetas), 2)
    values = np.transpose(t, (2, 1, 0, 3))  # theta, lambda, size, quantity
    write_stab(path,
               [("a", "m", "log", np.asarray(sizes, dtype=np.float64)),
                ("lambda", "m", "log", np.asarray(wavelengths,
                                                  dtype=np.float64)),
                ("theta", "rad", "lin", np.asarray(thetas,
                                                   dtype=np.float64))],
               [("Qabs", "1", "log"), ("Qabspol", "1", "lin")], values)
