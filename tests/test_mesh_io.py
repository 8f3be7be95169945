"""Mesh I/O on the host (no GPU): the binary mesh image round-trips every
Mesh::load array, rejects damaged files, the cached sphere equals the
generated one, and print_function's text follows src/phelm.cpp:248-309
(restated below from the reference's fprintf formats)."""
import math
import os

import numpy as np
import pytest

import mke_amd as P

KEYS = ("pr", "coords", "tri", "tri_zone", "adj_ptr", "adj", "is_boundary", "inner", "outer",
        "p2io")


def _same(a, b):
    return all(np.array_equal(np.asarray(a[k]), np.asarray(b[k])) for k in KEYS)


@pytest.mark.parametrize("level", [0, 2, 4])
def test_binary_image_round_trip(level, tmp_path):
    m = P.Mesh.sphere(level)
    path = str(tmp_path / "m.phgmesh")
    m.save(path)
    r = P.Mesh.load_file(path)
    assert (r.n_points, r.n_triangles, r.inner_size, r.outer_size, r.n_zones) == \
        (m.n_points, m.n_triangles, m.inner_size, m.outer_size, m.n_zones)
    assert _same(m.arrays(), r.arrays())
    rp0, ci0 = m.csr_structure()
    rp1, ci1 = r.csr_structure()
    assert np.array_equal(rp0, rp1) and np.array_equal(ci0, ci1)


def test_damaged_or_foreign_images_are_refused(tmp_path):
    m = P.Mesh.sphere(2)
    path = str(tmp_path / "m.phgmesh")
    m.save(path)
    data = bytearray(open(path, "rb").read())
    for name, blob in (("flip", data[:100] + bytes([data[100] ^ 1]) + data[101:]),
                       ("short", data[:-8]), ("text", P.sphere_mesh_text(1).encode()),
                       ("empty", b"")):
        bad = str(tmp_path / (name + ".phgmesh"))
        open(bad, "wb").write(bytes(blob))
        with pytest.raises(P.PhelmError):
            P.Mesh.load_file(bad)
    with pytest.raises(P.PhelmError):
        P.Mesh.load_file(str(tmp_path / "missing.phgmesh"))


def test_cached_sphere(tmp_path):
    d = str(tmp_path)
    a = P.Mesh.sphere_cached(3, d)
    assert os.path.exists(os.path.join(d, "sphere_L3.phgmesh"))
    b = P.Mesh.sphere_cached(3, d)
    assert _same(a.arrays(), P.Mesh.sphere(3).arrays()) and _same(a.arrays(), b.arrays())


def _print_function_ref(a, u, sphere_xyz):
    """print_function_ (src/phelm.cpp:248-296) with x/y/z null (u, v, 0) or
    the drivers' sphere maps (examples/test_baroclin.cpp:36-49)."""
    uv = a["coords"][:, 0, :]
    out = ["# points %d\n" % len(u)]
    for i in range(len(u)):
        s, t = uv[i]
        if sphere_xyz:
            x, y, z = math.cos(s) * math.cos(t), math.cos(s) * math.sin(t), math.sin(s)
        else:
            x, y, z = s, t, 0.0
        out.append("%.16f %.16f %.16f %.16f %.16f %.16f \n" % (x, y, z, u[i], s, t))
    out.append("# triangles %d\n" % len(a["tri"]))
    for tr in a["tri"]:
        out.append("%d %d %d\n" % (tr[0] + 1, tr[1] + 1, tr[2] + 1))
    out.append("# end \n")
    return "".join(out)


@pytest.mark.parametrize("sphere_xyz", [False, True])
def test_print_function_format(sphere_xyz):
    m = P.Mesh.sphere(1)
    a = m.arrays()
    u = np.sin(3.0 * a["pr"][:, 0]) + a["pr"][:, 2] ** 2
    got = m.print_function(u, sphere_xyz)
    want = _print_function_ref(a, u, sphere_xyz)
    if sphere_xyz:   # libm's cos/sin may differ from Python's in the last digit
        g, w = got.splitlines(), want.splitlines()
        assert len(g) == len(w)
        for x, y in zip(g, w):
            if x.startswith("#") or len(x.split()) == 3:
                assert x == y
            else:
                assert np.allclose([float(t) for t in x.split()], [float(t) for t in y.split()],
                                   rtol=0, atol=2e-16)
    else:
        assert got == want


def test_matrix_market_writer(tmp_path):
    """dump_operator's writer (utils/dump_operator_matrix.cpp counterpart):
    the Matrix Market text reads back as the same matrix, FP64 exactly with
    --double and rounded to float without; ELL order lists slot k of every
    row before slot k + 1."""
    sio = pytest.importorskip("scipy.io")
    from mke_amd.dump_operator import write_matrix_market
    rp = np.array([0, 2, 5, 6], np.int32)
    ci = np.array([0, 2, 0, 1, 2, 1], np.int32)
    v = np.array([1.0 / 3.0, -2.5e-17, math.pi, 4.0, -1e30, 0.1])
    dense = np.zeros((3, 3))
    for i in range(3):
        dense[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    for double, store in ((True, "CSR"), (False, "CSR"), (True, "ELL")):
        p = tmp_path / ("m_%d_%s.mtx" % (double, store))
        with open(p, "w") as f:
            write_matrix_market(f, rp, ci, v, 3, double, store)
        got = sio.mmread(str(p)).toarray()
        if double:
            assert np.array_equal(got, dense), store
        else:   # 9 significant digits name the float exactly
            assert np.array_equal(got.astype(np.float32), dense.astype(np.float32))
            assert not np.array_equal(got, dense)
        lines = open(p).read().split("\n")
        assert lines[1] == "3 3 6"
        if store == "ELL":   # slot 0 of rows 1..3, then slot 1, then slot 2
            pos = [tuple(map(int, ln.split()[:2])) for ln in lines[2:8]]
            assert pos == [(1, 1), (2, 1), (3, 2), (1, 3), (2, 2), (2, 3)], pos
