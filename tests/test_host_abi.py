"""Host side of the C ABI, no GPU needed: the library exports every entry
point include/phelm_gpu.h declares (and the Python mirror binds each one),
and the host-only entry points (mesh generation, Mesh::load, CSR structure)
agree with the oracle's restatement of the reference."""
import os
import re

import numpy as np
import pytest

import mke_amd as P
from mke_amd import phelm
from oracle import mesh_ref
from tests._common import oracle_mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "phelm_gpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phg_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported_and_bound():
    names = _declared()
    assert len(names) >= 30
    L = P.lib()
    for name in names:
        assert hasattr(L, name), name
        assert name in phelm.SIGNATURES, "no ctypes signature for " + name


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_mesh_text_matches_oracle_bytes(level):
    # utils/icosahedron.cpp:481-624 output (`--type full --coord local`)
    assert P.sphere_mesh_text(level) == mesh_ref.sphere_mesh_text(level)


@pytest.mark.parametrize("level", [1, 3])
def test_mesh_load_arrays_match_oracle(level):
    text, om = oracle_mesh(level)
    a = P.Mesh.load(text).arrays()
    ref = mesh_ref.load_text(text)
    for key in ("coords", "tri", "tri_zone", "adj_ptr", "adj", "is_boundary", "inner",
                "outer", "p2io"):
        assert np.array_equal(np.asarray(a[key]), np.asarray(ref[key]).reshape(np.shape(a[key]))), key


@pytest.mark.parametrize("level", [1, 2, 3, 4, 5])
def test_mesh_sizes_and_csr_structure(level):
    m = P.Mesh.sphere(level)
    V = 10 * 4 ** level + 2
    assert (m.n_points, m.n_triangles, m.inner_size, m.outer_size) == (V, 20 * 4 ** level, V - 1, 1)
    a = m.arrays()
    # the single outer point is the south pole, placed last by the ordering
    assert a["outer"][0] == V - 1 and abs(a["pr"][V - 1, 2] + 1.0) < 1e-14
    deg = np.diff(a["adj_ptr"])
    assert sorted(set(deg.tolist())) == [5, 6] and np.sum(deg == 5) == 12
    rp, ci = m.csr_structure()
    _, om = oracle_mesh(level)
    orp, oci = om.csr_structure()
    assert len(ci) == 70 * 4 ** level - 11
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    # columns ascending within each row, diagonal present
    for i in range(0, len(rp) - 1, max(1, (len(rp) - 1) // 200)):
        row = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row


def test_mesh_load_rejects_truncated_text():
    text = P.sphere_mesh_text(2)
    with pytest.raises(P.PhelmError, match="bad file format"):
        P.Mesh.load(text[: len(text) // 2])
    with pytest.raises(P.PhelmError):
        P.Mesh.load("")


def test_cpp_host_classes_compile():
    """include/phelm_gpu.hpp (the C++ Solver / SphereLaplace / SphereJacobian /
    SphereNorm / BarVortex classes over the C ABI) compiles with g++."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "syntax"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_argument_errors_without_a_gpu():
    """Entry points check their arguments before any device work and report
    PHG_ERR_ARG with a message, as the C++ classes then throw."""
    import ctypes as C
    L = P.lib()
    ERR_ARG = -1
    for name, args in (
        ("phg_barvortex_calc_L", (None, 1, None, None, None, None, None, None)),
        ("phg_barvortex_calc_L_1", (None, 1, None, None, None, None, None, None)),
        ("phg_barvortex_calc_LT", (None, 1, None, None, None, None, None, None)),
        ("phg_barvortex_calc_rp", (None, 1, None, None, None, None, None, None)),
        ("phg_barvortex_calc_bnd", (None, 1) + (None,) * 8),
        ("phg_barvortex_load_bnd", (None, 0, None, None, None)),
        ("phg_baroclin_calc_bnd", (None, 1) + (None,) * 10),
        ("phg_baroclin_calc_L", (None, 1) + (None,) * 9),
        ("phg_baroclin_calc_L_1", (None, 1) + (None,) * 9),
    ):
        assert getattr(L, name)(*args) == ERR_ARG, name
        assert L.phg_last_error().decode(), name
    m = P.Mesh.sphere(2)
    conf = phelm.BvConf(1.0, 0.0, 0.0, 1.0, 1.0, 0.5)
    h = C.c_void_p()
    lh = C.c_void_p(1)  # never dereferenced: the slot count is refused first
    assert L.phg_barvortex_create(m.h, C.byref(conf), lh, 3, None, C.byref(h)) == ERR_ARG
    assert "slots" in L.phg_last_error().decode()
    # a CSR with entries but no column or value array is refused before it is read
    import numpy as np
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    pi = rp.ctypes.data_as(C.POINTER(C.c_int))
    assert L.phg_matrix_from_csr(2, 2, pi, None, None, C.byref(h)) == ERR_ARG
    assert L.phg_matrix_from_csr(2, 2, pi, ci.ctypes.data_as(C.POINTER(C.c_int)), None,
                                 C.byref(h)) == ERR_ARG
    assert "col_idx or vals" in L.phg_last_error().decode()


def test_dot_noalloc_refuses_bad_sizes_before_device_work():
    """phg_dot_noalloc checks n against the 4096 block partials of its
    device scratch (n <= 4,194,304) and its pointers before any device work,
    so this runs without a GPU."""
    import ctypes as C
    L = P.lib()
    dummy = C.c_void_p(8)
    assert L.phg_dot_noalloc(4194305, dummy, dummy, dummy, None) == -1
    assert b"4194304" in L.phg_last_error()
    assert L.phg_dot_noalloc(-1, dummy, dummy, dummy, None) == -1
    assert L.phg_dot_noalloc(16, dummy, dummy, None, None) == -1
    assert b"null" in L.phg_last_error()
