"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
Build with ``make -C oracle``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

_D = C.c_double
_I = C.c_int
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
FX = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
FXY = C.CFUNCTYPE(C.c_double, C.c_double, C.c_double, C.c_void_p)

KIND_MASS, KIND_LAPLACE, KIND_HELMHOLTZ, KIND_JAC_MASS = 0, 1, 2, 3
# SphereJacobian operators (src/sjacobian.cpp:76-134)
KIND_DIFF1, KIND_DIFF2, KIND_DIFF1_COS, KIND_DIFF2_COS = 4, 5, 6, 7


class OrPoly(C.Structure):
    _fields_ = [("x_deg", _I), ("y_deg", _I), ("koef", _D * 16)]


class OrMesh(C.Structure):
    _fields_ = [("n_points", _I), ("n_triangles", _I), ("n_zones", _I),
                ("coords", _PD), ("tri", _PI), ("tri_zone", _PI),
                ("adj_ptr", _PI), ("adj", _PI), ("is_boundary", _PI),
                ("p2io", _PI), ("inner", _PI), ("outer", _PI),
                ("n_inner", _I), ("n_outer", _I)]


class OrGmresInfo(C.Structure):
    _fields_ = [("iterations", _I), ("restarts", _I), ("rel_residual", _D),
                ("converged", _I)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        for name in ("or_cos", "or_sin"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D]
        L.or_ipow.restype = _D
        L.or_ipow.argtypes = [_D, _I]
        L.or_use_libm_cos.argtypes = [_I]
        L.or_record_cos_args.argtypes = [_PD, C.c_long]
        L.or_recorded_cos_args.restype = C.c_long
        for name in ("or_cr_cos_bound", "or_cr_sin_bound"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_D, _D]
        L.or_cr_slow.restype = _D
        L.or_cr_slow.argtypes = [_D, _I]
        L.or_cr_fast_dd.argtypes = [_D, _I, _PD, _PD]
        L.or_gauss_kronrod15.restype = _D
        L.or_gauss_kronrod15.argtypes = [_D, _D, FX, C.c_void_p]
        L.or_cubature7.restype = _D
        L.or_cubature7.argtypes = [_D] * 6 + [FXY, C.c_void_p]
        for name in ("or_trapezoid_integral", "or_trapezoid_integral_cos",
                     "or_trapezoid_integral_sin", "or_trapezoid_integral_1_cos"):
            getattr(L, name).restype = _D
            getattr(L, name).argtypes = [_I, _I] + [_D] * 6
        L.or_integrate.restype = _D
        L.or_integrate.argtypes = [C.POINTER(OrPoly), _PD, _PD, _I]
        L.or_basis_p1.argtypes = [_PD, _PD, C.POINTER(OrPoly)]
        L.or_poly_mul.argtypes = [C.POINTER(OrPoly)] * 3
        L.or_poly_diff.argtypes = [C.POINTER(OrPoly), _I, C.POINTER(OrPoly)]
        L.or_poly_apply.restype = _D
        L.or_poly_apply.argtypes = [C.POINTER(OrPoly), _D, _D]
        L.or_element.restype = _D
        L.or_element.argtypes = [_I, C.POINTER(OrPoly), C.POINTER(OrPoly), _PD,
                                 _PD, _D, _D]
        L.or_csr_structure.restype = C.c_long
        L.or_csr_structure.argtypes = [C.POINTER(OrMesh), _PI, _PI]
        L.or_boundary_structure.restype = C.c_long
        L.or_boundary_structure.argtypes = [C.POINTER(OrMesh), _PI, _PI]
        L.or_assemble.argtypes = [C.POINTER(OrMesh), _I, _D, _D, _PI, _PI, _PD]
        L.or_full_structure.restype = C.c_long
        L.or_full_structure.argtypes = [C.POINTER(OrMesh), _PI, _PI]
        L.or_assemble_full.argtypes = [C.POINTER(OrMesh), _I, _D, _D, _PI, _PI, _PD]
        L.or_assemble_boundary.argtypes = [C.POINTER(OrMesh), _I, _D, _D, _PI,
                                           _PI, _PD]
        L.or_assemble_baroclin.argtypes = [C.POINTER(OrMesh), _PD, _PI, _PI, _PD]
        L.or_right_part_mass.argtypes = [C.POINTER(OrMesh), _PD, _PD]
        L.or_spmv.argtypes = [_I, _PI, _PI, _PD, _PD, _PD]
        L.or_gmres.restype = _I
        L.or_gmres.argtypes = [_I, _PI, _PI, _PD, _PD, _PD, _I, _D, _I,
                               C.POINTER(OrGmresInfo)]
        L.or_num_threads.restype = _I
        L.or_set_num_threads.argtypes = [_I]
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_PD)


def _pi(a):
    return a.ctypes.data_as(_PI)


def poly(x_deg, y_deg, koef):
    p = OrPoly()
    p.x_deg, p.y_deg = x_deg, y_deg
    for i, c in enumerate(koef):
        p.koef[i] = c
    return p


def poly_koef(p):
    return [p.koef[i] for i in range((p.x_deg + 1) * (p.y_deg + 1))]


def poly_print(p):
    """Polynom::print (src/polynom.cpp:98-145) as a string."""
    out = []
    for i in range(p.x_deg + 1):
        for j in range(p.y_deg + 1):
            k = p.koef[i * (p.y_deg + 1) + j]
            if abs(k) > 1e-8:
                if i and j:
                    out.append("%+.1f x^%d y^%d " % (k, i, j))
                elif i:
                    out.append("%+.1f x^%d " % (k, i))
                elif j:
                    out.append("%+.1f y^%d " % (k, j))
                else:
                    out.append("%+.1f " % k)
    return "".join(out)


class Mesh:
    """Holds the numpy arrays of a loaded mesh and its OrMesh view."""

    def __init__(self, arrays):
        a = arrays
        self.arrays = a
        self.coords = np.ascontiguousarray(a["coords"], dtype=np.float64)
        self.tri = np.ascontiguousarray(a["tri"], dtype=np.int32)
        self.tri_zone = np.ascontiguousarray(a["tri_zone"], dtype=np.int32)
        self.adj_ptr = np.ascontiguousarray(a["adj_ptr"], dtype=np.int32)
        self.adj = np.ascontiguousarray(a["adj"], dtype=np.int32)
        self.is_boundary = np.ascontiguousarray(a["is_boundary"], dtype=np.int32)
        self.p2io = np.ascontiguousarray(a["p2io"], dtype=np.int32)
        self.inner = np.ascontiguousarray(a["inner"], dtype=np.int32)
        self.outer = np.ascontiguousarray(a["outer"], dtype=np.int32)
        m = OrMesh()
        m.n_points = self.coords.shape[0]
        m.n_triangles = self.tri.shape[0]
        m.n_zones = self.coords.shape[1]
        m.coords, m.tri, m.tri_zone = _pd(self.coords), _pi(self.tri), _pi(self.tri_zone)
        m.adj_ptr, m.adj = _pi(self.adj_ptr), _pi(self.adj)
        m.is_boundary, m.p2io = _pi(self.is_boundary), _pi(self.p2io)
        m.inner, m.outer = _pi(self.inner), _pi(self.outer)
        m.n_inner, m.n_outer = len(self.inner), len(self.outer)
        self.c = m

    @property
    def n_inner(self):
        return len(self.inner)

    def csr_structure(self):
        n = self.n_inner
        rp = np.zeros(n + 1, dtype=np.int32)
        nnz = lib().or_csr_structure(C.byref(self.c), _pi(rp), None)
        ci = np.zeros(nnz, dtype=np.int32)
        lib().or_csr_structure(C.byref(self.c), _pi(rp), _pi(ci))
        return rp, ci

    def assemble(self, kind, c_mass=1.0, c_lap=0.0, structure=None):
        rp, ci = structure if structure is not None else self.csr_structure()
        vals = np.zeros(len(ci))
        lib().or_assemble(C.byref(self.c), kind, c_mass, c_lap, _pi(rp), _pi(ci),
                          _pd(vals))
        return rp, ci, vals

    def full_structure(self):
        V = self.coords.shape[0]
        rp = np.zeros(V + 1, dtype=np.int32)
        nnz = lib().or_full_structure(C.byref(self.c), _pi(rp), None)
        ci = np.zeros(nnz, dtype=np.int32)
        lib().or_full_structure(C.byref(self.c), _pi(rp), _pi(ci))
        return rp, ci

    def assemble_full(self, kind, c_mass=1.0, c_lap=0.0):
        """generate_full_matrix (src/generators.h:200-234), V x V."""
        rp, ci = self.full_structure()
        vals = np.zeros(len(ci))
        lib().or_assemble_full(C.byref(self.c), kind, c_mass, c_lap, _pi(rp), _pi(ci), _pd(vals))
        return rp, ci, vals

    def boundary_structure(self):
        n = self.n_inner
        rp = np.zeros(n + 1, dtype=np.int32)
        nnz = lib().or_boundary_structure(C.byref(self.c), _pi(rp), None)
        ci = np.zeros(max(nnz, 1), dtype=np.int32)
        lib().or_boundary_structure(C.byref(self.c), _pi(rp), _pi(ci))
        return rp, ci[:nnz]

    def assemble_boundary(self, kind, c_mass=1.0, c_lap=0.0):
        rp, ci = self.boundary_structure()
        vals = np.zeros(max(len(ci), 1))
        lib().or_assemble_boundary(C.byref(self.c), kind, c_mass, c_lap,
                                   _pi(rp), _pi(np.ascontiguousarray(ci) if len(ci) else np.zeros(1, np.int32)),
                                   _pd(vals))
        return rp, ci, vals[:len(ci)]


def baroclin_blocks(mesh, tau, sigma, mu, sigma1, mu1, alpha, theta=0.5, backward=False):
    """The nine n x n blocks of Baroclin's A_ (examples/baroclin.cpp:115-181),
    or of Ab_ with backward (integrate_backward_cb, :184-244), on the inner
    stencil: (rp, ci, vals9) with vals9[k] the values of block k in the order
    (1,1) (1,2) (2,1) (2,2) (2,4) (3,1) (3,3) (4,2) (4,4)."""
    rp, ci = mesh.csr_structure()
    vals = np.zeros(9 * len(ci))
    prm = np.array([tau, sigma, mu, sigma1, mu1, alpha, theta, 1.0 if backward else 0.0],
                   dtype=np.float64)
    lib().or_assemble_baroclin(C.byref(mesh.c), _pd(prm), _pi(rp), _pi(ci), _pd(vals))
    return rp, ci, vals.reshape(9, len(ci))


BAROCLIN_BLOCKS = ((0, 0), (0, 1), (1, 0), (1, 1), (1, 3), (2, 0), (2, 2), (3, 1), (3, 3))


def baroclin_matrix(rp, ci, vals9):
    """The 4n x 4n CSR of the nine blocks (block (r, c) at rows r n.., columns
    c n..), columns ascending in each row."""
    n = len(rp) - 1
    rows = [[] for _ in range(4)]
    for k, (r, c) in enumerate(BAROCLIN_BLOCKS):
        rows[r].append((c, k))
    brp = [0]
    bci, bv = [], []
    for r in range(4):
        blocks = sorted(rows[r])
        for i in range(n):
            for c, k in blocks:
                s, e = rp[i], rp[i + 1]
                bci.append(ci[s:e] + c * n)
                bv.append(vals9[k, s:e])
            brp.append(brp[-1] + sum(rp[i + 1] - rp[i] for _ in blocks))
    return (np.asarray(brp, dtype=np.int32), np.concatenate(bci).astype(np.int32),
            np.concatenate(bv))


def right_part_mass(mesh, F):
    """Element-order sums of F_j * mass element per inner row (the interior
    branch of Baroclin's right_part_cb under generate_right_part)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    b = np.zeros(mesh.n_inner)
    lib().or_right_part_mass(C.byref(mesh.c), _pd(F), _pd(b))
    return b


def spmv(rp, ci, vals, x):
    n = len(rp) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(n)
    lib().or_spmv(n, _pi(rp), _pi(ci), _pd(vals), _pd(x), _pd(y))
    return y


def gmres(rp, ci, vals, b, x0=None, restart=30, rtol=1e-10, max_iters=100000):
    n = len(rp) - 1
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    info = OrGmresInfo()
    lib().or_gmres(n, _pi(rp), _pi(ci), _pd(vals), _pd(b), _pd(x), restart,
                   rtol, max_iters, C.byref(info))
    return x, dict(iterations=info.iterations, restarts=info.restarts,
                   rel_residual=info.rel_residual, converged=bool(info.converged))
