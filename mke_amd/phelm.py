"""Python host mirror of the reference's solver/operator interface over the
C ABI of include/phelm_gpu.h (libphelm_hip.so, built in-tree for gfx950).

Names follow the reference: ``Mesh`` (src/mesh.h:354-427, Mesh::load),
``Solver`` (the external Solver<T>: ctor(n), add, mult_vector, solve),
``SphereLaplace`` (src/slaplace.h:60-108) and ``Helmholtz`` for the
BarVortex/SphereChafe implicit-step operator (examples/barvortex_impl.h:57-80,
examples/schafe.cpp:67-89).  Vectors are device tensors (torch is only used
for device memory and streams); all arithmetic runs in the HIP library.
There is no CPU fallback: if the library is missing, import fails.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libphelm_hip.so")

F64, F32, BF16 = 0, 1, 2
MASS, LAPLACE, HELMHOLTZ = 0, 1, 2
# SphereJacobian kinds (src/sjacobian.cpp:60-134)
JAC_MASS, DIFF1, DIFF2, DIFF1_COS, DIFF2_COS = 3, 4, 5, 6, 7
MGS, CGS2, CGS1 = 0, 1, 2
OK, ERR_NOCONV = 0, 1

_I, _D, _L, _P = C.c_int, C.c_double, C.c_long, C.c_void_p
_PI, _PD = C.POINTER(C.c_int), C.POINTER(C.c_double)


class GmresOpts(C.Structure):
    _fields_ = [("restart", _I), ("max_iters", _I), ("rtol", _D), ("ortho", _I),
                ("precision", _I), ("inner_rtol", _D), ("use_x0", _I), ("graph", _I),
                ("diag_scale", _I), ("persistent", _I)]


class GmresInfo(C.Structure):
    _fields_ = [("iterations", _I), ("restarts", _I), ("refinements", _I),
                ("rel_residual", _D), ("converged", _I), ("cycle_launches", _I),
                ("cycle_ms", _D)]


class BlockInfo(C.Structure):
    _fields_ = [("iterations", _I), ("cycles", _I), ("max_rel_residual", _D),
                ("converged", _I)]


# every symbol of include/phelm_gpu.h with (restype, argtypes)
SIGNATURES = {
    "phg_version": (C.c_char_p, []),
    "phg_last_error": (C.c_char_p, []),
    "phg_init": (_I, [_I]),
    "phg_mesh_sphere_text": (_I, [_I, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "phg_free": (None, [_P]),
    "phg_mesh_load": (_I, [C.c_char_p, C.c_size_t, C.POINTER(_P)]),
    "phg_mesh_sphere": (_I, [_I, C.POINTER(_P)]),
    "phg_mesh_info": (_I, [_P, _PI, _PI, _PI, _PI, _PI]),
    "phg_mesh_arrays": (_I, [_P, _PD, _PD, _PI, _PI, _PI, _PI, _PI, _PI, _PI, _PI]),
    "phg_mesh_destroy": (None, [_P]),
    "phg_mesh_csr_structure": (_I, [_P, _PI, _PI, C.POINTER(_L)]),
    "phg_matrix_from_triplets": (_I, [_I, _I, _L, _PI, _PI, _PD, C.POINTER(_P)]),
    "phg_matrix_from_csr": (_I, [_I, _I, _PI, _PI, _PD, C.POINTER(_P)]),
    "phg_assemble": (_I, [_P, _I, _D, _D, _P, C.POINTER(_P)]),
    "phg_assemble_boundary": (_I, [_P, _I, _D, _D, C.POINTER(_P)]),
    "phg_matrix_info": (_I, [_P, _PI, _PI, C.POINTER(_L)]),
    "phg_matrix_csr": (_I, [_P, _PI, _PI, _PD]),
    "phg_matrix_destroy": (None, [_P]),
    "phg_spmv": (_I, [_P, _P, _P, _P]),
    "phg_spmv_f32": (_I, [_P, _I, _P, _P, _P]),
    "phg_spmm_f32": (_I, [_P, _I, _I, _P, _P, _P]),
    "phg_dot": (_I, [_I, _P, _P, _PD, _P]),
    "phg_dot_noalloc": (_I, [_I, _P, _P, _P, _P]),
    "phg_nrm2": (_I, [_I, _P, _PD, _P]),
    "phg_axpby": (_I, [_I, _D, _P, _D, _P, _P, _P]),
    "phg_u2p": (_I, [_P, _P, _P, _P]),
    "phg_p2u": (_I, [_P, _P, _P, _P, _P]),
    "phg_proj_bnd": (_I, [_P, _P, _P, _P]),
    "phg_set_bnd": (_I, [_P, _P, _P, _P]),
    "phg_gmres_default_opts": (None, [C.POINTER(GmresOpts)]),
    "phg_gmres": (_I, [_P, _P, _P, C.POINTER(GmresOpts), C.POINTER(GmresInfo), _P]),
    "phg_block_gmres": (_I, [_P, _I, _P, _P, C.POINTER(GmresOpts), C.POINTER(BlockInfo), _P]),
    "phg_gmres_batch": (_I, [_P, _I, _P, _P, C.POINTER(GmresOpts), C.POINTER(GmresInfo), _P]),
    "phg_gmres_batch_tune": (_I, [_P, _I, C.POINTER(GmresOpts), _P, _PI, _PD]),
    "phg_gmres_batch_shape": (_I, [_I, _I, _I]),
    "phg_gmres_batch_set_shape": (_I, [_I, _I, _I, _I]),
    "phg_mesh_save": (_I, [_P, C.c_char_p]),
    "phg_mesh_load_file": (_I, [C.c_char_p, C.POINTER(_P)]),
    "phg_mesh_sphere_cached": (_I, [_I, C.c_char_p, C.POINTER(_P)]),
    "phg_print_function": (_I, [_P, _PD, _I, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]),
    "phg_mg_create": (_I, [_P, _I, _D, _D, _P, C.POINTER(_P)]),
    "phg_mg_destroy": (None, [_P]),
    "phg_mg_matrix": (_I, [_P, C.POINTER(_P)]),
    "phg_mg_levels": (_I, [_P]),
    "phg_gmres_batch_mg": (_I, [_P, _I, _P, _P, C.POINTER(GmresOpts), C.POINTER(GmresInfo), _P]),
    "phg_assemble_full": (_I, [_P, _I, _D, _D, _P, C.POINTER(_P)]),
    "phg_bv_default_opts": (None, [_P]),
    "phg_barvortex_estimate_bytes": (C.c_size_t, [_I, _I]),
    "phg_barvortex_create": (_I, [_P, _P, _P, _I, _P, C.POINTER(_P)]),
    "phg_barvortex_destroy": (None, [_P]),
    "phg_barvortex_slots": (_I, [_P]),
    "phg_barvortex_calc": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "phg_barvortex_load": (_I, [_P, _I, _P, _P]),
    "phg_barvortex_iterate": (_I, [_P, _P, _P, _P, _P]),
    "phg_barvortex_store": (_I, [_P, _I, _P, _P]),
    "phg_barvortex_calc_rp": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_load_rp": (_I, [_P, _I, _P, _P, _P]),
    "phg_barvortex_calc_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_baroclin_calc_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_load_bnd": (_I, [_P, _I, _P, _P, _P]),
    "phg_barvortex_calc_L": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_calc_L_1": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_calc_LT": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_calc_L_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_calc_L_1_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_barvortex_calc_LT_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_bc_default_opts": (None, [_P]),
    "phg_baroclin_create": (_I, [_P, _P, _P, _I, _P, C.POINTER(_P)]),
    "phg_baroclin_destroy": (None, [_P]),
    "phg_baroclin_slots": (_I, [_P]),
    "phg_baroclin_matrix": (_I, [_P, C.POINTER(_P)]),
    "phg_baroclin_calc": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_baroclin_calc_L": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_baroclin_calc_L_1": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_baroclin_calc_L_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_baroclin_calc_L_1_bnd": (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "phg_serve_default_config": (None, [_P]),
    "phg_serve_create": (_I, [_P, C.POINTER(_P)]),
    "phg_serve_destroy": (None, [_P]),
    "phg_serve_slots": (_I, [_P, _I]),
    "phg_serve_add_member": (_I, [_P, _I, _PD, _PI]),
    "phg_serve_member_state": (_I, [_P, _I, _PD]),
    "phg_serve_run": (_I, [_P, _I, _PI, _PD, _PD, _PD, _P, _P]),
    "phg_jacobian_create": (_I, [_P, _P, C.POINTER(_P)]),
    "phg_jacobian_destroy": (None, [_P]),
    "phg_jacobian_matrix": (_I, [_P, _I, C.POINTER(_P)]),
    "phg_jacobian_calc2": (_I, [_P, _I, _P, _P, _P, C.POINTER(GmresOpts), C.POINTER(GmresInfo), _P]),
    "phg_jacobian_calc2t": (_I, [_P, _I, _P, _P, _P, C.POINTER(GmresOpts), C.POINTER(GmresInfo), _P]),
}

_lib = None


class PhelmError(RuntimeError):
    pass


def lib():
    """Loads libphelm_hip.so; raises if it was not built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PhelmError("libphelm_hip.so not built: run __graft_entry__.build() "
                             "or python -m mke_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what="phelm call"):
    if rc < 0:
        raise PhelmError("%s failed (%d): %s" % (what, rc, lib().phg_last_error().decode()))
    return rc


def _ptr(t):
    """Device pointer of a torch tensor (or an int address)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    return getattr(stream, "cuda_stream", stream)


def _pi(a):
    return a.ctypes.data_as(_PI)


def _pd(a):
    return a.ctypes.data_as(_PD)


def batch_shape(n_rows, nrhs, precision=F64):
    """Rows per thread the batched solves use for this shape (process-wide)."""
    return check(lib().phg_gmres_batch_shape(int(n_rows), int(nrhs), int(precision)),
                 "phg_gmres_batch_shape")


def set_batch_shape(n_rows, nrhs, precision, rows_per_thread):
    """Pins the rows per thread of a batched-solve shape (1, 2, 4, 8 or 16; 0
    returns it to the untuned default)."""
    check(lib().phg_gmres_batch_set_shape(int(n_rows), int(nrhs), int(precision),
                                          int(rows_per_thread)), "phg_gmres_batch_set_shape")


def sphere_mesh_text(level):
    """Text of ``mke_mesh_sphere --type full --coord local --iter level``."""
    buf, n = C.c_void_p(), C.c_size_t()
    check(lib().phg_mesh_sphere_text(level, C.byref(buf), C.byref(n)), "phg_mesh_sphere_text")
    try:
        return C.string_at(buf, n.value).decode()
    finally:
        lib().phg_free(buf)


class Mesh:
    """Mesh::load result (src/mesh.cpp:323-519)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if isinstance(handle, int) else handle
        a = [C.c_int() for _ in range(5)]
        check(lib().phg_mesh_info(self.h, *[C.byref(v) for v in a]), "phg_mesh_info")
        (self.n_points, self.n_triangles, self.inner_size, self.outer_size,
         self.n_zones) = [v.value for v in a]

    @classmethod
    def load(cls, text):
        if isinstance(text, str):
            text = text.encode()
        h = C.c_void_p()
        check(lib().phg_mesh_load(text, len(text), C.byref(h)), "phg_mesh_load")
        return cls(h)

    @classmethod
    def sphere(cls, level):
        h = C.c_void_p()
        check(lib().phg_mesh_sphere(level, C.byref(h)), "phg_mesh_sphere")
        return cls(h)

    @classmethod
    def load_file(cls, path):
        """A mesh from its binary image (save)."""
        h = C.c_void_p()
        check(lib().phg_mesh_load_file(os.fsencode(path), C.byref(h)), "phg_mesh_load_file")
        return cls(h)

    @classmethod
    def sphere_cached(cls, level, directory):
        """The level's sphere from directory/sphere_L<level>.phgmesh, written
        there on first use."""
        h = C.c_void_p()
        check(lib().phg_mesh_sphere_cached(level, os.fsencode(directory), C.byref(h)),
              "phg_mesh_sphere_cached")
        return cls(h)

    def save(self, path):
        """Writes the binary image of the mesh."""
        check(lib().phg_mesh_save(self.h, os.fsencode(path)), "phg_mesh_save")

    def print_function(self, u, sphere_xyz=False):
        """print_function (src/phelm.cpp:248-309) text of the host values u."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.n_points,):
            raise PhelmError("print_function: u must have n_points values")
        buf, n = C.c_void_p(), C.c_size_t()
        check(lib().phg_print_function(self.h, _pd(u), int(sphere_xyz), C.byref(buf),
                                       C.byref(n)), "phg_print_function")
        try:
            return C.string_at(buf, n.value).decode()
        finally:
            lib().phg_free(buf)

    def arrays(self):
        V, T, nz = self.n_points, self.n_triangles, self.n_zones
        out = dict(pr=np.zeros((V, 3)), coords=np.zeros((V, nz, 2)),
                   tri=np.zeros((T, 3), np.int32), tri_zone=np.zeros(T, np.int32),
                   adj_ptr=np.zeros(V + 1, np.int32), adj=np.zeros(3 * T, np.int32),
                   is_boundary=np.zeros(V, np.int32), inner=np.zeros(self.inner_size, np.int32),
                   outer=np.zeros(max(self.outer_size, 1), np.int32), p2io=np.zeros(V, np.int32))
        check(lib().phg_mesh_arrays(self.h, _pd(out["pr"]), _pd(out["coords"]), _pi(out["tri"]),
                                    _pi(out["tri_zone"]), _pi(out["adj_ptr"]), _pi(out["adj"]),
                                    _pi(out["is_boundary"]), _pi(out["inner"]),
                                    _pi(out["outer"]), _pi(out["p2io"])), "phg_mesh_arrays")
        out["outer"] = out["outer"][:self.outer_size]
        return out

    def zone0_xy(self):
        """(x, y) = MeshPoint::x(), y() (src/mesh.h:126-148) of every point:
        its zone-0 chart coordinates, where the reference evaluates its
        callbacks (proj, conf.rp)."""
        return self.arrays()["coords"][:, 0, :].copy()

    def csr_structure(self):
        rp = np.zeros(self.inner_size + 1, np.int32)
        nnz = C.c_long()
        check(lib().phg_mesh_csr_structure(self.h, _pi(rp), None, C.byref(nnz)))
        ci = np.zeros(max(nnz.value, 1), np.int32)
        check(lib().phg_mesh_csr_structure(self.h, _pi(rp), _pi(ci), C.byref(nnz)))
        return rp, ci[:nnz.value]

    def u2p(self, u, p, stream=None):
        check(lib().phg_u2p(self.h, _ptr(u), _ptr(p), _stream(stream)), "phg_u2p")

    def p2u(self, p, bnd, u, stream=None):
        check(lib().phg_p2u(self.h, _ptr(p), _ptr(bnd), _ptr(u), _stream(stream)), "phg_p2u")

    def proj_bnd(self, u, b, stream=None):
        """b[k] = u[outer[k]] (proj_bnd, src/proj.cpp:140-149)."""
        check(lib().phg_proj_bnd(self.h, _ptr(u), _ptr(b), _stream(stream)), "phg_proj_bnd")

    def set_bnd(self, u, bnd, stream=None):
        """u's boundary points from bnd (None: zero), inner points untouched
        (set_bnd, src/proj.cpp:109-133)."""
        check(lib().phg_set_bnd(self.h, _ptr(bnd) if bnd is not None else None, _ptr(u),
                                _stream(stream)), "phg_set_bnd")

    def close(self):
        if self.h:
            lib().phg_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """A device CSR matrix (the storage behind Solver<T>)."""

    def __init__(self, handle, owned=True):
        self.h = handle
        self.owned = owned
        r, c, nnz = C.c_int(), C.c_int(), C.c_long()
        check(lib().phg_matrix_info(self.h, C.byref(r), C.byref(c), C.byref(nnz)))
        self.n_rows, self.n_cols, self.nnz = r.value, c.value, nnz.value

    @classmethod
    def assemble(cls, mesh, kind, c_mass=1.0, c_lap=0.0, stream=None):
        h = C.c_void_p()
        check(lib().phg_assemble(mesh.h, kind, c_mass, c_lap, _stream(stream), C.byref(h)),
              "phg_assemble")
        return cls(h)

    @classmethod
    def assemble_full(cls, mesh, kind, c_mass=1.0, c_lap=0.0, stream=None):
        """generate_full_matrix: every point x every point (SphereNorm's matrix
        with MASS)."""
        h = C.c_void_p()
        check(lib().phg_assemble_full(mesh.h, kind, c_mass, c_lap, _stream(stream), C.byref(h)),
              "phg_assemble_full")
        return cls(h)

    @classmethod
    def assemble_boundary(cls, mesh, kind, c_mass=1.0, c_lap=0.0):
        h = C.c_void_p()
        check(lib().phg_assemble_boundary(mesh.h, kind, c_mass, c_lap, C.byref(h)),
              "phg_assemble_boundary")
        return cls(h)

    @classmethod
    def from_csr(cls, n_rows, n_cols, row_ptr, col_idx, vals):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        v = np.ascontiguousarray(vals, np.float64)
        h = C.c_void_p()
        check(lib().phg_matrix_from_csr(n_rows, n_cols, _pi(rp), _pi(ci) if len(ci) else None,
                                        _pd(v) if len(v) else None, C.byref(h)),
              "phg_matrix_from_csr")
        return cls(h)

    @classmethod
    def from_triplets(cls, n_rows, n_cols, rows, cols, vals):
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        v = np.ascontiguousarray(vals, np.float64)
        h = C.c_void_p()
        check(lib().phg_matrix_from_triplets(n_rows, n_cols, len(r), _pi(r), _pi(c), _pd(v),
                                             C.byref(h)), "phg_matrix_from_triplets")
        return cls(h)

    def csr(self):
        rp = np.zeros(self.n_rows + 1, np.int32)
        ci = np.zeros(max(self.nnz, 1), np.int32)
        v = np.zeros(max(self.nnz, 1))
        check(lib().phg_matrix_csr(self.h, _pi(rp), _pi(ci), _pd(v)), "phg_matrix_csr")
        return rp, ci[:self.nnz], v[:self.nnz]

    def mult_vector(self, y, x, stream=None):
        """y = A x (Solver::mult_vector), FP64 device tensors."""
        check(lib().phg_spmv(self.h, _ptr(x), _ptr(y), _stream(stream)), "phg_spmv")

    def mult_vector_f32(self, y, x, prec=F32, stream=None):
        check(lib().phg_spmv_f32(self.h, prec, _ptr(x), _ptr(y), _stream(stream)), "phg_spmv_f32")

    def mult_matrix_f32(self, Y, X, nrhs, prec=F32, stream=None):
        check(lib().phg_spmm_f32(self.h, prec, nrhs, _ptr(X), _ptr(Y), _stream(stream)),
              "phg_spmm_f32")

    def solve(self, x, b, restart=30, rtol=1e-10, max_iters=200000, ortho=MGS, precision=F64,
              inner_rtol=1e-5, use_x0=False, graph=True, persistent=True, stream=None):
        """Solves A x = b (Solver::solve) with GMRES(restart); returns the info dict.
        persistent: MGS restart cycles as one cooperative kernel (same bits as
        a kernel per MGS pass)."""
        o = GmresOpts()
        lib().phg_gmres_default_opts(C.byref(o))
        o.restart, o.max_iters, o.rtol = restart, max_iters, rtol
        o.ortho, o.precision, o.inner_rtol, o.use_x0 = ortho, precision, inner_rtol, int(use_x0)
        o.graph = int(graph)
        o.persistent = int(persistent)
        info = GmresInfo()
        check(lib().phg_gmres(self.h, _ptr(b), _ptr(x), C.byref(o), C.byref(info),
                              _stream(stream)), "phg_gmres")
        return dict(iterations=info.iterations, restarts=info.restarts,
                    refinements=info.refinements, rel_residual=info.rel_residual,
                    converged=bool(info.converged), cycle_launches=info.cycle_launches,
                    cycle_ms=info.cycle_ms)

    def solve_block(self, X, B, restart=30, rtol=1e-10, max_iters=100000, precision=F32,
                    inner_rtol=1e-5, use_x0=False, ortho=CGS2, stream=None):
        """Block GMRES for 64 right-hand sides: X, B are n x 64 FP64 device
        tensors (row-major).  Returns the info dict."""
        o = GmresOpts()
        lib().phg_gmres_default_opts(C.byref(o))
        o.restart, o.max_iters, o.rtol, o.precision = restart, max_iters, rtol, precision
        o.inner_rtol, o.use_x0, o.ortho = inner_rtol, int(use_x0), ortho
        info = BlockInfo()
        check(lib().phg_block_gmres(self.h, 64, _ptr(B), _ptr(X), C.byref(o), C.byref(info),
                                    _stream(stream)), "phg_block_gmres")
        return dict(iterations=info.iterations, cycles=info.cycles,
                    max_rel_residual=info.max_rel_residual, converged=bool(info.converged))

    def solve_batch(self, X, B, restart=30, rtol=1e-10, max_iters=200000, ortho=CGS2,
                    precision=F64, inner_rtol=1e-5, use_x0=False, graph=True, diag_scale=False,
                    stream=None):
        """K independent solves A x_q = b_q sharing the matrix (1 <= K <= 64):
        X, B are n x K FP64 device tensors (row-major).  Returns one info
        dict per column."""
        K = int(B.shape[1]) if B.dim() == 2 else 1
        o = GmresOpts()
        lib().phg_gmres_default_opts(C.byref(o))
        o.restart, o.max_iters, o.rtol, o.ortho = restart, max_iters, rtol, ortho
        o.precision, o.inner_rtol, o.use_x0, o.graph = precision, inner_rtol, int(use_x0), int(graph)
        o.diag_scale = int(diag_scale)
        infos = (GmresInfo * K)()
        rc = lib().phg_gmres_batch(self.h, K, _ptr(B), _ptr(X), C.byref(o), infos, _stream(stream))
        check(rc, "phg_gmres_batch")
        return [dict(iterations=i.iterations, restarts=i.restarts, refinements=i.refinements,
                     rel_residual=i.rel_residual, converged=bool(i.converged)) for i in infos]

    def tune_batch(self, nrhs, restart=30, ortho=CGS2, precision=F64, stream=None):
        """Autotunes the rows per thread of the batched solves for this
        matrix's (n, nrhs, precision) shape: times one restart cycle per
        candidate (1, 2, 4, 8, 16), keeps the fastest for the process and returns
        (rows_per_thread, {candidate: ms})."""
        o = GmresOpts()
        lib().phg_gmres_default_opts(C.byref(o))
        o.restart, o.ortho, o.precision = restart, ortho, precision
        best = C.c_int(0)
        ms = (C.c_double * 5)()
        check(lib().phg_gmres_batch_tune(self.h, int(nrhs), C.byref(o), _stream(stream),
                                         C.byref(best), ms), "phg_gmres_batch_tune")
        return best.value, {r: ms[i] for i, r in enumerate((1, 2, 4, 8, 16))}

    def close(self):
        if self.h:
            if self.owned:
                lib().phg_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """Solver<T>(n) of the reference's solver library: add(i, j, a) stages
    triplets on the host (duplicates accumulate in call order); the first
    mult_vector/solve builds the device matrix."""

    def __init__(self, n_rows, n_cols=None):
        self.n_rows = n_rows
        self.n_cols = n_rows if n_cols is None else n_cols
        self._r, self._c, self._v = [], [], []
        self._m = None

    def add(self, i, j, a):
        self._r.append(i)
        self._c.append(j)
        self._v.append(a)
        self._m = None

    def matrix(self):
        if self._m is None:
            self._m = Matrix.from_triplets(self.n_rows, self.n_cols, self._r, self._c, self._v)
        return self._m

    def mult_vector(self, y, x, stream=None):
        self.matrix().mult_vector(y, x, stream)

    def solve(self, x, b, **kw):
        return self.matrix().solve(x, b, **kw)


def _opts(restart, rtol, max_iters, ortho, precision, inner_rtol=1e-5, use_x0=False, graph=True):
    o = GmresOpts()
    lib().phg_gmres_default_opts(C.byref(o))
    o.restart, o.max_iters, o.rtol, o.ortho = restart, max_iters, rtol, ortho
    o.precision, o.inner_rtol, o.use_x0, o.graph = precision, inner_rtol, int(use_x0), int(graph)
    return o


class Multigrid:
    """V-cycle preconditioner over the nested sphere meshes of levels 2..L and
    its fine operator (phg_mg_*); solve_batch runs the right-preconditioned
    batched GMRES (FP64)."""

    def __init__(self, mesh, kind=LAPLACE, c_mass=1.0, c_lap=0.0, stream=None):
        self.m = mesh
        self.h = C.c_void_p()
        check(lib().phg_mg_create(mesh.h, kind, c_mass, c_lap, _stream(stream), C.byref(self.h)),
              "phg_mg_create")
        mh = C.c_void_p()
        check(lib().phg_mg_matrix(self.h, C.byref(mh)))
        self.A = Matrix(mh, owned=False)
        self.levels = lib().phg_mg_levels(self.h)

    def solve_batch(self, X, B, restart=30, rtol=1e-10, max_iters=10000, ortho=CGS2,
                    use_x0=False, graph=True, precision=F64, inner_rtol=1e-5, stream=None):
        """K right-preconditioned solves (one V-cycle per step); precision F32
        or BF16 runs the Krylov vectors and the V-cycle in FP32 under FP64
        refinement."""
        K = int(B.shape[1]) if B.dim() == 2 else 1
        o = _opts(restart, rtol, max_iters, ortho, precision, inner_rtol=inner_rtol,
                  use_x0=use_x0, graph=graph)
        infos = (GmresInfo * K)()
        check(lib().phg_gmres_batch_mg(self.h, K, _ptr(B), _ptr(X), C.byref(o), infos,
                                       _stream(stream)), "phg_gmres_batch_mg")
        return [dict(iterations=i.iterations, restarts=i.restarts, rel_residual=i.rel_residual,
                     converged=bool(i.converged)) for i in infos]

    def close(self):
        if self.h:
            lib().phg_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BvConf(C.Structure):
    _fields_ = [("tau", _D), ("sigma", _D), ("mu", _D), ("k1", _D), ("k2", _D),
                ("theta", _D)]


class BvOpts(C.Structure):
    _fields_ = [("max_iters", _I), ("tol", _D), ("rtol_mass", _D), ("rtol_helmholtz", _D),
                ("rtol_poisson", _D), ("precision", _I), ("max_solver_iters", _I),
                ("stall_floor", _D)]


class BvInfo(C.Structure):
    _fields_ = [("iterations", _I), ("rel_change", _D), ("converged", _I), ("stalled", _I),
                ("diverged", _I),
                ("solver_iterations", _I), ("it_mass", _I), ("it_jacobian", _I),
                ("it_helmholtz", _I), ("it_poisson", _I)]


def _bv_info(i):
    return dict(iterations=i.iterations, rel_change=i.rel_change, converged=bool(i.converged),
                stalled=bool(i.stalled), diverged=bool(i.diverged),
                solver_iterations=i.solver_iterations, it_mass=i.it_mass,
                it_jacobian=i.it_jacobian, it_helmholtz=i.it_helmholtz, it_poisson=i.it_poisson)


def bv_opts(**kw):
    o = BvOpts()
    lib().phg_bv_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class BarVortex:
    """BarVortex (examples/barvortex.h:62-155): the implicit step
    BarVortex::calc on the GPU for up to `slots` members at once; lh is the
    Coriolis + orography mesh function (device tensor of n_points values)."""

    def __init__(self, mesh, lh, tau, sigma=0.0, mu=0.0, k1=1.0, k2=1.0, theta=0.5, slots=1,
                 stream=None):
        self.m = mesh
        self.conf = BvConf(tau, sigma, mu, k1, k2, theta)
        self.h = C.c_void_p()
        check(lib().phg_barvortex_create(mesh.h, C.byref(self.conf), _ptr(lh), slots,
                                         _stream(stream), C.byref(self.h)), "phg_barvortex_create")
        self.slots = slots

    def calc(self, u1, u, opts=None, stream=None, rp=None, bnd_u=None, bnd_w=None):
        """u, u1: (nmembers, n_points) FP64 device tensors (or (n_points,));
        rp: the right part conf.rp(x, y, t, mu, sigma) at every point for the
        step's t, same shape as u (or None); bnd_u, bnd_w: the new-time
        boundary values of psi and of lap psi, (nmembers, outer_size) (or
        (outer_size,)) FP64 device tensors, or None for zeros."""
        k = 1 if u.dim() == 1 else int(u.shape[0])
        if rp is not None and tuple(rp.shape) != tuple(u.shape):
            raise PhelmError("BarVortex.calc: rp must have the shape of u")
        self._check_bnd("BarVortex.calc", k, bnd_u, bnd_w)
        infos = (BvInfo * k)()
        opt = lambda t: _ptr(t) if t is not None else None
        rc = lib().phg_barvortex_calc_bnd(self.h, k, _ptr(u), opt(rp), opt(bnd_u), opt(bnd_w),
                                          _ptr(u1), C.byref(opts) if opts is not None else None,
                                          infos, _stream(stream))
        check(rc, "phg_barvortex_calc")
        return [_bv_info(i) for i in infos]

    def _check_bnd(self, what, k, *bnds):
        for b in bnds:
            if b is not None and (b.numel() != k * self.m.outer_size or
                                  str(b.dtype) != "torch.float64" or not b.is_contiguous()):
                raise PhelmError("%s: bnd_u/bnd_w must be contiguous FP64 with "
                                 "outer_size values per member" % what)

    def rp_values(self, fn, t):
        """conf.rp evaluated as BarVortex::calc does (barvortex_impl.h:233-244):
        fn(x, y, t, mu, sigma) at the zone-0 chart coordinates, a host array."""
        if not hasattr(self, "_xy"):
            self._xy = self.m.zone0_xy()
        x, y = self._xy[:, 0], self._xy[:, 1]
        return np.asarray(fn(x, y, t, self.conf.mu, self.conf.sigma), dtype=np.float64)

    def _lin(self, name, out, a, z, opts, stream, bnd_u=None, bnd_w=None):
        k = 1 if a.dim() == 1 else int(a.shape[0])
        if tuple(z.shape) != tuple(a.shape) or tuple(out.shape) != tuple(a.shape):
            raise PhelmError("BarVortex.%s: arguments must have one shape" % name)
        self._check_bnd("BarVortex." + name, k, bnd_u, bnd_w)
        infos = (BvInfo * k)()
        opt = lambda t: _ptr(t) if t is not None else None
        rc = getattr(lib(), "phg_barvortex_%s_bnd" % name)(
            self.h, k, _ptr(a), _ptr(z), opt(bnd_u), opt(bnd_w), _ptr(out),
            C.byref(opts) if opts is not None else None, infos, _stream(stream))
        check(rc, "phg_barvortex_" + name)
        return [_bv_info(i) for i in infos]

    def calc_L(self, u1, u, z, opts=None, stream=None, bnd_u=None, bnd_w=None):
        """BarVortex::calc_L(u1, u, z, bnd_u, bnd_w, t)
        (examples/barvortex_impl.h:326-400); bnd_u, bnd_w as in calc (None:
        zeros)."""
        return self._lin("calc_L", u1, u, z, opts, stream, bnd_u, bnd_w)

    def calc_L_1(self, u1, u, z, opts=None, stream=None, bnd_u=None, bnd_w=None):
        """BarVortex::calc_L_1 (examples/barvortex_impl.h:402-467): adds to
        u1's content, as the reference does."""
        return self._lin("calc_L_1", u1, u, z, opts, stream, bnd_u, bnd_w)

    def calc_LT(self, v1, v, z, opts=None, stream=None, bnd_u=None, bnd_w=None):
        """BarVortex::calc_LT(v1, v, z, bnd_u, bnd_w, t)
        (examples/barvortex_impl.h:469-535)."""
        return self._lin("calc_LT", v1, v, z, opts, stream, bnd_u, bnd_w)

    def S_step(self, ans, F, opts=None, stream=None):
        """BarVortex::S_step (examples/barvortex_impl.h:538-542): calc(ans, F, 0, 0, 0)."""
        return self.calc(ans, F, opts, stream)

    def L_step(self, ans, F, z, opts=None, stream=None):
        """BarVortex::L_step (examples/barvortex_impl.h:548-554)."""
        return self.calc_L(ans, F, z, opts, stream)

    def L_1_step(self, ans, F, z, opts=None, stream=None):
        """BarVortex::L_1_step (examples/barvortex_impl.h:556-562): calc_L_1
        into zeros (a new buffer, so ans may be F or z)."""
        import torch
        tmp = torch.zeros_like(F)
        infos = self.calc_L_1(tmp, F, z, opts, stream)
        ans.copy_(tmp)
        return infos

    def LT_step(self, ans, F, z, opts=None, stream=None):
        """BarVortex::LT_step (examples/barvortex_impl.h:564-570)."""
        return self.calc_LT(ans, F, z, opts, stream)

    def load(self, slot, u, stream=None, rp=None, bnd_u=None, bnd_w=None):
        n_points = self.m.n_points
        for t in (u, rp):
            if t is not None and (t.numel() != n_points or str(t.dtype) != "torch.float64"
                                  or not t.is_contiguous()):
                raise PhelmError("BarVortex.load: u/rp must be contiguous FP64 with n_points "
                                 "values")
        self._check_bnd("BarVortex.load", 1, bnd_u, bnd_w)
        check(lib().phg_barvortex_load_rp(self.h, slot, _ptr(u), _ptr(rp) if rp is not None else None,
                                          _stream(stream)), "phg_barvortex_load")
        if bnd_u is not None or bnd_w is not None:
            opt = lambda t: _ptr(t) if t is not None else None
            check(lib().phg_barvortex_load_bnd(self.h, slot, opt(bnd_u), opt(bnd_w),
                                               _stream(stream)), "phg_barvortex_load_bnd")

    def iterate(self, opts=None, stream=None):
        state = (C.c_int * self.slots)()
        infos = (BvInfo * self.slots)()
        check(lib().phg_barvortex_iterate(self.h, C.byref(opts) if opts is not None else None,
                                          state, infos, _stream(stream)), "phg_barvortex_iterate")
        return list(state), [_bv_info(i) for i in infos]

    def store(self, slot, u1, stream=None):
        check(lib().phg_barvortex_store(self.h, slot, _ptr(u1), _stream(stream)),
              "phg_barvortex_store")

    def close(self):
        if self.h:
            lib().phg_barvortex_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BcConf(C.Structure):
    _fields_ = [("tau", _D), ("sigma", _D), ("mu", _D), ("sigma1", _D), ("mu1", _D),
                ("alpha", _D)]


class BcOpts(C.Structure):
    _fields_ = [("max_iters", _I), ("tol", _D), ("rtol_mass", _D), ("rtol_system", _D),
                ("max_solver_iters", _I), ("precondition", _I)]


class BcInfo(C.Structure):
    _fields_ = [("iterations", _I), ("distance", _D), ("converged", _I), ("it_mass", _I),
                ("it_jacobian", _I), ("it_system", _I)]


def bc_opts(**kw):
    o = BcOpts()
    lib().phg_bc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Baroclin:
    """Baroclin (examples/baroclin.h:60-292): the two-layer implicit step
    Baroclin::calc on the GPU for up to `slots` members at once; lh is the
    Coriolis + orography mesh function (device tensor of n_points values)."""

    def __init__(self, mesh, lh, tau, sigma, mu, sigma1, mu1, alpha, slots=1, stream=None):
        self.m = mesh
        self.conf = BcConf(tau, sigma, mu, sigma1, mu1, alpha)
        self.h = C.c_void_p()
        check(lib().phg_baroclin_create(mesh.h, C.byref(self.conf), _ptr(lh), slots,
                                        _stream(stream), C.byref(self.h)), "phg_baroclin_create")
        self.slots = slots

    def matrix(self):
        """A_ (4n x 4n), owned by this object."""
        h = C.c_void_p()
        check(lib().phg_baroclin_matrix(self.h, C.byref(h)), "phg_baroclin_matrix")
        return Matrix(h, owned=False)

    def calc(self, u11, u21, u1, u2, f=None, g=None, opts=None, stream=None, bnd=None):
        """u1, u2, u11, u21: (nmembers, n_points) FP64 device tensors (or
        (n_points,)); f, g: the forcing at the mesh points for this step's t
        (device tensors of n_points values) or None; bnd: the boundary values
        of w1, w2, u1, u2, (nmembers, outer_size) (or (outer_size,)) FP64
        device tensor, or None for zeros.  Returns one info dict per member."""
        k = 1 if u1.dim() == 1 else int(u1.shape[0])
        self._check_bnd("Baroclin.calc", k, bnd)
        infos = (BcInfo * k)()
        rc = lib().phg_baroclin_calc_bnd(self.h, k, _ptr(u1), _ptr(u2),
                                         _ptr(f) if f is not None else None,
                                         _ptr(g) if g is not None else None,
                                         _ptr(bnd) if bnd is not None else None, _ptr(u11),
                                         _ptr(u21), C.byref(opts) if opts is not None else None,
                                         infos, _stream(stream))
        check(rc, "phg_baroclin_calc")
        return [dict(iterations=i.iterations, distance=i.distance, converged=bool(i.converged),
                     it_mass=i.it_mass, it_jacobian=i.it_jacobian, it_system=i.it_system)
                for i in infos]

    def _check_bnd(self, what, k, bnd):
        if bnd is not None and (bnd.numel() != k * self.m.outer_size or
                                str(bnd.dtype) != "torch.float64" or not bnd.is_contiguous()):
            raise PhelmError("%s: bnd must be contiguous FP64 with outer_size values per "
                             "member" % what)

    def _lin(self, name, u11, u21, u1, u2, z1, z2, opts, stream, bnd):
        k = 1 if u1.dim() == 1 else int(u1.shape[0])
        for t in (u2, z1, z2, u11, u21):
            if tuple(t.shape) != tuple(u1.shape):
                raise PhelmError("Baroclin.%s: arguments must have one shape" % name)
        self._check_bnd("Baroclin." + name, k, bnd)
        infos = (BcInfo * k)()
        rc = getattr(lib(), "phg_baroclin_" + name + "_bnd")(
            self.h, k, _ptr(u1), _ptr(u2), _ptr(z1), _ptr(z2),
            _ptr(bnd) if bnd is not None else None, _ptr(u11), _ptr(u21),
            C.byref(opts) if opts is not None else None, infos, _stream(stream))
        check(rc, "phg_baroclin_" + name)
        return [dict(iterations=i.iterations, distance=i.distance, converged=bool(i.converged),
                     it_mass=i.it_mass, it_jacobian=i.it_jacobian, it_system=i.it_system)
                for i in infos]

    def calc_L(self, u11, u21, u1, u2, z1, z2, opts=None, stream=None, bnd=None):
        """Baroclin::calc_L (examples/baroclin.cpp:607-802): the step
        linearised about (z1, z2), one pass with A_; bnd as in calc (the
        boundary values of every Laplacian, of the right part and of the
        results), or None for zeros."""
        return self._lin("calc_L", u11, u21, u1, u2, z1, z2, opts, stream, bnd)

    def calc_L_1(self, u11, u21, u1, u2, z1, z2, opts=None, stream=None, bnd=None):
        """Baroclin::calc_L_1 (examples/baroclin.cpp:804-986): up to 10 passes
        with Ab_ until the distance is below 1e-8; bnd as in calc_L (the
        right part's boundary branch from Ab_'s elements)."""
        return self._lin("calc_L_1", u11, u21, u1, u2, z1, z2, opts, stream, bnd)

    def L_step(self, u11, u21, u1, u2, z1, z2, opts=None, stream=None):
        return self.calc_L(u11, u21, u1, u2, z1, z2, opts, stream)

    def L_1_step(self, u11, u21, u1, u2, z1, z2, opts=None, stream=None):
        return self.calc_L_1(u11, u21, u1, u2, z1, z2, opts, stream)

    def LT_step(self, u11, u21, u1, u2, z1, z2, opts=None, stream=None):
        """The reference's calc_LT has an empty body (examples/baroclin.cpp:
        988-993): the outputs stay as they are."""
        return []

    def close(self):
        if self.h:
            lib().phg_baroclin_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServeConfig(C.Structure):
    _fields_ = [("n_levels", _I), ("levels", _I * 8), ("slots", _I * 8),
                ("hbm_budget", C.c_size_t), ("coriolis", _D), ("conf", BvConf),
                ("opts", BvOpts), ("verbose", _I), ("substeps", _I * 8)]


class ServeStats(C.Structure):
    _fields_ = [("completed", _I), ("passes", _L), ("wall_s", _D), ("throughput", _D),
                ("p50_latency_s", _D), ("p99_latency_s", _D), ("mean_latency_s", _D),
                ("p50_service_s", _D), ("mean_occupancy", _D), ("device_bytes", C.c_size_t)]


class Server:
    """Continuous-batching server of BarVortex implicit steps (phg_serve_*):
    one engine per mesh level, slots sized by the HBM budget, member states
    resident on the device."""

    def __init__(self, levels=(5, 6, 7, 8), slots=None, hbm_budget=None, precision=F64,
                 verbose=False, substeps=None, **conf):
        c = ServeConfig()
        lib().phg_serve_default_config(C.byref(c))
        c.n_levels = len(levels)
        for i, lv in enumerate(levels):
            c.levels[i] = lv
            c.slots[i] = 0 if slots is None else slots[i]
            c.substeps[i] = 1 if substeps is None else substeps[i]
        if hbm_budget is not None:
            c.hbm_budget = int(hbm_budget)
        for k, v in conf.items():
            setattr(c.conf, k, v)
        c.opts.precision = precision
        c.verbose = int(verbose)
        self.cfg = c
        self.h = C.c_void_p()
        check(lib().phg_serve_create(C.byref(c), C.byref(self.h)), "phg_serve_create")
        self.levels = list(levels)

    def slots(self, level):
        return lib().phg_serve_slots(self.h, level)

    def add_member(self, level, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        mid = C.c_int()
        check(lib().phg_serve_add_member(self.h, level, _pd(u), C.byref(mid)),
              "phg_serve_add_member")
        return mid.value

    def member_state(self, member, n_points):
        out = np.zeros(n_points)
        check(lib().phg_serve_member_state(self.h, member, _pd(out)), "phg_serve_member_state")
        return out

    def run(self, members, arrivals):
        n = len(members)
        m = np.ascontiguousarray(members, dtype=np.int32)
        a = np.ascontiguousarray(arrivals, dtype=np.float64)
        start, finish = np.zeros(n), np.zeros(n)
        infos = (BvInfo * max(n, 1))()
        st = ServeStats()
        check(lib().phg_serve_run(self.h, n, _pi(m), _pd(a), _pd(start), _pd(finish), infos,
                                  C.byref(st)), "phg_serve_run")
        stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
        return start, finish, [_bv_info(infos[r]) for r in range(n)], stats

    def close(self):
        if self.h:
            lib().phg_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SphereJacobian:
    """SphereJacobian (src/sjacobian.h:65-149): the nine cubature matrices
    assembled on the GPU; calc2 evaluates up to 8 (u, v) pairs with one
    batched mass solve of the 8 distinct right-hand sides per pair."""

    MATRICES = ("idt_", "diff1_", "diff2_", "diff1_cos_", "diff2_cos_",
                "diff1_rp_", "diff2_rp_", "diff1_cos_rp_", "diff2_cos_rp_")

    def __init__(self, mesh, stream=None, restart=8, rtol=1e-13, ortho=CGS1,
                 precision=F64, diag_scale=True):
        self.m = mesh
        self.h = C.c_void_p()
        check(lib().phg_jacobian_create(mesh.h, _stream(stream), C.byref(self.h)),
              "phg_jacobian_create")
        self.opts = _opts(restart, rtol, 200000, ortho, precision)
        self.opts.diag_scale = int(diag_scale)   # the library's default mass-solve options
        for which, name in enumerate(self.MATRICES):
            mh = C.c_void_p()
            check(lib().phg_jacobian_matrix(self.h, which, C.byref(mh)))
            setattr(self, name, Matrix(mh, owned=False))

    def _run(self, fn, ans, u, v, stream):
        npairs = 1 if u.dim() == 1 else int(u.shape[0])
        infos = (GmresInfo * (8 * npairs))()
        rc = fn(self.h, npairs, _ptr(u), _ptr(v), _ptr(ans), C.byref(self.opts), infos,
                _stream(stream))
        check(rc, "phg_jacobian_calc2")
        return [dict(iterations=i.iterations, rel_residual=i.rel_residual,
                     converged=bool(i.converged)) for i in infos]

    def calc2(self, ans, u, v, stream=None):
        """ans[k] = J(u[k], v[k]) at the inner points; u, v: (npairs, n_points)
        or (n_points,) FP64 device tensors."""
        return self._run(lib().phg_jacobian_calc2, ans, u, v, stream)

    def calc2t(self, ans, u, v, stream=None):
        return self._run(lib().phg_jacobian_calc2t, ans, u, v, stream)

    def calc1(self, ans, u, v, bnd, p1, stream=None):
        """calc2 then p2u with the boundary values bnd (src/sjacobian.cpp:448-453);
        p1 is an inner-sized work tensor."""
        infos = self.calc2(p1, u, v, stream)
        self.m.p2u(p1, bnd, ans, stream)
        return infos

    def close(self):
        if self.h:
            lib().phg_jacobian_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def helmholtz_coeffs(tau=1.0, sigma=0.0, theta=1.0, mu=1e-2):
    """c_mass, c_lap of bv_private::integrate_cb (examples/barvortex_impl.h:57-80):
    (1/tau + sigma*theta) * M - theta*mu * L."""
    return 1.0 / tau + sigma * theta, -theta * mu


class SphereLaplace:
    """SphereLaplace<T, Matrix> (src/slaplace.h:60-108): idt_ (mass),
    laplace_ (stiffness) and the boundary matrices, assembled on the GPU, with
    calc1/calc2/solve (src/impl/slaplace_impl.h:84-192).  Vectors are FP64
    device tensors: full-mesh (V) for F, bnd of outer size."""

    def __init__(self, mesh, stream=None):
        self.m = mesh
        self.idt_ = Matrix.assemble(mesh, MASS, stream=stream)
        self.laplace_ = Matrix.assemble(mesh, LAPLACE, stream=stream)
        self.bnd1_ = Matrix.assemble_boundary(mesh, MASS)       # laplace_bnd1_cb
        self.bnd2_ = Matrix.assemble_boundary(mesh, HELMHOLTZ, 0.0, -1.0)  # laplace_bnd2_cb
        self.bnd3_ = Matrix.assemble_boundary(mesh, LAPLACE)    # slaplace_integrate_cb
        self._mg = None
        self.stream = stream

    def _vec(self, n):
        import torch
        return torch.zeros(n, dtype=torch.float64, device="cuda")

    def calc2(self, ans, F, rtol=1e-13):
        """ans (inner) = idt^-1 (laplace_ u2p(F) + bnd3_ proj_bnd(F))."""
        s = _stream(self.stream)
        rs, os_ = self.m.inner_size, self.m.outer_size
        inn, out, tmp = self._vec(rs), self._vec(rs), self._vec(max(os_, 1))
        check(lib().phg_u2p(self.m.h, _ptr(F), _ptr(inn), s), "u2p")
        check(lib().phg_proj_bnd(self.m.h, _ptr(F), _ptr(tmp), s), "proj_bnd")
        self.laplace_.mult_vector(out, inn, self.stream)
        self.bnd3_.mult_vector(inn, tmp, self.stream)
        check(lib().phg_axpby(rs, 1.0, _ptr(out), 1.0, _ptr(inn), _ptr(out), s), "vec_sum")
        return self.idt_.solve(ans, out, ortho=CGS2, rtol=rtol, stream=self.stream)

    def calc1(self, ans, F, bnd=None, rtol=1e-13):
        """calc2, then p2u with the boundary values bnd (None: zero)."""
        p = self._vec(self.m.inner_size)
        info = self.calc2(p, F, rtol)
        check(lib().phg_p2u(self.m.h, _ptr(p), _ptr(bnd) if bnd is not None else None,
                            _ptr(ans), _stream(self.stream)), "p2u")
        return info

    def solve(self, ans, F, bnd=None, rtol=1e-10, max_iters=2000):
        """laplace_ x = idt_ u2p(F) (+ bnd2_ bnd), ans = p2u(x, bnd): the
        Poisson solve, multigrid-preconditioned GMRES."""
        s = _stream(self.stream)
        rs = self.m.inner_size
        inn, b = self._vec(rs), self._vec(rs)
        check(lib().phg_u2p(self.m.h, _ptr(F), _ptr(inn), s), "u2p")
        self.idt_.mult_vector(b, inn, self.stream)
        if bnd is not None:
            self.bnd2_.mult_vector(inn, bnd, self.stream)
            check(lib().phg_axpby(rs, 1.0, _ptr(b), 1.0, _ptr(inn), _ptr(b), s), "vec_sum")
        if self._mg is None:
            self._mg = Multigrid(self.m, LAPLACE)
        x = self._vec(rs)
        info = self._mg.solve_batch(x.reshape(-1, 1), b.reshape(-1, 1), rtol=rtol,
                                    max_iters=max_iters)[0]
        check(lib().phg_p2u(self.m.h, _ptr(x), _ptr(bnd) if bnd is not None else None,
                            _ptr(ans), s), "p2u")
        return info


class SphereNorm:
    """SphereNorm<T> (src/norm.h:58-120): NORM_ = generate_full_matrix with
    sphere_scalar_cb; scalar(u, v) = u . NORM_ v (fast_scalar,
    src/mesh.h:550-591), norm = sqrt(scalar(u, u)), dist = norm(u - v)."""

    def __init__(self, mesh, stream=None):
        self.m = mesh
        self.NORM_ = Matrix.assemble_full(mesh, MASS, stream=stream)
        self.stream = stream

    def scalar(self, u, v):
        import torch
        tmp = torch.zeros(self.m.n_points, dtype=torch.float64, device="cuda")
        self.NORM_.mult_vector(tmp, v, self.stream)
        r = C.c_double()
        check(lib().phg_dot(self.m.n_points, _ptr(u), _ptr(tmp), C.byref(r),
                            _stream(self.stream)), "vec_scalar2")
        return r.value

    def norm(self, u):
        return math.sqrt(self.scalar(u, u))

    def dist(self, u, v):
        import torch
        d = torch.zeros(self.m.n_points, dtype=torch.float64, device="cuda")
        check(lib().phg_axpby(self.m.n_points, 1.0, _ptr(u), -1.0, _ptr(v), _ptr(d),
                              _stream(self.stream)), "vec_diff")
        return self.norm(d)


def schafe_f(x, y, t, mu, sigma):
    """SphereChafe_Private::f (examples/schafe.cpp:55-63): the forcing for
    which x sin(y + t) cos^4 x solves du/dt = mu Delta u - sigma u + f
    (examples/test_schafe.cpp:25-34); ipow(c, 2) = (1 c) c."""
    x = np.asarray(x, dtype=np.float64)
    c = np.cos(x)
    c2 = (1.0 * c) * c
    s = np.sin(y + t)
    return c2 * (x * np.cos(y + t) * c2 + sigma * s * c2 * x + 9 * mu * s * np.sin(x) * c
                 - 15 * mu * s * x + 20 * mu * s * x * c2)


class SphereChafe:
    """SphereChafe<T> (examples/schafe.h:52-97): du/dt = mu Delta u - sigma u
    + f, one step per solve() (examples/schafe_impl.h:126-206).  A_ =
    (1/tau + sigma/2) M - (mu/2) L and its boundary matrix bnd_
    (schafe_integrate_cb, examples/schafe.cpp:67-89) are GPU-assembled; the
    right part u/tau + (mu/2) calc2(u) - (sigma/2) u + f(x, y, t) (f at the
    inner points' zone-0 chart coordinates, on the host as the reference
    evaluates it), times idt_, plus bnd_ bnd; A_ x = right part by GMRES;
    ans = p2u(x, bnd).  Vectors are FP64 device tensors."""

    def __init__(self, mesh, tau, sigma, mu, f=schafe_f, stream=None):
        self.m = mesh
        self.tau_, self.sigma_, self.mu_ = tau, sigma, mu
        self.f = f
        self.laplace_ = SphereLaplace(mesh, stream=stream)
        cm, cl = 1.0 / tau + sigma * 0.5, -0.5 * mu
        self.A_ = Matrix.assemble(mesh, HELMHOLTZ, cm, cl, stream=stream)
        self.bnd_ = Matrix.assemble_boundary(mesh, HELMHOLTZ, cm, cl)
        inner = mesh.arrays()["inner"]
        xy = mesh.zone0_xy()
        self._x, self._y = xy[inner, 0], xy[inner, 1]
        self.stream = stream

    def solve(self, ans, X0, bnd=None, t=0.0, restart=30, rtol=1e-10, max_iters=20000,
              ortho=MGS):
        import torch
        s = _stream(self.stream)
        rs = self.m.inner_size
        vec = lambda: torch.zeros(rs, dtype=torch.float64, device="cuda")
        u, du, crp, x = vec(), vec(), vec(), vec()
        L = lib()
        check(L.phg_u2p(self.m.h, _ptr(X0), _ptr(u), s), "u2p")
        self.laplace_.calc2(du, X0)
        check(L.phg_axpby(rs, 1.0 / self.tau_, _ptr(u), self.mu_ * 0.5, _ptr(du), _ptr(du), s),
              "vec_sum1")
        check(L.phg_axpby(rs, 1.0, _ptr(du), -self.sigma_ * 0.5, _ptr(u), _ptr(du), s),
              "vec_sum1")
        rp = torch.as_tensor(np.ascontiguousarray(
            self.f(self._x, self._y, t, self.mu_, self.sigma_) if self.f is not None
            else np.zeros(rs)), dtype=torch.float64, device="cuda")
        check(L.phg_axpby(rs, 1.0, _ptr(du), 1.0, _ptr(rp), _ptr(u), s), "vec_sum")
        self.laplace_.idt_.mult_vector(crp, u, self.stream)
        if bnd is not None:
            tmp = vec()
            self.bnd_.mult_vector(tmp, bnd, self.stream)
            check(L.phg_axpby(rs, 1.0, _ptr(crp), 1.0, _ptr(tmp), _ptr(crp), s), "vec_sum")
        info = self.A_.solve(x, crp, restart=restart, rtol=rtol, max_iters=max_iters,
                             ortho=ortho, stream=self.stream)
        check(L.phg_p2u(self.m.h, _ptr(x), _ptr(bnd) if bnd is not None else None, _ptr(ans),
                        s), "p2u")
        return info


class Helmholtz:
    """Implicit-step operator A = c_mass*M + c_lap*L of BarVortex
    (examples/barvortex_impl.h:57-80, :149-153) and SphereChafe
    (examples/schafe.cpp:67-89), assembled per triangle on the GPU."""

    def __init__(self, mesh, tau=1.0, sigma=0.0, theta=1.0, mu=1e-2, stream=None):
        self.m = mesh
        self.c_mass, self.c_lap = helmholtz_coeffs(tau, sigma, theta, mu)
        self.A_ = Matrix.assemble(mesh, HELMHOLTZ, self.c_mass, self.c_lap, stream=stream)
        self.bnd_ = Matrix.assemble_boundary(mesh, HELMHOLTZ, self.c_mass, self.c_lap)
