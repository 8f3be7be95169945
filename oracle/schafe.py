"""CPU restatement of SphereChafe (examples/schafe.h:52-97,
examples/schafe_impl.h:102-206, examples/schafe.cpp:55-89) and of the
reference driver's analytic case (examples/test_schafe.cpp:25-34, 61-125).

TEST INFRASTRUCTURE ONLY: the checker for the GPU SphereChafe; the product
never imports it.

A_ = schafe_integrate_cb: integrate_cos(phi_i phi_j) (1/tau + sigma/2) +
slaplace(phi_j, phi_i) (-mu/2), as the oracle's Helmholtz kind; bnd_ the same
callback on the boundary columns.  solve() follows schafe_impl.h:126-206 in
its vector order; SphereLaplace::calc2 and A_.solve are sparse LU solves
(the answer any converged GMRES approximates).
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O


def schafe_f(x, y, t, mu, sigma):
    """SphereChafe_Private::f (examples/schafe.cpp:55-63); ipow(c, 2) =
    (1 c) c, cos/sin of the C library (the reference's right part is
    evaluated on the host)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    c = np.cos(x)
    c2 = (1.0 * c) * c
    s = np.sin(y + t)
    return c2 * (x * np.cos(y + t) * c2 + sigma * s * c2 * x + 9 * mu * s * np.sin(x) * c
                 - 15 * mu * s * x + 20 * mu * s * x * c2)


def schafe_ans(x, y, t):
    """ans (examples/test_schafe.cpp:25-29): x sin(y + t) cos^4 x, ipow as
    repeated multiplication."""
    c = np.cos(x)
    c4 = (((1.0 * c) * c) * c) * c
    return x * np.sin(y + t) * c4


def _csr(t, n_cols):
    rp, ci, v = t
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, n_cols))


class SphereChafe:
    def __init__(self, mesh, xy, tau, sigma, mu):
        self.m = mesh
        self.tau, self.sigma, self.mu = tau, sigma, mu
        self.xy = np.asarray(xy, dtype=np.float64)   # zone-0 chart coordinates, every point
        n = mesh.n_inner
        cm, cl = 1.0 / tau + sigma * 0.5, -0.5 * mu  # schafe_integrate_cb
        self.A = mesh.assemble(O.KIND_HELMHOLTZ, cm, cl)
        self.bndA = mesh.assemble_boundary(O.KIND_HELMHOLTZ, cm, cl)
        self.idt = mesh.assemble(O.KIND_MASS)
        self.lap = mesh.assemble(O.KIND_LAPLACE)
        self.bnd3 = mesh.assemble_boundary(O.KIND_LAPLACE)
        self.lu_idt = spla.splu(_csr(self.idt, n).tocsc())
        self.lu_A = spla.splu(_csr(self.A, n).tocsc())
        self.V = len(mesh.p2io)

    def calc2(self, u):
        """SphereLaplace::calc2 (src/impl/slaplace_impl.h:153-178)."""
        out = O.spmv(*self.lap, u[self.m.inner])
        if len(self.bnd3[1]):
            out = out + O.spmv(*self.bnd3, np.ascontiguousarray(u[self.m.outer]))
        return self.lu_idt.solve(out)

    def solve(self, X0, bnd, t):
        """SphereChafe::solve (examples/schafe_impl.h:126-206): one step from
        X0 (every point) with the new boundary values bnd (outer order, or
        None)."""
        X0 = np.asarray(X0, dtype=np.float64)
        u = X0[self.m.inner]
        du = self.calc2(X0)
        du = (1.0 / self.tau) * u + (self.mu * 0.5) * du
        du = 1.0 * du + (-self.sigma * 0.5) * u
        inner = self.m.inner
        rp = schafe_f(self.xy[inner, 0], self.xy[inner, 1], t, self.mu, self.sigma)
        u = du + rp
        crp = O.spmv(*self.idt, u)
        if bnd is not None:
            crp = crp + O.spmv(*self.bndA, np.ascontiguousarray(bnd, dtype=np.float64))
        x = self.lu_A.solve(crp)
        out = np.zeros(self.V)
        out[inner] = x
        if bnd is not None:
            out[self.m.outer] = bnd
        return out


def norm_dist(mesh, u, v):
    """SphereNorm::dist (src/norm.h:58-120)."""
    d = np.asarray(u) - np.asarray(v)
    nm = mesh.assemble_full(O.KIND_MASS)
    return math.sqrt(float(np.dot(d, O.spmv(*nm, d))))
