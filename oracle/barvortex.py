"""CPU restatement of BarVortex::calc (examples/barvortex_impl.h:171-310),
SphereLaplace::calc1/solve (src/impl/slaplace_impl.h:84-192) and SphereNorm
(src/norm.h:58-120).

TEST INFRASTRUCTURE ONLY: the checker for the GPU BarVortex step; the product
never imports it.

Matrices are the oracle's assemblies of the reference callbacks.  Every
solve of the reference goes through its external solver library (GMRES,
parity unpinned); here each is a sparse LU factorisation, i.e. the exact
answer any converged solve approximates.  Vector operations follow the
reference's order; the linal helpers it calls are read from their call
sites: vec_sum1(r, a, b, k1, k2) = k1 a + k2 b, vec_sum2(r, a, b, k) =
a + k b, vec_sum = a + b, vec_diff = a - b.  Boundary (pole) values are zero,
as the reference drivers pass them (examples/test_barvortex.cpp:445).
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O
from .sjacobian import SphereJacobian


def _csr(t, n_cols):
    rp, ci, v = t
    return sp.csr_matrix((v, ci, rp), shape=(len(rp) - 1, n_cols))


class BarVortex:
    def __init__(self, mesh, tau, sigma, mu, k1, k2, theta, lh):
        self.m = mesh
        self.tau, self.sigma, self.mu = tau, sigma, mu
        self.k1, self.k2, self.theta = k1, k2, theta
        self.lh = np.asarray(lh, dtype=np.float64)
        n, no = mesh.n_inner, len(mesh.outer)
        self.idt = mesh.assemble(O.KIND_MASS)
        self.lap = mesh.assemble(O.KIND_LAPLACE)
        self.bnd3 = mesh.assemble_boundary(O.KIND_LAPLACE)
        cm, cl = 1.0 / tau + sigma * theta, -theta * mu      # integrate_cb
        self.A = mesh.assemble(O.KIND_HELMHOLTZ, cm, cl)
        self.norm_m = mesh.assemble_full(O.KIND_MASS)
        self.J = SphereJacobian(mesh)
        self.lu_idt = spla.splu(_csr(self.idt, n).tocsc())
        self.lu_A = spla.splu(_csr(self.A, n).tocsc())
        self.lu_lap = spla.splu(_csr(self.lap, n).tocsc())
        self.V = len(mesh.p2io)

    def p2u(self, p):
        u = np.zeros(self.V)
        u[self.m.inner] = p
        return u

    def calc1(self, u):
        """SphereLaplace::calc1 with bnd = 0: p2u(idt^-1 (laplace u_in + bnd3 u_bnd))."""
        u_in, u_bnd = u[self.m.inner], u[self.m.outer]
        out = O.spmv(*self.lap, u_in)
        brp, bci, bv = self.bnd3
        tmp = O.spmv(brp, bci, bv, np.ascontiguousarray(u_bnd)) if len(bci) else 0.0 * out
        out = out + tmp
        return self.p2u(self.lu_idt.solve(out))

    def poisson(self, w):
        """SphereLaplace::solve with bnd = 0: p2u(laplace^-1 (idt u2p(w)))."""
        b = O.spmv(*self.idt, w[self.m.inner])
        return self.p2u(self.lu_lap.solve(b))

    def norm(self, u):
        return math.sqrt(float(np.dot(u, O.spmv(*self.norm_m, u))))

    def calc(self, u, max_iters=1000, tol=1e-14):
        u = np.asarray(u, dtype=np.float64)
        th, k1, k2 = self.theta, self.k1, self.k2
        nr0 = self.norm(u)
        w = self.calc1(u)
        dw = self.calc1(w)
        FC = (1.0 / self.tau) * w + (self.mu * (1.0 - th)) * dw
        FC = 1.0 * FC + (-self.sigma * (1.0 - th)) * w
        u_n = np.zeros(self.V)
        w_n = np.zeros(self.V)
        it = 0
        rel = float("nan")
        for it in range(1, max_iters + 1):
            tmp1 = (k1 * th) * w_n + (k1 * (1.0 - th)) * w
            tmp1 = tmp1 + k2 * self.lh
            tmp2 = th * u_n + (1.0 - th) * u
            jac = self.J.calc2(tmp2, tmp1)
            F = FC[self.m.inner] - jac
            rp = O.spmv(*self.idt, F)
            w_n = self.p2u(self.lu_A.solve(rp))
            u_n1 = self.poisson(w_n)
            nr = self.norm(u_n1 - u_n)
            u_n = u_n1
            rel = nr / nr0 if nr0 > 0 else nr
            if rel < tol or math.isnan(nr):
                break
        return u_n, it, rel
