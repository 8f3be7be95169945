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
a + k b, vec_sum = a + b, vec_diff = a - b.  Boundary (pole) values bnd_u
(of psi) and bnd_w (of its Laplacian) are optional everywhere (None: zero, as
the reference drivers pass them, examples/test_barvortex.cpp:445).
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
        # Ab_: integrate_backward_cb (barvortex_impl.h:83-103)
        self.Ab = mesh.assemble(O.KIND_HELMHOLTZ, 1.0 / tau - sigma * (1.0 - theta),
                                (1.0 - theta) * mu)
        self.bndA = mesh.assemble_boundary(O.KIND_HELMHOLTZ, cm, cl)    # bnd_ (integrate_cb)
        # bndb_: integrate_backward_cb on the boundary columns (barvortex_impl.h:141-153)
        self.bndAb = mesh.assemble_boundary(O.KIND_HELMHOLTZ, 1.0 / tau - sigma * (1.0 - theta),
                                            (1.0 - theta) * mu)
        self.bnd2 = mesh.assemble_boundary(O.KIND_HELMHOLTZ, 0.0, -1.0)  # laplace_bnd2_cb
        # barvortex_impl.h:287-293 ADDS bnd_ bnd_w to the right part, where the
        # Dirichlet elimination of the same operator subtracts it; +1 follows
        # the reference, -1 is the consistent elimination (used to pin the rest).
        self.bnd_sign = 1.0
        self.norm_m = mesh.assemble_full(O.KIND_MASS)
        self.J = SphereJacobian(mesh)
        self.lu_idt = spla.splu(_csr(self.idt, n).tocsc())
        self.lu_A = spla.splu(_csr(self.A, n).tocsc())
        self.lu_Ab = spla.splu(_csr(self.Ab, n).tocsc())
        self.lu_lap = spla.splu(_csr(self.lap, n).tocsc())
        self.V = len(mesh.p2io)

    def p2u(self, p, bnd=None):
        """p2u (src/proj.cpp:49-75): inner values, boundary values from bnd
        (indexed by outer number) or 0."""
        u = np.zeros(self.V)
        u[self.m.inner] = p
        if bnd is not None:
            u[self.m.outer] = np.asarray(bnd, dtype=np.float64)
        return u

    def _bmul(self, t, bnd):
        brp, bci, bv = t
        return O.spmv(brp, bci, bv, np.ascontiguousarray(bnd, dtype=np.float64))

    def calc1(self, u, bnd=None):
        """SphereLaplace::calc1 (src/impl/slaplace_impl.h:153-192):
        p2u(idt^-1 (laplace u_in + bnd3 u_bnd), bnd)."""
        u_in, u_bnd = u[self.m.inner], u[self.m.outer]
        out = O.spmv(*self.lap, u_in)
        tmp = self._bmul(self.bnd3, u_bnd) if len(self.bnd3[1]) else 0.0 * out
        out = out + tmp
        return self.p2u(self.lu_idt.solve(out), bnd)

    def poisson(self, w, bnd=None):
        """SphereLaplace::solve (src/impl/slaplace_impl.h:84-130):
        p2u(laplace^-1 (idt u2p(w) + bnd2 bnd), bnd), bnd2 = -(the stiffness
        boundary block) (laplace_bnd2_cb, src/slaplace.cpp:159-169)."""
        b = O.spmv(*self.idt, w[self.m.inner])
        if bnd is not None:
            b = b + self._bmul(self.bnd2, bnd)
        return self.p2u(self.lu_lap.solve(b), bnd)

    def norm(self, u):
        return math.sqrt(float(np.dot(u, O.spmv(*self.norm_m, u))))

    def calc(self, u, max_iters=1000, tol=1e-14, forcing=None, bnd_u=None, bnd_w=None):
        """One step; forcing: the right-part callback conf.rp(x, y, t, mu, sigma)
        evaluated at every mesh point for this step's t (or None), added to
        FC as barvortex_impl.h:233-244 does; bnd_u, bnd_w: the boundary (pole)
        values of psi and of its Laplacian at the new time (or None: 0)."""
        u = np.asarray(u, dtype=np.float64)
        th, k1, k2 = self.theta, self.k1, self.k2
        nr0 = self.norm(u)
        w = self.calc1(u, bnd_w)
        dw = self.calc1(w)
        FC = (1.0 / self.tau) * w + (self.mu * (1.0 - th)) * dw
        FC = 1.0 * FC + (-self.sigma * (1.0 - th)) * w
        if forcing is not None:
            FC = FC + np.asarray(forcing, dtype=np.float64)
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
            if bnd_w is not None:
                rp = rp + self.bnd_sign * self._bmul(self.bndA, bnd_w)
            w_n = self.p2u(self.lu_A.solve(rp), bnd_w)
            u_n1 = self.poisson(w_n, bnd_u)
            nr = self.norm(u_n1 - u_n)
            u_n = u_n1
            rel = nr / nr0 if nr0 > 0 else nr
            if rel < tol or math.isnan(nr):
                break
        return u_n, it, rel


    # ---- linearised and adjoint steps (barvortex_impl.h:326-571) ----
    def jac1(self, u, v):
        """SphereJacobian::calc1 with bnd = 0 (src/sjacobian.cpp:456-461)."""
        return self.p2u(self.J.calc2(u, v))

    def jac1t(self, u, v):
        """SphereJacobian::calc1t with bnd = 0 (src/sjacobian.cpp:463-468)."""
        return self.p2u(self.J.calc2t(u, v))

    def _solve_full(self, lu, u1, bndm=None, bnd_w=None):
        """The reference's right part idt u2p(u1), plus bndm bnd_w when
        bnd_w is given (barvortex_impl.h:389-397), then phelm::solve
        (mesh.h:719-746): p2u(A^-1 rp, bnd_w)."""
        rp = O.spmv(*self.idt, u1[self.m.inner])
        if bnd_w is not None:
            rp = rp + self._bmul(bndm, bnd_w)
        return self.p2u(lu.solve(rp), bnd_w)

    def _lin_terms(self, u, z, bnd_w=None):
        z_lapl = self.calc1(z, bnd_w)
        pt1 = self.calc1(u, bnd_w)
        jac1 = self.jac1(u, self.lh)
        jac2 = self.jac1(z, pt1)
        jac3 = self.jac1(u, z_lapl)
        pt3 = np.zeros(self.V)
        pt3 = pt3 + self.k2 * jac1                          # vec_sum2
        pt3 = pt3 + self.k1 * jac2
        pt3 = pt3 + self.k1 * jac3
        return pt1, pt3

    def calc_L(self, u, z, bnd_u=None, bnd_w=None):
        """BarVortex::calc_L (barvortex_impl.h:326-400)."""
        th, mu, sigma, tau = self.theta, self.mu, self.sigma, self.tau
        pt1, pt3 = self._lin_terms(np.asarray(u, dtype=np.float64),
                                   np.asarray(z, dtype=np.float64), bnd_w)
        pt3 = pt3 * -1.0
        pt2 = self.calc1(pt1)
        pt2 = pt2 * ((1.0 - th) * mu)
        pt1 = pt1 * (1.0 / tau - (1.0 - th) * sigma)
        u1 = np.zeros(self.V)
        u1 = u1 + pt1
        u1 = u1 + pt2
        u1 = u1 + pt3
        u1 = self._solve_full(self.lu_A, u1, self.bndA, bnd_w)
        return self.poisson(u1, bnd_u)

    def calc_L_1(self, u1, u, z, bnd_u=None, bnd_w=None):
        """BarVortex::calc_L_1 (barvortex_impl.h:402-467): adds to u1 as the
        reference does (its caller L_1_step passes zeros)."""
        th, mu, sigma, tau = self.theta, self.mu, self.sigma, self.tau
        pt1, pt3 = self._lin_terms(np.asarray(u, dtype=np.float64),
                                   np.asarray(z, dtype=np.float64), bnd_w)
        pt2 = self.calc1(pt1)
        pt2 = pt2 * (-th * mu)
        pt1 = pt1 * (1.0 / tau + th * sigma)
        r = np.asarray(u1, dtype=np.float64)
        r = r + pt1
        r = r + pt2
        r = r + pt3
        r = self._solve_full(self.lu_Ab, r, self.bndAb, bnd_w)
        return self.poisson(r, bnd_u)

    def calc_LT(self, v, z, bnd_u=None, bnd_w=None):
        """BarVortex::calc_LT (barvortex_impl.h:469-535)."""
        th, mu, sigma, tau = self.theta, self.mu, self.sigma, self.tau
        k1, k2 = self.k1, self.k2
        v = np.asarray(v, dtype=np.float64)
        lz = self.calc1(np.asarray(z, dtype=np.float64), bnd_w)
        v1 = self.poisson(v, bnd_u)
        v1 = self._solve_full(self.lu_A, v1, self.bndA, bnd_w)
        pt1 = self.calc1(v1, bnd_w)
        pt1 = pt1 * (1.0 / tau - (1.0 - th) * sigma)
        pt2 = self.calc1(v1, bnd_w)
        pt2 = self.calc1(pt2)
        pt2 = pt2 * ((1.0 - th) * mu)
        tmp = self.jac1t(v1, self.lh)
        p_lapl = self.jac1t(z, v1)
        p_lapl = self.calc1(p_lapl)
        h1 = self.jac1t(v1, lz)
        h1 = k1 * h1 + k1 * p_lapl                          # vec_sum1
        h1 = h1 + k2 * tmp                                  # vec_sum2
        h1 = h1 * -1.0
        out = np.zeros(self.V)
        out = out + pt1
        out = out + pt2
        out = out + h1
        # set_bnd(v1, bnd_u) (src/proj.cpp:143-160)
        out[self.m.outer] = 0.0 if bnd_u is None else np.asarray(bnd_u, dtype=np.float64)
        return out


def an1(x, y, t):
    """The reference driver's analytic solution (examples/test_barvortex.cpp:216-219)."""
    return x * np.sin(y + t) * np.cos(x) ** 4


def rp1(x, y, t, mu, sigma):
    """The forcing that makes an1 a solution without Coriolis
    (examples/test_barvortex.cpp:226-242), term for term."""
    c, s = np.cos(x), np.sin(x)
    sy, cy = np.sin(y + t), np.cos(y + t)
    return (390 * mu * sy * x * c ** 2 + 147 * mu * sy * s * c - 400 * mu * sy * x * c ** 4
            - 360 * mu * sy * s * c ** 3 - 20 * sigma * sy * c ** 4 * x
            + 15 * sigma * sy * c ** 2 * x - 9 * sigma * sy * c ** 3 * s
            + 30 * cy * c ** 4 * sy * x * x * s + 9 * cy * c ** 6 * sy * s - 45 * mu * sy * x
            - 9 * cy * c ** 5 * sy * x - 20 * x * cy * c ** 4 - 9 * cy * c ** 3 * s
            + 15 * x * cy * c ** 2)
