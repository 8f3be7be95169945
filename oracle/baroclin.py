"""CPU restatement of Baroclin::Baroclin and Baroclin::calc
(examples/baroclin.cpp:69-181, 445-606): the two-layer baroclinic implicit
step with its 4n x 4n block system.

TEST INFRASTRUCTURE ONLY: the checker for the GPU Baroclin step; the product
never imports it.

The nine blocks of A_ are the oracle's literal restatement of integrate_cb
(oracle/or_fem.c: or_assemble_baroclin); the right part is generate_right_part
with right_part_cb's interior branch (or_right_part_mass) and its boundary
branch (baroclin.cpp:259-323): for a boundary column j0 every element a of
integrate_cb's four block columns adds -bnd[j0] a to its block row, one bnd
for w1, w2, u1 and u2 (BW1 = BW2 = BU1 = BU2 = bnd, :541-544); here per block
with the block's boundary matrix (block_kinds).  The linearised steps calc_L
and calc_L_1 (:607-986) use the same bnd for every Laplacian (z's included),
the right part (A_'s elements, or Ab_'s through right_part_backward_cb,
:341-436; block_kinds_backward) and the four p2u.  The drivers pass zero
boundary values (examples/test_baroclin.cpp:215).  A_.solve of the reference goes through its
external solver library (GMRES, parity unpinned); here it is a sparse LU
factorisation of the 4n x 4n matrix.  SphereLaplace::calc1, SphereJacobian and
SphereNorm are the oracle's (oracle/barvortex.py, oracle/sjacobian.py).
Vector helpers as read from their call sites: vec_sum1(r, a, b, k1, k2) =
k1 a + k2 b, vec_sum(r, a, b) = a + b.
"""
import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import oracle as O
from .barvortex import BarVortex


def block_kinds(tau, sigma, mu, sigma1, mu1, alpha, theta):
    """integrate_cb's nine element formulas (baroclin.cpp:142-178) as
    (kind, c_mass, c_lap), in the order of oracle.BAROCLIN_BLOCKS; each
    block is c_mass (mass element) + c_lap (laplace element)."""
    H, M = O.KIND_HELMHOLTZ, O.KIND_MASS
    a2 = alpha * alpha
    return [(H, 1.0 / tau + 0.5 * theta * sigma, -(theta * mu)),
            (H, (-0.5 * sigma) * theta, 0.0),
            (H, 0.5 * theta * sigma, 0.0),
            (H, 1.0 / tau + 0.5 * theta * sigma + a2 * theta * mu1, -(theta * mu)),
            (H, -a2 / tau - a2 * theta * sigma1, 0.0),
            (M, 1.0, 0.0),
            (H, 0.0, -1.0),
            (M, 1.0, 0.0),
            (H, 0.0, -1.0)]


def block_kinds_backward(tau, sigma, mu, sigma1, mu1, alpha, theta):
    """integrate_backward_cb's nine element formulas (baroclin.cpp:184-244)
    as (kind, c_mass, c_lap), in the order of oracle.BAROCLIN_BLOCKS."""
    H, M = O.KIND_HELMHOLTZ, O.KIND_MASS
    a2 = alpha * alpha
    return [(H, 1.0 / tau - 0.5 * (1.0 - theta) * sigma, (1.0 - theta) * mu),
            (H, (0.5 * sigma) * (1.0 - theta), 0.0),
            (H, -0.5 * (1.0 - theta) * sigma, 0.0),
            (H, 1.0 / tau - 0.5 * (1.0 - theta) * sigma - a2 * (1.0 - theta) * mu1,
             (1.0 - theta) * mu),
            (H, -a2 / tau + a2 * (1.0 - theta) * sigma1, 0.0),
            (M, 1.0, 0.0),
            (H, 0.0, -1.0),
            (M, 1.0, 0.0),
            (H, 0.0, -1.0)]


class Baroclin:
    def __init__(self, mesh, tau, sigma, mu, sigma1, mu1, alpha, lh):
        self.m = mesh
        self.tau, self.sigma, self.mu = tau, sigma, mu
        self.sigma1, self.mu1, self.alpha = sigma1, mu1, alpha
        self.theta = 0.5                                    # baroclin.cpp:80
        self.lh = np.asarray(lh, dtype=np.float64)
        n = mesh.n_inner
        self.n = n
        rp, ci, v9 = O.baroclin_blocks(mesh, tau, sigma, mu, sigma1, mu1, alpha, self.theta)
        self.blocks = (rp, ci, v9)
        brp, bci, bv = O.baroclin_matrix(rp, ci, v9)
        self.A = (brp, bci, bv)
        self.lu_A = spla.splu(sp.csr_matrix((bv, bci, brp), shape=(4 * n, 4 * n)).tocsc())
        self._lu_Ab = None
        self.kinds = block_kinds(tau, sigma, mu, sigma1, mu1, alpha, self.theta)
        self.bnd_blocks = [mesh.assemble_boundary(kd, cm, cl) for kd, cm, cl in self.kinds]
        self.kinds_b = block_kinds_backward(tau, sigma, mu, sigma1, mu1, alpha, self.theta)
        self._bnd_blocks_b = None
        # SphereLaplace::calc1, SphereJacobian and SphereNorm of the same mesh
        self.bv = BarVortex(mesh, tau=1.0, sigma=0.0, mu=0.0, k1=0.0, k2=0.0, theta=0.5,
                            lh=np.zeros(len(mesh.p2io)))
        self.V = len(mesh.p2io)

    def calc1(self, u, bnd=None):
        return self.bv.calc1(u, bnd)

    def bnd_rhs(self, bnd, backward=False):
        """right_part_cb's boundary branch: the 4n vector of -sum_c
        (block (r, c) boundary columns) bnd per block row r; with backward,
        right_part_backward_cb's (baroclin.cpp:341-436), on Ab_'s elements."""
        rs = self.n
        out = np.zeros(4 * rs)
        if backward and self._bnd_blocks_b is None:
            self._bnd_blocks_b = [self.m.assemble_boundary(kd, cm, cl)
                                  for kd, cm, cl in self.kinds_b]
        blocks = self._bnd_blocks_b if backward else self.bnd_blocks
        for (r, c), (brp, bci, bv) in zip(O.BAROCLIN_BLOCKS, blocks):
            if len(bci):
                out[r * rs:(r + 1) * rs] += O.spmv(brp, bci, -bv,
                                                   np.ascontiguousarray(bnd, dtype=np.float64))
        return out

    def jac(self, u, v):
        return self.bv.J.calc2(u, v)

    def dist(self, u, v):
        return self.bv.norm(u - v)

    def calc(self, u1, u2, f=None, g=None, max_iters=20, tol=1e-8, bnd=None):
        """One step (u1, u2) -> (u11, u21); f, g: the forcing callbacks'
        values at every mesh point for this step's t (or None); bnd: the
        boundary values of w1, w2, u1, u2 (one array, outer order) or None
        for 0.  Returns (u11, u21, iterations, last max distance)."""
        th, tau = self.theta, self.tau
        sigma, mu, sigma1, mu1, alpha = self.sigma, self.mu, self.sigma1, self.mu1, self.alpha
        a2 = alpha * alpha
        inner = self.m.inner
        rs = self.n
        u1 = np.asarray(u1, dtype=np.float64)
        u2 = np.asarray(u2, dtype=np.float64)
        w1 = self.calc1(u1, bnd)
        w2 = self.calc1(u2, bnd)
        dw1 = self.calc1(w1, bnd)
        dw2 = self.calc1(w2, bnd)
        brhs = self.bnd_rhs(bnd) if bnd is not None else None
        s1 = lambda a, b, k1, k2: k1 * a + k2 * b  # noqa: E731  vec_sum1
        FC = s1(w1, w2, -0.5 * (1.0 - th) * sigma, 0.5 * (1.0 - th) * sigma)
        FC = s1(FC, dw1, 1.0, mu * (1.0 - th))
        FC = s1(FC, w1, 1.0, 1.0 / tau)
        GC = s1(w1, w2, -0.5 * (1.0 - th) * sigma, -0.5 * (1.0 - th) * sigma)
        GC = s1(GC, dw2, 1.0, mu * (1.0 - th))
        GC = s1(GC, w2, 1.0, 1.0 / tau)
        GC = s1(GC, u2, 1.0, -alpha * alpha / tau)
        GC = s1(GC, w2, 1.0, -alpha * alpha * mu1 * (1 - th))
        GC = s1(GC, u2, 1.0, alpha * alpha * sigma1 * (1 - th))
        u1_n, u2_n, w1_n, w2_n = u1.copy(), u2.copy(), w1.copy(), w2.copy()
        if f is not None:
            FC = FC + np.asarray(f, dtype=np.float64)
        if g is not None:
            GC = GC + alpha * alpha * np.asarray(g, dtype=np.float64)
        nr = float("nan")
        it = 0
        for it in range(1, max_iters + 1):
            tmp1 = s1(u1, u1_n, 1.0 - th, th)
            tmp2 = s1(w1, w1_n, 1.0 - th, th) + self.lh
            jac1 = self.jac(tmp1, tmp2)
            tmp1 = s1(u2, u2_n, 1.0 - th, th)
            tmp2 = s1(w2, w2_n, 1.0 - th, th)
            jac2 = self.jac(tmp1, tmp2)
            F = s1(jac1, jac2, -1.0, -1.0)
            tmp1 = s1(u1, u1_n, 1.0 - th, th)
            tmp2 = s1(w2, w2_n, 1.0 - th, th)
            jac1 = self.jac(tmp1, tmp2)
            tmp1 = s1(u2, u2_n, 1.0 - th, th)
            tmp2 = s1(w1, w1_n, 1.0 - th, th) + self.lh
            jac2 = self.jac(tmp1, tmp2)
            tmp1 = s1(u1, u1_n, 1.0 - th, th)
            tmp2 = s1(u2, u2_n, 1.0 - th, th)
            jac3 = self.jac(tmp1, tmp2)
            G = s1(jac1, jac2, -1.0, -1.0)
            G = s1(G, jac3, 1.0, a2)
            F = F + FC[inner]
            G = G + GC[inner]
            rp = np.zeros(4 * rs)
            rp[:rs] = O.right_part_mass(self.m, F)
            rp[rs:2 * rs] = O.right_part_mass(self.m, G)
            if brhs is not None:
                rp = rp + brhs
            ans = self.lu_A.solve(rp)
            w1_n = self.bv.p2u(ans[:rs], bnd)
            w2_n = self.bv.p2u(ans[rs:2 * rs], bnd)
            u1_n1 = self.bv.p2u(ans[2 * rs:3 * rs], bnd)
            u2_n1 = self.bv.p2u(ans[3 * rs:], bnd)
            nr = max(self.dist(u1_n1, u1_n), self.dist(u2_n1, u2_n))
            u1_n, u2_n = u1_n1, u2_n1
            if nr < tol:
                break
        return u1_n, u2_n, it, nr


    # ---- linearised steps (examples/baroclin.cpp:607-986) ---------------------
    @property
    def lu_Ab(self):
        """Ab_ (integrate_backward_cb, baroclin.cpp:184-244), factorised."""
        if self._lu_Ab is None:
            n = self.n
            rp, ci, v9 = O.baroclin_blocks(self.m, self.tau, self.sigma, self.mu, self.sigma1,
                                           self.mu1, self.alpha, self.theta, backward=True)
            self.Ab_blocks = (rp, ci, v9)
            brp, bci, bv = O.baroclin_matrix(rp, ci, v9)
            self.Ab = (brp, bci, bv)
            self._lu_Ab = spla.splu(sp.csr_matrix((bv, bci, brp), shape=(4 * n, 4 * n)).tocsc())
        return self._lu_Ab

    def lin_fc(self, which, u1, u2, bnd=None):
        """The explicit parts FC, GC of calc_L (which 0, baroclin.cpp:684-700)
        or calc_L_1 (which 1, :868-884) and the Laplacians w1, w2 they use,
        with the boundary values bnd (or None for 0)."""
        th, tau = self.theta, self.tau
        sigma, mu, sigma1, mu1, alpha = self.sigma, self.mu, self.sigma1, self.mu1, self.alpha
        s1 = lambda a, b, k1, k2: k1 * a + k2 * b  # noqa: E731  vec_sum1
        w1, w2 = self.calc1(u1, bnd), self.calc1(u2, bnd)
        dw1, dw2 = self.calc1(w1, bnd), self.calc1(w2, bnd)
        if which == 0:
            FC = s1(w1, w2, -0.5 * (1.0 - th) * sigma, 0.5 * (1.0 - th) * sigma)
            FC = s1(FC, dw1, 1.0, mu * (1.0 - th))
            FC = s1(FC, w1, 1.0, 1.0 / tau)
            GC = s1(w1, w2, -0.5 * (1.0 - th) * sigma, -0.5 * (1.0 - th) * sigma)
            GC = s1(GC, dw2, 1.0, mu * (1.0 - th))
            GC = s1(GC, w2, 1.0, 1.0 / tau)
            GC = s1(GC, u2, 1.0, -alpha * alpha / tau)
            GC = s1(GC, w2, 1.0, -alpha * alpha * mu1 * (1 - th))
            GC = s1(GC, u2, 1.0, alpha * alpha * sigma1 * (1 - th))
        else:
            FC = s1(w1, w2, 0.5 * th * sigma, -0.5 * th * sigma)
            FC = s1(FC, dw1, 1.0, -mu * th)
            FC = s1(FC, w1, 1.0, 1.0 / tau)
            GC = s1(w1, w2, 0.5 * th * sigma, 0.5 * th * sigma)
            GC = s1(GC, dw2, 1.0, -mu * th)
            GC = s1(GC, w2, 1.0, 1.0 / tau)
            GC = s1(GC, u2, 1.0, -alpha * alpha / tau)
            GC = s1(GC, w2, 1.0, alpha * alpha * mu1 * th)
            GC = s1(GC, u2, 1.0, -alpha * alpha * sigma1 * th)
        return FC, GC, w1, w2

    def calc_lin(self, which, u1, u2, z1, z2, tol=1e-8, bnd=None, full=False):
        """Baroclin::calc_L (which 0, baroclin.cpp:620-802: one pass with A_)
        or calc_L_1 (which 1, :804-986: at most 10 passes with Ab_ until the
        distance is below tol).  The Jacobian arguments are the code's,
        including J(L z1, avg w2) and J(L z2, avg u1) in G where the comments
        write J(z1, .) and J(., L z2).  bnd: the one boundary array the code
        passes to every calc1 (u, w and z: :675-682, :859-866), to the right
        part's boundary branch (BW1 = BW2 = BU1 = BU2 = bnd, :708-711 with
        A_'s elements, :892-895 with Ab_'s) and to the four p2u (:783-786,
        :967-970); None for 0.  Returns (u11, u21, passes, distance), with
        full also the last pass's w1_n, w2_n."""
        th, a2 = self.theta, self.alpha * self.alpha
        inner, rs = self.m.inner, self.n
        f64 = lambda a: np.asarray(a, dtype=np.float64)  # noqa: E731
        u1, u2, z1, z2 = f64(u1), f64(u2), f64(z1), f64(z2)
        s1 = lambda a, b, k1, k2: k1 * a + k2 * b  # noqa: E731  vec_sum1
        FC, GC, w1, w2 = self.lin_fc(which, u1, u2, bnd)
        dz1, dz2 = self.calc1(z1, bnd), self.calc1(z2, bnd)
        if which == 0:
            ka, kb, lu, passes = 1.0 - th, th, self.lu_A, 1
        else:
            ka, kb, lu, passes = th, 1.0 - th, self.lu_Ab, 10
        brhs = self.bnd_rhs(bnd, backward=which == 1) if bnd is not None else None
        u1_n, u2_n, w1_n, w2_n = u1.copy(), u2.copy(), w1.copy(), w2.copy()
        dz1lh = dz1 + self.lh                                     # vec_sum
        J = self.jac
        nr, it = float("nan"), 0
        for it in range(1, passes + 1):
            mu1_, mu2_ = s1(u1, u1_n, ka, kb), s1(u2, u2_n, ka, kb)
            mw1_, mw2_ = s1(w1, w1_n, ka, kb), s1(w2, w2_n, ka, kb)
            if which == 0:
                F = s1(J(mu1_, dz1lh), J(z1, mw1_), -1.0, -1.0)
                F = s1(F, J(z2, mw2_), 1.0, -1.0)
                F = s1(F, J(mu2_, dz2), 1.0, -1.0)
                G = s1(J(dz1, mw2_), J(dz2, mu1_), -1.0, -1.0)
                G = s1(G, J(mu2_, dz1lh), 1.0, -1.0)
                G = s1(G, J(z2, mw1_), 1.0, -1.0)
                G = s1(G, J(z1, mu2_), 1.0, a2)
                G = s1(G, J(mu1_, z2), 1.0, a2)
            else:
                F = J(mu1_, dz1lh) + J(z1, mw1_)                   # vec_sum
                F = F + J(z2, mw2_)
                F = F + J(mu2_, dz2)
                G = J(dz1, mw2_) + J(dz2, mu1_)
                G = G + J(mu2_, dz1lh)
                G = G + J(z2, mw1_)
                G = s1(G, J(z1, mu2_), 1.0, -a2)
                G = s1(G, J(mu1_, z2), 1.0, -a2)
            F = F + FC[inner]
            G = G + GC[inner]
            rp = np.zeros(4 * rs)
            rp[:rs] = O.right_part_mass(self.m, F)
            rp[rs:2 * rs] = O.right_part_mass(self.m, G)
            if brhs is not None:
                rp = rp + brhs
            ans = lu.solve(rp)
            w1_n = self.bv.p2u(ans[:rs], bnd)
            w2_n = self.bv.p2u(ans[rs:2 * rs], bnd)
            u1_n1 = self.bv.p2u(ans[2 * rs:3 * rs], bnd)
            u2_n1 = self.bv.p2u(ans[3 * rs:], bnd)
            nr = max(self.dist(u1_n1, u1_n), self.dist(u2_n1, u2_n))
            u1_n, u2_n = u1_n1, u2_n1
            if nr < tol:
                break
        if full:
            return u1_n, u2_n, it, nr, w1_n, w2_n
        return u1_n, u2_n, it, nr


def manufactured(pr_zone0, t, sigma, mu, sigma1, mu1, alpha):
    """The reference driver's analytic pair and forcing
    (examples/test_baroclin.cpp:77-131, 176-184): u1 = x sin(y+t) cos^4 x,
    u2 = x cos(y+t) cos^4 x with x, y the zone-0 chart coordinates, and the
    forcing rp_f1 / rp_g1 that makes them an exact solution without Coriolis
    (zero_coriolis).  pr_zone0: (V, 2) array of (x, y).  Returns
    (u1, u2, f, g)."""
    x, y = pr_zone0[:, 0], pr_zone0[:, 1]
    c, s = np.cos(x), np.sin(x)
    sy, cy = np.sin(y + t), np.cos(y + t)
    u1 = x * sy * c ** 4
    u2 = x * cy * c ** 4
    f = (-45 * mu * sy * x - 4.5 * sigma * c ** 3 * sy * s + 4.5 * sigma * c ** 3 * s * cy
         - 10 * sigma * c ** 4 * x * sy + 10 * sigma * c ** 4 * x * cy + 7.5 * sigma * c ** 2 * x * sy
         - 7.5 * sigma * c ** 2 * x * cy - 360 * mu * sy * s * c ** 3 + 147 * mu * sy * s * c
         - 400 * mu * sy * x * c ** 4 + 390 * mu * sy * x * c ** 2 - 20 * x * cy * c ** 4
         - 9 * cy * c ** 3 * s + 15 * x * cy * c ** 2)
    a2 = alpha * alpha
    g = (20 * x * sy * c ** 4 + 9 * sy * c ** 3 * s - 15 * x * sy * c ** 2
         + 390 * mu * cy * x * c ** 2 - 10 * sigma * c ** 4 * x * cy - 4.5 * sigma * c ** 3 * sy * s
         - a2 * c ** 7 * x - 45 * mu * cy * x + 18 * c ** 6 * cy ** 2 * s - 9 * c ** 6 * s
         + 9 * x * c ** 5 - 4.5 * sigma * c ** 3 * s * cy - 30 * x * x * c ** 4 * s
         - 18 * x * c ** 5 * cy ** 2 + a2 * c ** 4 * x * sy + 7.5 * sigma * c ** 2 * x * sy
         + 7.5 * sigma * c ** 2 * x * cy - 400 * mu * cy * x * c ** 4 + 147 * mu * cy * s * c
         + 60 * x * x * c ** 4 * cy ** 2 * s - 360 * mu * cy * s * c ** 3 - 10 * sigma * c ** 4 * x * sy
         + 4 * a2 * c ** 6 * x * x * s - 9 * a2 * c ** 3 * mu1 * cy * s - 20 * a2 * c ** 4 * mu1 * cy * x
         + 15 * a2 * c ** 2 * mu1 * cy * x - a2 * c ** 4 * sigma1 * x * cy) / a2
    return u1, u2, f, g


def norm_full(bc, u):
    return math.sqrt(float(np.dot(u, O.spmv(*bc.bv.norm_m, u))))
