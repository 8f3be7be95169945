"""Pins the oracle's Baroclin restatement (oracle/baroclin.py,
examples/baroclin.cpp:69-181, 445-606) on the CPU:

* the nine blocks of A_ are the linear combinations of the (pinned) mass and
  stiffness matrices that integrate_cb's element formulas give, to rounding;
* the reference driver's manufactured solution (examples/test_baroclin.cpp:
  77-131, 176-184: u1 = x sin(y+t) cos^4 x, u2 = x cos(y+t) cos^4 x with the
  forcing rp_f1 / rp_g1, no Coriolis): one step's tendency (u(tau) - u(0)) /
  tau approaches the exact one as the mesh is refined, for both layers;
* a mistake planted in a copy of the oracle (the alpha^2 J(u1, u2) term of
  the second layer dropped) breaks that convergence, so the check tells such
  mistakes apart;
* the linearised steps calc_L / calc_L_1 (baroclin.cpp:607-986): Ab_'s blocks
  are integrate_backward_cb's mass/stiffness combinations; without Coriolis
  and about z = 0 (no Jacobian terms) calc_L is A_^-1 Ab_ and calc_L_1 is
  Ab_^-1 A_ on consistent states, so L_1(L(u)) = u exactly; the Jacobian part
  of calc_L equals a central difference of the nonlinear step about z to
  O(tau) in the two subspaces where the code's J(L z1, avg w2) and
  J(L z2, avg u1) (the comments write J(z1, .) and J(., L z2)) vanish:
  (u1, z1) with u2 = z2 = 0, and (u2, z2) with u1 = z1 = 0 (linearising
  about -z misses by 2x); with the Jacobian terms L_1(L(u)) = u + O(tau^2).
"""
import numpy as np
import pytest

pytest.importorskip("scipy")

from oracle import mesh_ref, oracle as O
from oracle.baroclin import Baroclin, manufactured

PRM = dict(sigma=1.6e-2, mu=8e-2, sigma1=1.6e-2, mu1=8e-2, alpha=1.0)
TAU = 1e-3


def _mesh(level):
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))
    M = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(M.p2io), -1, 2)[:, 0, :]
    return M, xy


def test_blocks_are_the_mass_stiffness_combinations():
    M, _ = _mesh(3)
    tau, th = TAU, 0.5
    s, mu, s1, mu1, al = PRM["sigma"], PRM["mu"], PRM["sigma1"], PRM["mu1"], PRM["alpha"]
    rp, ci, v9 = O.baroclin_blocks(M, tau, s, mu, s1, mu1, al, th)
    m = M.assemble(O.KIND_MASS)[2]
    lap = M.assemble(O.KIND_LAPLACE)[2]
    want = [(1 / tau + 0.5 * th * s) * m - th * mu * lap, -0.5 * s * th * m, 0.5 * th * s * m,
            (1 / tau + 0.5 * th * s + al * al * th * mu1) * m - th * mu * lap,
            (-al * al / tau - al * al * th * s1) * m, m, -lap, m, -lap]
    for k in range(9):
        assert np.abs(v9[k] - want[k]).max() <= 1e-15 * np.abs(want[k]).max(), k
    # the 4n x 4n CSR: block (r, c) at rows r n.., columns c n..
    brp, bci, bv = O.baroclin_matrix(rp, ci, v9)
    n = M.n_inner
    assert len(brp) == 4 * n + 1 and brp[-1] == 9 * len(ci)
    assert all(np.all(np.diff(bci[brp[i]:brp[i + 1]]) > 0) for i in range(0, 4 * n, 97))


def _tendency_errors(cls, levels):
    out = []
    for lv in levels:
        M, xy = _mesh(lv)
        bc = cls(M, tau=TAU, lh=np.zeros(len(xy)), **PRM)
        u1, u2, f, g = manufactured(xy, 0.0, **PRM)
        r1, r2, _, _ = manufactured(xy, TAU, **PRM)
        v1, v2, it, nr = bc.calc(u1, u2, f, g)
        assert nr < 1e-8 and it <= 20
        e1 = bc.dist(v1 - u1, r1 - u1) / bc.bv.norm(r1 - u1)
        e2 = bc.dist(v2 - u2, r2 - u2) / bc.bv.norm(r2 - u2)
        out.append((e1, e2))
    return out


def _converges(errs):
    """Levels 3 -> 4: layer 1 tendency error 0.118 -> 0.040 (x3.0), layer 2
    0.190 -> 0.079 (x2.4) for the restatement."""
    (a1, a2), (b1, b2) = errs
    return b1 < 0.05 and b1 < a1 / 2.5 and b2 < 0.1 and b2 < a2 / 2.2


def test_manufactured_solution_tendency_converges():
    errs = _tendency_errors(Baroclin, (3, 4))
    assert _converges(errs), errs


class _NoJac3(Baroclin):
    """Planted mistake: the alpha^2 J(u1, u2) term of layer 2 dropped."""
    calls = 0

    def jac(self, u, v):                        # every fifth call is jac3
        _NoJac3.calls += 1
        r = super().jac(u, v)
        return 0.0 * r if _NoJac3.calls % 5 == 0 else r


class _NoG(Baroclin):
    """Planted mistake: layer 2's forcing g not added to GC."""

    def calc(self, u1, u2, f=None, g=None, **k):
        return super().calc(u1, u2, f, None, **k)


@pytest.mark.parametrize("wrong", [_NoJac3, _NoG])
def test_planted_mistake_breaks_convergence(wrong):
    errs = _tendency_errors(wrong, (3, 4))
    assert not _converges(errs), errs


def test_backward_blocks_are_the_mass_stiffness_combinations():
    M, _ = _mesh(3)
    tau, th = TAU, 0.5
    s, mu, s1, mu1, al = PRM["sigma"], PRM["mu"], PRM["sigma1"], PRM["mu1"], PRM["alpha"]
    rp, ci, v9 = O.baroclin_blocks(M, tau, s, mu, s1, mu1, al, th, backward=True)
    m = M.assemble(O.KIND_MASS)[2]
    lap = M.assemble(O.KIND_LAPLACE)[2]
    b = 1.0 - th
    want = [(1 / tau - 0.5 * b * s) * m + b * mu * lap, 0.5 * s * b * m, -0.5 * b * s * m,
            (1 / tau - 0.5 * b * s - al * al * b * mu1) * m + b * mu * lap,
            (-al * al / tau + al * al * b * s1) * m, m, -lap, m, -lap]
    for k in range(9):
        assert np.abs(v9[k] - want[k]).max() <= 1e-15 * np.abs(want[k]).max(), k


def _lin_fields(level):
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))
    M = O.Mesh(a)
    pr = np.asarray(a["pr"]).reshape(-1, 3)
    x, y, z = pr[:, 0], pr[:, 1], pr[:, 2]

    def field(seed):
        r = np.random.default_rng(seed)
        f = (r.standard_normal() * x * y + r.standard_normal() * (x * x - y * y) * z
             + r.standard_normal() * np.sin(2 * x + y) * z + 0.3 * r.standard_normal() * y)
        return f - f[M.outer[0]]

    return M, field, 2.0 * np.clip(z, -1, 1)


def _rel(bc, a, b, ref):
    return float(np.hypot(bc.dist(a[0], b[0]), bc.dist(a[1], b[1]))
                 / np.hypot(bc.bv.norm(ref[0]), bc.bv.norm(ref[1])))


def test_linear_steps_without_jacobian_are_inverse():
    M, field, lh = _lin_fields(3)
    zero = np.zeros(len(lh))
    bc = Baroclin(M, tau=1e-2, lh=zero, **PRM)
    u = (field(1), field(2))
    L = bc.calc_lin(0, u[0], u[1], zero, zero)[:2]
    assert _rel(bc, L, u, u) > 1e-3                        # the step moves the state
    back = bc.calc_lin(1, L[0], L[1], zero, zero)
    assert back[2] == 2 and back[3] == 0.0                 # linear: the second pass repeats
    assert _rel(bc, back[:2], u, u) <= 1e-12


def test_linear_step_jacobian_part_is_the_derivative():
    M, field, lh = _lin_fields(3)
    zero = np.zeros(len(lh))
    u1, u2, z1, z2 = field(1), field(2), field(3), field(4)
    for du, zz in (((u1, zero), (z1, zero)), ((zero, u2), (zero, z2))):
        rel = []
        for tau in (4e-3, 1e-3):
            bc = Baroclin(M, tau=tau, lh=lh, **PRM)
            b0 = Baroclin(M, tau=tau, lh=zero, **PRM)
            eps = 1e-5
            sp = bc.calc(zz[0] + eps * du[0], zz[1] + eps * du[1], max_iters=100, tol=1e-15)
            sm = bc.calc(zz[0] - eps * du[0], zz[1] - eps * du[1], max_iters=100, tol=1e-15)
            D = [(sp[k] - sm[k]) / (2 * eps) for k in (0, 1)]
            L = bc.calc_lin(0, du[0], du[1], zz[0], zz[1])[:2]
            P = b0.calc_lin(0, du[0], du[1], zero, zero)[:2]
            jl = [L[k] - P[k] for k in (0, 1)]
            jd = [D[k] - P[k] for k in (0, 1)]
            rel.append(_rel(bc, jl, jd, jl))
            bad = bc.calc_lin(0, du[0], du[1], -zz[0], -zz[1])[:2]
            assert _rel(bc, [bad[k] - P[k] for k in (0, 1)], jd, jl) > 1.5
        assert rel[0] < 0.1 and rel[1] < 0.03 and rel[1] < rel[0] / 3.0, rel


def test_inverse_step_inverts_to_second_order():
    M, field, lh = _lin_fields(3)
    u = (field(1), field(2))
    z1, z2 = field(3), field(4)
    errs = []
    for tau in (1e-3, 2.5e-4):
        bc = Baroclin(M, tau=tau, lh=lh, **PRM)
        L = bc.calc_lin(0, u[0], u[1], z1, z2)[:2]
        back = bc.calc_lin(1, L[0], L[1], z1, z2)
        errs.append(_rel(bc, back[:2], u, u))
    assert errs[1] < errs[0] / 10.0 and errs[1] < 1e-4, errs


def test_block_kinds_reproduce_the_blocks():
    """The (kind, c_mass, c_lap) table behind the boundary blocks assembles
    the nine inner blocks of or_assemble_baroclin's literal formulas."""
    from oracle.baroclin import block_kinds
    M, _ = _mesh(3)
    rp, ci, v9 = O.baroclin_blocks(M, TAU, PRM["sigma"], PRM["mu"], PRM["sigma1"], PRM["mu1"],
                                   PRM["alpha"], 0.5)
    kinds = block_kinds(TAU, PRM["sigma"], PRM["mu"], PRM["sigma1"], PRM["mu1"], PRM["alpha"], 0.5)
    for k, (kd, cm, cl) in enumerate(kinds):
        got = M.assemble(kd, cm, cl)[2]
        assert np.abs(got - v9[k]).max() <= 1e-14 * np.abs(v9[k]).max(), k


def test_boundary_values_are_a_dirichlet_elimination():
    """right_part_cb's boundary branch (baroclin.cpp:259-323) is the
    elimination of known boundary values: for any four full fields whose
    boundary values all equal bnd, A_ times their inner values equals the
    inner rows of the FULL block products (generate_full_matrix, another
    assembly route) plus bnd_rhs(bnd).  A dropped block, a wrong sign or a
    boundary index mix-up fails it."""
    M, _ = _mesh(3)
    bc = Baroclin(M, TAU, lh=np.zeros(len(M.p2io)), **PRM)
    rng = np.random.default_rng(5)
    V, inner, outer, n = len(M.p2io), M.inner, M.outer, M.n_inner
    bnd = rng.standard_normal(len(outer))
    X = rng.standard_normal((4, V))
    X[:, outer] = bnd
    rhs = np.zeros(4 * n)
    for (r, c), (kd, cm, cl) in zip(O.BAROCLIN_BLOCKS, bc.kinds):
        frp, fci, fv = M.assemble_full(kd, cm, cl)
        rhs[r * n:(r + 1) * n] += O.spmv(frp, fci, fv, X[c])[inner]
    brp, bci, bv = bc.A
    lhs = O.spmv(brp, bci, bv, np.concatenate([X[c][inner] for c in range(4)]))
    err = np.abs(lhs - (rhs + bc.bnd_rhs(bnd))).max()
    assert err <= 1e-10 * np.abs(rhs).max(), err
    assert np.abs(bc.bnd_rhs(bnd)).max() > 1e-3 * np.abs(rhs).max()   # the term matters


def test_step_takes_the_boundary_values():
    """The step writes bnd into the four results' boundary points, zero
    boundary values give the step without them bit for bit, and non-zero
    ones change it."""
    M, xy = _mesh(3)
    bc = Baroclin(M, TAU, lh=np.zeros(len(M.p2io)), **PRM)
    u1, u2 = manufactured(xy, 0.0, **PRM)[:2]
    a = bc.calc(u1, u2, max_iters=5)
    b = bc.calc(u1, u2, max_iters=5, bnd=np.zeros(len(M.outer)))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    c = bc.calc(u1, u2, max_iters=5, bnd=np.full(len(M.outer), 0.5))
    assert np.all(c[0][M.outer] == 0.5) and np.all(c[1][M.outer] == 0.5)
    assert bc.dist(c[0], a[0]) > 1e-6 * bc.dist(a[0], 0 * a[0])


def _no_jacobian(bc):
    bc.jac = lambda u, v: np.zeros(bc.n)
    return bc


def _unreduced_residual(bc, kinds, X, F, G):
    """Inner rows of the FULL block system (generate_full_matrix per block
    element formula) applied to the four full fields X, minus the right part
    [M F; M G; 0; 0]; relative to the largest block product."""
    M, inner, n = bc.m, bc.m.inner, bc.n
    res = np.zeros(4 * n)
    res[:n] -= O.right_part_mass(M, F)
    res[n:2 * n] -= O.right_part_mass(M, G)
    scale = 0.0
    for (r, c), (kd, cm, cl) in zip(O.BAROCLIN_BLOCKS, kinds):
        frp, fci, fv = M.assemble_full(kd, cm, cl)
        t = O.spmv(frp, fci, fv, X[c])[inner]
        res[r * n:(r + 1) * n] += t
        scale = max(scale, np.abs(t).max())
    return np.abs(res).max() / scale


def test_linear_steps_take_the_boundary_values():
    """Boundary (pole) values in calc_L / calc_L_1 (baroclin.cpp:620-986):
    * with every Jacobian term removed, calc_L is the nonlinear step's linear
      part, so with the same bnd it equals Baroclin::calc (whose boundary
      handling the two tests above pin); dropping bnd moves it far more;
    * with every Jacobian term removed, either step's answer, completed with
      bnd at the boundary points, solves the inner rows of the UNREDUCED
      system of its own element formulas (A_'s for calc_L, Ab_'s for
      calc_L_1; generate_full_matrix, another assembly route); at sigma,
      mu != 0 the two sets differ, so the other step's formulas, a dropped
      boundary term or a wrong sign fail it;
    * the results hold bnd at the boundary points, zero values give the plain
      steps bit for bit and non-zero ones change them.
    Where bnd enters L(z) (calc1(z, bnd), :681-682, :865-866: only the
    Jacobian terms see it) follows the code: pinned by reading only."""
    M, field, lh = _lin_fields(3)
    u1, u2, z1, z2 = field(1), field(2), field(3), field(4)
    bnd = np.full(len(M.outer), 0.37)
    bc = _no_jacobian(Baroclin(M, tau=1e-2, lh=lh, **PRM))
    L = bc.calc_lin(0, u1, u2, z1, z2, bnd=bnd)
    S = bc.calc(u1, u2, max_iters=3, bnd=bnd)
    assert _rel(bc, L[:2], S[:2], S[:2]) <= 1e-13
    assert _rel(bc, bc.calc_lin(0, u1, u2, z1, z2)[:2], S[:2], S[:2]) > 1e-4
    for which, kinds, other in ((0, bc.kinds, bc.kinds_b), (1, bc.kinds_b, bc.kinds)):
        r = bc.calc_lin(which, u1, u2, z1, z2, bnd=bnd, full=True)
        X = [r[4], r[5], r[0], r[1]]
        for x in X:
            assert np.all(x[M.outer] == 0.37)
        FC, GC = bc.lin_fc(which, u1, u2, bnd)[:2]
        F, G = FC[M.inner], GC[M.inner]
        assert _unreduced_residual(bc, kinds, X, F, G) <= 1e-12, which
        assert _unreduced_residual(bc, other, X, F, G) > 1e-6, which
        Xz = [x.copy() for x in X]
        for x in Xz:
            x[M.outer] = 0.0                           # the answer without its boundary values
        assert _unreduced_residual(bc, kinds, Xz, F, G) > 1e-6, which
    # zero values: the plain steps' bits; non-zero values: other results
    full = Baroclin(M, tau=1e-2, lh=lh, **PRM)
    zero = np.zeros(len(M.outer))
    for which in (0, 1):
        a = full.calc_lin(which, u1, u2, z1, z2)
        b = full.calc_lin(which, u1, u2, z1, z2, bnd=zero)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        c = full.calc_lin(which, u1, u2, z1, z2, bnd=bnd)
        assert np.all(c[0][M.outer] == 0.37) and np.all(c[1][M.outer] == 0.37)
        assert _rel(full, c[:2], a[:2], a[:2]) > 1e-4


def test_backward_boundary_values_are_a_dirichlet_elimination():
    """right_part_backward_cb's boundary branch (baroclin.cpp:341-436) is the
    elimination of known boundary values for Ab_: for four full fields whose
    boundary values all equal bnd, Ab_ times their inner values equals the
    inner rows of the FULL products of integrate_backward_cb's formulas plus
    bnd_rhs(bnd, backward=True)."""
    M, _ = _mesh(3)
    bc = Baroclin(M, TAU, lh=np.zeros(len(M.p2io)), **PRM)
    bc.lu_Ab                                            # builds Ab_
    rng = np.random.default_rng(6)
    V, inner, outer, n = len(M.p2io), M.inner, M.outer, M.n_inner
    bnd = rng.standard_normal(len(outer))
    X = rng.standard_normal((4, V))
    X[:, outer] = bnd
    rhs = np.zeros(4 * n)
    for (r, c), (kd, cm, cl) in zip(O.BAROCLIN_BLOCKS, bc.kinds_b):
        frp, fci, fv = M.assemble_full(kd, cm, cl)
        rhs[r * n:(r + 1) * n] += O.spmv(frp, fci, fv, X[c])[inner]
    brp, bci, bv = bc.Ab
    lhs = O.spmv(brp, bci, bv, np.concatenate([X[c][inner] for c in range(4)]))
    err = np.abs(lhs - (rhs + bc.bnd_rhs(bnd, backward=True))).max()
    assert err <= 1e-10 * np.abs(rhs).max(), err
    err_fwd = np.abs(lhs - (rhs + bc.bnd_rhs(bnd))).max()
    assert err_fwd > 1e-6 * np.abs(rhs).max(), err_fwd
