"""Pins the oracle's SphereJacobian matrices (src/sjacobian.cpp:60-134,
7-point rule of src/quadrature.c:130-171) without reusing its formulas:

* DIFF1/DIFF2 integrate d_x phi_j phi_i (degree 1), which the 7-point rule
  integrates exactly, so every entry equals grad(phi_j)|_T * |T| / 3 summed
  over the triangles of the row's patch, computed here from the vertex
  coordinates by a 3x3 solve;
* the cos-weighted kinds are compared with a 12x12-point Gauss-Legendre
  rule on the Duffy-collapsed square, and their error must fall at the
  rate a degree-5 rule gives on these integrands;
* sum_j phi_j = 1 gives zero row sums of every derivative matrix once the
  boundary (south-pole) columns are added.
"""
import functools

import numpy as np
import pytest

from oracle import oracle as O
from tests._common import oracle_mesh

DIFF_KINDS = (O.KIND_DIFF1, O.KIND_DIFF2, O.KIND_DIFF1_COS, O.KIND_DIFF2_COS)


def _dense(rp, ci, vals, n_cols):
    A = np.zeros((len(rp) - 1, n_cols))
    for i in range(len(rp) - 1):
        for c in range(rp[i], rp[i + 1]):
            A[i, ci[c]] += vals[c]
    return A


def _gl_duffy(n=12):
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1.0)
    w = 0.5 * w
    U, Vv = np.meshgrid(g, g, indexing="ij")
    W = np.outer(w, w)
    return U.ravel(), Vv.ravel(), W.ravel()


_GL = _gl_duffy()


def _tri_integral(xs, ys, f):
    """int_T f(x, y) dx dy by the Duffy map (u, v) -> v0 + u(v1-v0) + uv(v2-v1)."""
    u, v, w = _GL
    x = xs[0] + u * (xs[1] - xs[0]) + u * v * (xs[2] - xs[1])
    y = ys[0] + u * (ys[1] - ys[0]) + u * v * (ys[2] - ys[1])
    det = abs((xs[1] - xs[0]) * (ys[2] - ys[1]) - (ys[1] - ys[0]) * (xs[2] - xs[1]))
    return float(np.sum(w * u * f(x, y)) * det)


def _basis(xs, ys):
    """Coefficients (a, b, c) of phi_k = a + b x + c y with phi_k(v_l) = delta_kl."""
    Amat = np.column_stack([np.ones(3), xs, ys])
    return np.linalg.solve(Amat, np.eye(3))  # column k holds (a, b, c) of phi_k


@functools.lru_cache(maxsize=32)
def _expected(level, kind, exact):
    """Dense inner x inner and inner x outer matrices of `kind`, each row in the
    chart of the first triangle around its vertex (src/generators.cpp:76)."""
    _, M = oracle_mesh(level)
    n, no = M.n_inner, len(M.outer)
    E = np.zeros((n, n))
    Eb = np.zeros((n, no))
    outer_idx = {int(p): j for j, p in enumerate(M.outer)}
    for i, p in enumerate(M.inner):
        z = M.tri_zone[M.adj[M.adj_ptr[p]]]
        for e in range(M.adj_ptr[p], M.adj_ptr[p + 1]):
            t = M.adj[e]
            pts = M.tri[t]
            xs = M.coords[pts, z, 0]
            ys = M.coords[pts, z, 1]
            B = _basis(xs, ys)
            li = int(np.where(pts == p)[0][0])
            area = 0.5 * abs((xs[0] - xs[2]) * (ys[1] - ys[2]) - (ys[0] - ys[2]) * (xs[1] - xs[2]))
            for a in range(3):
                q = int(pts[a])
                if kind in (O.KIND_DIFF1, O.KIND_DIFF2) and exact:
                    g = B[1, a] if kind == O.KIND_DIFF1 else B[2, a]
                    val = g * area / 3.0
                else:
                    ai, bi, ci_ = B[:, li]
                    phi_i = lambda x, y: ai + bi * x + ci_ * y
                    if kind == O.KIND_JAC_MASS:
                        aj, bj, cj = B[:, a]
                        f = lambda x, y: np.cos(x) * phi_i(x, y) * (aj + bj * x + cj * y)
                    else:
                        g = B[1, a] if kind in (O.KIND_DIFF1, O.KIND_DIFF1_COS) else B[2, a]
                        w = np.cos if kind in (O.KIND_DIFF1_COS, O.KIND_DIFF2_COS) else (lambda x: 1.0)
                        f = lambda x, y, g=g, w=w: w(x) * g * phi_i(x, y)
                    val = _tri_integral(xs, ys, f)
                if M.is_boundary[q]:
                    Eb[i, outer_idx[q]] += val
                else:
                    E[i, M.p2io[q]] += val
    return E, Eb


def _oracle_dense(level, kind):
    _, M = oracle_mesh(level)
    rp, ci, v = M.assemble(kind)
    brp, bci, bv = M.assemble_boundary(kind)
    return _dense(rp, ci, v, M.n_inner), _dense(brp, bci, bv, len(M.outer))


@pytest.mark.parametrize("kind", [O.KIND_DIFF1, O.KIND_DIFF2])
@pytest.mark.parametrize("level", [2, 3])
def test_diff_matrices_equal_gradient_form(level, kind):
    E, Eb = _expected(level, kind, True)
    A, Ab = _oracle_dense(level, kind)
    scale = np.abs(E).max(axis=1, keepdims=True)
    assert np.all(np.abs(A - E) <= 1e-12 * scale)
    assert np.all(np.abs(Ab - Eb) <= 1e-12 * scale)
    # the pole column is used by the rows of its six neighbours only
    _, M = oracle_mesh(level)
    assert len(M.boundary_structure()[1]) == 6


@pytest.mark.parametrize("kind", [O.KIND_JAC_MASS, O.KIND_DIFF1_COS, O.KIND_DIFF2_COS])
def test_cos_kinds_converge_as_degree5_rule(kind):
    errs = []
    for level in (2, 3):
        E, Eb = _expected(level, kind, False)
        A, Ab = _oracle_dense(level, kind)
        scale = np.abs(E).max()
        errs.append(max(np.abs(A - E).max(), np.abs(Ab - Eb).max()) / scale)
    # the rule's error term holds the 6th derivatives of the integrand; with
    # basis gradients ~1/h the error relative to the entries is O(h^4) for
    # phi_i phi_j cos x (ratio 16 per level) and O(h^5) for d phi_j phi_i cos x
    # (ratio 32).  A wrong table or a dropped weight gives an O(1) error.
    ratio = 10.0 if kind == O.KIND_JAC_MASS else 20.0
    assert errs[0] < 1e-6, errs
    assert errs[1] < errs[0] / ratio, errs


@pytest.mark.parametrize("kind", DIFF_KINDS)
def test_diff_row_sums_vanish(kind):
    A, Ab = _oracle_dense(3, kind)
    s = A.sum(axis=1) + Ab.sum(axis=1)
    scale = np.abs(A).max(axis=1)
    assert np.all(np.abs(s) <= 1e-13 * scale)


def test_jacobian_mass_close_to_integrate1_mass():
    # id_cb of SphereJacobian (7-point rule) and of SphereLaplace (GK15
    # trapezoids) integrate the same cos x phi_i phi_j
    _, M = oracle_mesh(3)
    _, _, mj = M.assemble(O.KIND_JAC_MASS)
    _, _, mm = M.assemble(O.KIND_MASS)
    assert np.abs(mj - mm).max() <= 1e-6 * np.abs(mm).max()


# ---------------------------------------------------------------- calc2 --
def _ring_distance_to_zone_change(M):
    """Graph distance (CSR stencil rings) of every inner row to the nearest
    row assembled in another chart."""
    rp, ci = M.csr_structure()
    n = M.n_inner
    zone = np.array([M.tri_zone[M.adj[M.adj_ptr[p]]] for p in M.inner])
    dist = np.full(n, -1)
    front = [i for i in range(n) if np.any(zone[ci[rp[i]:rp[i + 1]]] != zone[i])]
    dist[front] = 0
    d = 0
    while front:
        d += 1
        nxt = []
        for i in front:
            for j in ci[rp[i]:rp[i + 1]]:
                if dist[j] < 0:
                    dist[j] = d
                    nxt.append(j)
        front = nxt
    return dist


@functools.lru_cache(maxsize=8)
def _jacobian_error(level):
    """max |calc2(z, x) - y| over rows at least 2^(level-1) rings (a fixed
    distance on the sphere) away from a change of chart.  For the coordinate
    functions of the unit sphere J(z, x) = (u_lam v_phi - u_phi v_lam)/cos phi
    = cos phi sin lam = y."""
    from oracle import mesh_ref
    from oracle.sjacobian import SphereJacobian
    text, M = oracle_mesh(level)
    pr = np.asarray(mesh_ref.load_text(text)["pr"]).reshape(-1, 3)
    J = SphereJacobian(M)
    ans = J.calc2(pr[:, 2], pr[:, 0])
    err = np.abs(ans - pr[M.inner, 1])
    far = _ring_distance_to_zone_change(M) >= 2 ** (level - 1)
    return err[far].max(), J, pr


def test_jacobian_converges_inside_charts():
    # The reference assembles each row in its own chart, so the mass solves
    # mix derivatives taken in different charts and the error is O(1) along
    # chart boundaries; away from them calc2 converges at first order.  A
    # swapped operator pair or sign in calc2_1..4 gives O(1) errors here.
    e3, e4, e5 = (_jacobian_error(k)[0] for k in (3, 4, 5))
    assert e3 < 2e-2 and e4 < 0.6 * e3 and e5 < 0.6 * e4, (e3, e4, e5)


def test_jacobian_antisymmetry():
    _, J, pr = _jacobian_error(3)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(len(pr)), rng.standard_normal(len(pr))
    a = J.calc2(u, v)
    assert np.array_equal(a, -J.calc2(v, u))   # exact: the four variants pair up
    assert np.all(J.calc2(u, u) == 0.0)
    assert np.abs(J.calc2(np.ones(len(pr)), v)).max() < 1e-12 * np.abs(a).max()
    assert np.array_equal(J.calc2t(u, v), -a)
