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
  mistakes apart.
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
