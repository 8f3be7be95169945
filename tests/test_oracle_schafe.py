"""The oracle's SphereChafe (oracle/schafe.py) pinned by the reference
driver's analytic case (examples/test_schafe.cpp:25-34, 61-125): ans =
x sin(y + t) cos^4 x with the forcing SphereChafe_Private::f
(examples/schafe.cpp:55-63), tau-steps with the boundary values of ans.

The step evaluates f at the start of the step (schafe_impl.h:172-180), so
the scheme is first order in time: at a fixed end time the error is
c tau plus a smaller spatial part, so its differences halve with tau.  Planted mistakes in the
oracle (the sign of the sigma term of the right part, sigma left out of A_,
one term of f dropped) make it inconsistent, and the error then stays at
the size of the mistake instead of halving."""
import numpy as np
import pytest

pytest.importorskip("scipy")

import oracle.schafe as S
from oracle import mesh_ref
from oracle import oracle as O

MU, SIGMA, T_END = 1.0, 70.0, 0.04


def _mesh(level):
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))
    om = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(om.p2io), -1, 2)[:, 0, :]
    return om, xy


def run(om, xy, tau, plant=None):
    sc = S.SphereChafe(om, xy, tau, SIGMA, MU)
    if plant == "rhs_sigma_sign":
        sc.sigma = -SIGMA          # the right part's -sigma/2 u becomes +sigma/2 u
    elif plant == "A_without_sigma":
        import scipy.sparse.linalg as spla
        sc.lu_A = spla.splu(S._csr(om.assemble(O.KIND_HELMHOLTZ, 1.0 / tau, -0.5 * MU),
                                   om.n_inner).tocsc())
    U = S.schafe_ans(xy[:, 0], xy[:, 1], 0.0)
    for i in range(int(round(T_END / tau))):
        B = S.schafe_ans(xy[om.outer, 0], xy[om.outer, 1], tau * (i + 1))
        U = sc.solve(U, B, tau * i)
    return S.norm_dist(om, U, S.schafe_ans(xy[:, 0], xy[:, 1], T_END))


def test_analytic_case_first_order_in_time():
    om, xy = _mesh(4)
    e = [run(om, xy, tau) for tau in (0.01, 0.005, 0.0025)]
    # e(tau) = c tau + (spatial part): the differences halve with tau
    r = (e[0] - e[1]) / (e[1] - e[2])
    assert 1.8 < r < 2.4 and e[2] < 1e-3, (e, r)


@pytest.mark.parametrize("plant", ["rhs_sigma_sign", "A_without_sigma", "f_term"])
def test_planted_mistakes_break_convergence(plant, monkeypatch):
    om, xy = _mesh(3)
    ok = [run(om, xy, tau) for tau in (0.01, 0.005)]
    if plant == "f_term":
        good = S.schafe_f

        def bad_f(x, y, t, mu, sigma):   # the 9 mu sin(y + t) sin x cos x term dropped
            c = np.cos(x)
            return good(x, y, t, mu, sigma) - c * c * 9 * mu * np.sin(y + t) * np.sin(x) * c
        monkeypatch.setattr(S, "schafe_f", bad_f)
    e = [run(om, xy, tau, plant) for tau in (0.01, 0.005)]
    assert e[1] > 4 * ok[1] and e[0] / e[1] < 1.5, (plant, e, ok)
