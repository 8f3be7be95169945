"""The reference's own known-answer runs on the oracle, with the
reference's thresholds:

* SphereLaplace (examples/test_slaplace.cpp:99-199, the double run):
  L1 invert err = dist(solve(func_lapl, bnd), func), L2 laplace err =
  dist(calc1(func, bnd = func_lapl), func_lapl), L3 laplace err =
  dist(func, solve(calc1(...), bnd = func)); scripts/test_slaplace.lua:
  e3 <= 1e-7 at levels 3-5, e1(L3)/e1(L4) >= 3.3 and e1(L4)/e1(L5) >= 3.3;
* SphereJacobian (examples/test_barvortex.cpp:74-98): jacobian err =
  dist(calc1(f1, f2, bnd = an), an); scripts/test_jacobian.lua: <= 1 at
  level 4.
dist is SphereNorm's (sphere_scalar_cb over generate_full_matrix)."""
import numpy as np
import pytest

pytest.importorskip("scipy")

from oracle import kat_fields as K
from oracle import mesh_ref
from oracle import oracle as O
from oracle.barvortex import BarVortex as OB
from oracle.sjacobian import SphereJacobian


def _setup(level):
    a = mesh_ref.load_text(mesh_ref.sphere_mesh_text(level))
    om = O.Mesh(a)
    xy = np.asarray(a["coords"]).reshape(len(om.p2io), -1, 2)[:, 0, :]
    # the BarVortex restatement carries SphereLaplace's calc1 / solve and SphereNorm
    ob = OB(om, tau=1.0, sigma=0.0, mu=0.0, k1=0.0, k2=0.0, theta=0.5, lh=np.zeros(len(xy)))
    return om, xy, ob


def slaplace_errors(level):
    om, xy, ob = _setup(level)
    x, y = xy[:, 0], xy[:, 1]
    xo, yo = xy[om.outer, 0], xy[om.outer, 1]
    F, rans = K.slaplace_func_lapl(x, y), K.slaplace_func(x, y)
    e1 = ob.norm(ob.poisson(F, K.slaplace_func(xo, yo)) - rans)
    U, LU = K.slaplace_func(x, y), K.slaplace_func_lapl(x, y)
    LU1 = ob.calc1(U, K.slaplace_func_lapl(xo, yo))
    e2 = ob.norm(LU - LU1)
    e3 = ob.norm(U - ob.poisson(LU1, K.slaplace_func(xo, yo)))
    return e1, e2, e3


def test_slaplace_run_meets_reference_thresholds():
    e = {lv: slaplace_errors(lv) for lv in (3, 4, 5)}
    print({lv: ["%.2e" % v for v in e[lv]] for lv in e})
    for lv in e:
        assert e[lv][2] <= 1e-7, (lv, e[lv])
    assert e[3][0] / e[4][0] >= 3.3 and e[4][0] / e[5][0] >= 3.3, e
    # L2 (calc1 against the analytic Laplacian) falls with refinement too
    assert e[5][1] < e[4][1] < e[3][1], e


def jacobian_errors(level):
    """(the reference's jacobian err over the whole sphere, the rms error at
    latitudes |x| < 1, the same near the poles |x| >= 1)."""
    om, xy, ob = _setup(level)
    x, y = xy[:, 0], xy[:, 1]
    F1, F2 = K.jacobian_f1(x, y), K.jacobian_f2(x, y)
    rans = K.jacobian_an(x, y)
    bnd = K.jacobian_an(xy[om.outer, 0], xy[om.outer, 1])
    d = ob.p2u(SphereJacobian(om).calc2(F1, F2), bnd) - rans
    band = np.abs(x) < 1.0
    return (ob.norm(d), float(np.sqrt(np.mean(d[band] ** 2))),
            float(np.sqrt(np.mean(d[~band & (np.abs(np.cos(x)) > 1e-8)] ** 2))))


def test_jacobian_run_and_the_pole_singularity():
    """The reference's threshold (jacobian err <= 1 at level 4) cannot hold on
    the full sphere: test_jacobian_an divides by cos x, and the north pole is
    an inner point with x = pi/2 in the zone-0 chart (cos x = 6.1e-17), where
    the analytic value is -1.6e16 and the discrete Jacobian is O(10).  Away
    from the poles the discrete Jacobian converges to the analytic one (the
    rms error at |x| < 1 falls more than 2x per level); next to them it does
    not (the charts' Jacobian mixing, O(1))."""
    e = {lv: jacobian_errors(lv) for lv in (3, 4, 5)}
    print({lv: ["%.3e" % v for v in e[lv]] for lv in e})
    assert e[4][0] > 1e14                      # dominated by the north pole point
    assert e[3][1] > 2 * e[4][1] > 4 * e[5][1] and e[5][1] < 5e-3, e
    assert all(e[lv][2] > 0.1 for lv in e)     # no convergence next to the poles
