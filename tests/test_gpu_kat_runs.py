"""The reference's known-answer runs through the product classes on the GPU
(mke_amd.SphereLaplace, SphereJacobian, SphereNorm over the C ABI), with the
reference's thresholds and against the oracle's values of the same runs
(tests/test_oracle_kat_runs.py):

* SphereLaplace (examples/test_slaplace.cpp:99-199; scripts/
  test_slaplace.lua:20-55): e3 <= 1e-7 at levels 3-5, e1 ratios >= 3.3
  from level 3 to 4 to 5; e1 and e2 equal the oracle's to 1e-4 of their
  size (the GPU Poisson solve stops at a relative residual of 1e-12, the
  oracle's is a direct solve);
* SphereJacobian (examples/test_barvortex.cpp:74-98): the error against
  the analytic answer away from the poles equals the oracle's to 1e-6 of
  it, and the whole-sphere error is the oracle's (dominated by the north
  pole point, where the analytic answer divides by cos x = 6.1e-17)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytest.importorskip("scipy")

import mke_amd as P
from oracle import kat_fields as K
from tests.test_oracle_kat_runs import jacobian_errors, slaplace_errors

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def _mesh(level):
    pm = P.Mesh.sphere(level)
    xy = pm.zone0_xy()
    outer = pm.arrays()["outer"]
    return pm, xy[:, 0], xy[:, 1], xy[outer, 0], xy[outer, 1]


def gpu_slaplace_errors(level):
    pm, x, y, xo, yo = _mesh(level)
    lap = P.SphereLaplace(pm)
    nr = P.SphereNorm(pm)
    V = pm.n_points
    F, rans = _dev(K.slaplace_func_lapl(x, y)), _dev(K.slaplace_func(x, y))
    ans = torch.zeros(V, dtype=torch.float64, device=DEV)
    info = lap.solve(ans, F, _dev(K.slaplace_func(xo, yo)), rtol=1e-12)
    assert info["rel_residual"] <= 1e-11, info
    e1 = nr.dist(ans, rans)
    U, LU = _dev(K.slaplace_func(x, y)), _dev(K.slaplace_func_lapl(x, y))
    LU1 = torch.zeros(V, dtype=torch.float64, device=DEV)
    lap.calc1(LU1, U, _dev(K.slaplace_func_lapl(xo, yo)))
    e2 = nr.dist(LU, LU1)
    back = torch.zeros(V, dtype=torch.float64, device=DEV)
    lap.solve(back, LU1, _dev(K.slaplace_func(xo, yo)), rtol=1e-12)
    e3 = nr.dist(U, back)
    return e1, e2, e3


def test_slaplace_run_on_gpu():
    print("SphereLaplace known-answer run, levels 3-5", flush=True)
    e = {lv: gpu_slaplace_errors(lv) for lv in (3, 4, 5)}
    print({lv: ["%.3e" % v for v in e[lv]] for lv in e}, flush=True)
    for lv in e:
        assert e[lv][2] <= 1e-7, (lv, e[lv])
        o = slaplace_errors(lv)
        assert abs(e[lv][0] - o[0]) <= 1e-4 * o[0] and abs(e[lv][1] - o[1]) <= 1e-4 * o[1], \
            (lv, e[lv], o)
    assert e[3][0] / e[4][0] >= 3.3 and e[4][0] / e[5][0] >= 3.3, e


@pytest.mark.parametrize("level", [4, 5])
def test_jacobian_run_on_gpu(level):
    print("SphereJacobian known-answer run, level %d" % level, flush=True)
    pm, x, y, xo, yo = _mesh(level)
    J = P.SphereJacobian(pm)
    ans = torch.zeros(pm.n_points, dtype=torch.float64, device=DEV)
    p1 = torch.zeros(pm.inner_size, dtype=torch.float64, device=DEV)
    infos = J.calc1(ans, _dev(K.jacobian_f1(x, y)), _dev(K.jacobian_f2(x, y)),
                    _dev(K.jacobian_an(xo, yo)), p1)
    assert all(i["converged"] for i in infos)
    d = ans.cpu().numpy() - K.jacobian_an(x, y)
    band = np.abs(x) < 1.0
    rms = float(np.sqrt(np.mean(d[band] ** 2)))
    whole = P.SphereNorm(pm).dist(ans, _dev(K.jacobian_an(x, y)))
    o = jacobian_errors(level)
    print("  gpu: whole %.3e, band rms %.3e; oracle: %.3e, %.3e" % (whole, rms, o[0], o[1]),
          flush=True)
    assert abs(rms - o[1]) <= 1e-6 * o[1], (rms, o)
    assert abs(whole - o[0]) <= 1e-6 * o[0], (whole, o)
