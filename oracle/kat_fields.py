"""The analytic fields of the reference's known-answer drivers, evaluated as
the drivers evaluate them (C library sin/cos, ipow(v, p) as repeated
multiplication from 1).

TEST INFRASTRUCTURE ONLY.

* func / func_lapl / bnd: examples/test_slaplace.cpp:25-72 (SphereLaplace
  run: L1 invert, L2 laplace, L3 laplace; thresholds
  scripts/test_slaplace.lua:20-55);
* test_jacobian_f1 / f2 / an: examples/test_barvortex.cpp:57-73 (the
  SphereJacobian run of :74-98; threshold scripts/test_jacobian.lua:24-34).
All take the zone-0 chart coordinates (x, y) = MeshPoint::x(), y()."""
import numpy as np


def _ipow2(v):
    return (1.0 * v) * v


def slaplace_func(x, y):
    sx, cx, cy = np.sin(x), np.cos(x), np.cos(y)
    return 0.5 * _ipow2(cx * cy - 1.0) + _ipow2(cx * cy - sx) + 2.0 * _ipow2(sx - 3)


def slaplace_func_lapl(x, y):
    sx, cx, sy, cy = np.sin(x), np.cos(x), np.sin(y), np.cos(y)
    ans = np.zeros_like(np.asarray(x, dtype=np.float64))
    ans = ans + (sx * sx * (cx * cy - 1) * cy)
    ans = ans + -(cx * cy - 1) * cy
    ans = ans + -2 * sx * (cx * cy - sx) * (-sx * cy - cx)
    ans = ans + -2 * (cx * cy - sx) * cy
    ans = np.where(np.abs(ans) > 1e-15, ans / cx, ans)
    ans = ans + sx * sx * cy * cy
    ans = ans + -cx * (cx * cy - 1) * cy
    ans = ans + sy * sy
    ans = ans + 2 * _ipow2(-sx * cy - cx)
    ans = ans + 2 * (cx * cy - sx) * (-cx * cy + sx)
    ans = ans + 2 * sy * sy
    ans = ans + -8 * (sx - 3.0) * sx + 4.0 * cx * cx
    return ans


def jacobian_f1(x, y):
    return np.sin(x) * np.sin(y)


def jacobian_f2(x, y):
    return np.cos(x) * np.cos(y)


def jacobian_an(x, y):
    return (-np.sin(x) * np.cos(y) * np.sin(x) * np.cos(y)
            + np.cos(x) * np.sin(y) * np.cos(x) * np.sin(y)) / np.cos(x)
