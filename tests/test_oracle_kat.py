"""Pins the oracle against the reference's own known-answer tests.

tests/test_quadrature.c:27-40, tests/test_integrate.cpp:26-274 and
tests/test_polynom.cpp:20-56 of the reference, re-run against the oracle
with the reference's 1e-14 tolerance; the GK15 and 7-point rules are also
compared bit for bit with the reference's own src/quadrature.c built into
oracle/_ref (when that library was built here)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

L = O.lib()
REF_SO = os.path.join(os.path.dirname(O.__file__), "_ref", "libquadrature_ref.so")


def check(a, b, tol=1e-14):
    assert abs(a - b) <= tol, "%.16f != %.16f" % (a, b)


def test_gk15_kat():
    # test_quadrature.c:27-40
    f1 = O.FX(lambda x, d: x * x)
    f2 = O.FX(lambda x, d: x / math.cos(x))
    check(L.or_gauss_kronrod15(0.0, 1.0, f1, None), 0.3333333333333333333)
    check(L.or_gauss_kronrod15(-1.0, 1.0, f2, None), 0.0)


def test_gk15_polynomial_exactness():
    # the 15-point Kronrod rule integrates degree <= 22 exactly up to the
    # 15-digit truncation of its tables (a few 1e-15 relative)
    for deg in range(0, 23):
        f = O.FX(lambda x, d, deg=deg: x ** deg)
        got = L.or_gauss_kronrod15(0.0, 1.0, f, None)
        assert abs(got - 1.0 / (deg + 1)) < 5e-14 * 10 ** (deg / 22.0), deg


def _tri(p1, p2, p3):
    xs = (C.c_double * 3)(p1[0], p2[0], p3[0])
    ys = (C.c_double * 3)(p1[1], p2[1], p3[1])
    return xs, ys


def _xmone_ymtwo():
    # (P2X - 1) * (P2Y - 2): polynom.h operators
    a = O.poly(1, 0, [-1.0, 1.0])
    b = O.poly(0, 1, [-2.0, 1.0])
    r = O.OrPoly()
    L.or_poly_mul(C.byref(a), C.byref(b), C.byref(r))
    return r


def test_integrate_simple():
    # test_integrate.cpp:26-46 and :48-54
    xs, ys = _tri((0, 0), (1, 0), (1, 1))
    p = _xmone_ymtwo()
    check(L.or_integrate(C.byref(p), xs, ys, 0), 0.2916666666666666)
    check(L.or_trapezoid_integral(0, 0, 1, 0, 0, 1, 0.5, 1), 0.125)


def test_integrate_cos_sin_closed_forms():
    # test_integrate.cpp:56-89
    xs, ys = _tri((0, 0), (1, 0), (1, 1))
    p = _xmone_ymtwo()
    check(L.or_integrate(C.byref(p), xs, ys, 1),
          (4.0 * math.sin(1.) - 9.0 * math.cos(1.)) / 2.0 + 1.0)
    check(L.or_integrate(C.byref(p), xs, ys, 2),
          5.0 - (9.0 * math.sin(1.) + 4.0 * math.cos(1.)) / 2.0)


def test_integrate_1_cos_closed_form():
    # no reference KAT (print-only, test_integrate.cpp:91): pin by the closed
    # form of int_0^1 int_0^x (x-1)(y-2)/cos x dy dx, evaluated by scipy quad
    from scipy import integrate
    xs, ys = _tri((0, 0), (1, 0), (1, 1))
    p = _xmone_ymtwo()
    inner = lambda x: (x - 1.0) * (x * x / 2.0 - 2.0 * x) / math.cos(x)
    ref, _ = integrate.quad(inner, 0.0, 1.0, epsabs=1e-15, epsrel=1e-15)
    # GK15 on a unit interval with a pole-free integrand: ~1e-13 accurate
    check(L.or_integrate(C.byref(p), xs, ys, 3), ref, 1e-12)


SPHERE_AREA = [
    # test_integrate.cpp:146-274: (triangle, expected)
    (((0, 0), (0, math.pi / 8.0), (math.pi / 8.0, 0)), 0.0761204674887130),
    (((0, math.pi / 8.0), (0, math.pi / 4.0), (math.pi / 8, math.pi / 8.0)), 0.0761204674887130),
    (((0, math.pi / 8.0), (math.pi / 8, 0), (math.pi / 8, math.pi / 8.0)), 0.0741589649823731),
    (((math.pi / 8.0, 0), (math.pi / 4, 0), (math.pi / 8, math.pi / 8.0)), 0.0664933188536525),
    (((0, 0), (math.pi / 8.0, 0), (0, math.pi / 8.0)), 0.0761204674887130),
    (((math.pi / 8.0, 0), (math.pi / 4.0, 0), (math.pi / 8, math.pi / 8.0)), 0.0664933188536525),
    (((math.pi / 8.0, 0), (0, math.pi / 8), (math.pi / 8, math.pi / 8.0)), 0.0741589649823731),
]


@pytest.mark.parametrize("tri,expected", SPHERE_AREA)
def test_sphere_area(tri, expected):
    xs, ys = _tri(*tri)
    one = O.poly(0, 0, [1.0])
    check(L.or_integrate(C.byref(one), xs, ys, 1), expected)


def test_polynom_ops_print():
    # test_polynom.cpp:20-56 (Polynom ops and print strings)
    x = O.poly(1, 0, [0.0, 1.0])
    y = O.poly(0, 1, [0.0, 1.0])
    assert O.poly_print(x) == "+1.0 x^1 "
    assert O.poly_print(y) == "+1.0 y^1 "

    def lin(base, c, s):  # (base - c) * s
        k = O.poly_koef(base)
        k[0] -= c
        return O.poly(base.x_deg, base.y_deg, [v * s for v in k])

    def sub(a, b):
        xd, yd = max(a.x_deg, b.x_deg), max(a.y_deg, b.y_deg)
        def k(p, i, j):
            return 0.0 if i > p.x_deg or j > p.y_deg else p.koef[i * (p.y_deg + 1) + j]
        return O.poly(xd, yd, [k(a, i, j) - k(b, i, j) for i in range(xd + 1) for j in range(yd + 1)])

    assert O.poly_print(lin(x, 1, 2 - 5)) == "+3.0 -3.0 x^1 "
    assert O.poly_print(lin(y, 11, -9 - 5)) == "+154.0 -14.0 y^1 "
    p1 = sub(lin(x, 1, 2 - 5), lin(y, 11, -9 - 5))
    assert O.poly_print(p1) == "-151.0 +14.0 y^1 -3.0 x^1 "
    k1 = 1.0 / 15.0
    p1 = O.poly(p1.x_deg, p1.y_deg, [v * k1 for v in O.poly_koef(p1)])
    assert O.poly_print(p1) == "-10.1 +0.9 y^1 -0.2 x^1 "
    p2 = sub(lin(x, 2, 2 - 10), lin(y, -11, 10 + 5))
    assert O.poly_print(p2) == "-149.0 -15.0 y^1 -8.0 x^1 "
    pr = O.OrPoly()
    L.or_poly_mul(C.byref(p1), C.byref(p2), C.byref(pr))
    assert O.poly_print(pr) == "+1499.9 +11.9 y^1 -14.0 y^2 +110.3 x^1 -4.5 x^1 y^1 +1.6 x^2 "
    d = O.OrPoly()
    L.or_poly_diff(C.byref(p1), 0, C.byref(d))
    assert O.poly_print(d) == "-0.2 "


def test_p1_basis_is_nodal_and_partition_of_unity():
    rng = np.random.default_rng(3)
    for _ in range(50):
        pts = rng.uniform(-1.5, 4.7, size=(3, 2))
        xs, ys = _tri(*pts)
        basis = (O.OrPoly * 3)()
        L.or_basis_p1(xs, ys, basis)
        for k in range(3):
            for l in range(3):
                v = L.or_poly_apply(C.byref(basis[k]), pts[l][0], pts[l][1])
                assert abs(v - (1.0 if k == l else 0.0)) < 1e-9
        q = rng.uniform(-1, 4, size=2)
        s = sum(L.or_poly_apply(C.byref(basis[k]), q[0], q[1]) for k in range(3))
        assert abs(s - 1.0) < 1e-9


def test_cubature7_degree5_exact():
    # 7-point rule is exact for total degree <= 5 (quadrature.c:130-171):
    # int over (0,0),(1,0),(0,1) of x^a y^b = a! b! / (a+b+2)!
    for a in range(6):
        for b in range(6 - a):
            f = O.FXY(lambda x, y, d, a=a, b=b: x ** a * y ** b)
            got = L.or_cubature7(0, 0, 1, 0, 0, 1, f, None)
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(got - exact) < 1e-14, (a, b)


def test_cos_against_c_library():
    # the oracle's cos/sin is the correctly rounded one (pinned against
    # mpmath in test_cr_trig.py); the C library's is within 1 ulp of it and
    # equal almost everywhere on the chart latitude range
    xs = np.concatenate([np.linspace(-3.5, 3.5, 20001),
                         np.random.default_rng(0).uniform(-1.6, 1.6, 20000)])
    worst, differ = 0.0, 0
    for x in xs:
        for ours, ref in ((L.or_cos(float(x)), math.cos(float(x))),
                          (L.or_sin(float(x)), math.sin(float(x)))):
            if abs(ref) > 1e-3:
                worst = max(worst, abs(ours - ref) / math.ulp(ref))
            differ += ours != ref
    assert worst <= 1.0, worst
    assert differ <= 0.01 * 2 * len(xs), differ


@pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built (no reference tree)")
def test_quadrature_bitwise_vs_reference_build():
    ref = C.CDLL(REF_SO)
    ref.gauss_kronrod15.restype = C.c_double
    ref.gauss_kronrod15.argtypes = [C.c_double, C.c_double, O.FX, C.c_void_p]
    ref.cubature7.restype = C.c_double
    ref.cubature7.argtypes = [C.c_double] * 6 + [O.FXY, C.c_void_p]
    rng = np.random.default_rng(7)
    fs = [O.FX(lambda x, d: math.cos(x) * x ** 3 - 0.3 * x),
          O.FX(lambda x, d: 1.0 / math.cos(x) + x * x),
          O.FX(lambda x, d: math.exp(x))]
    for _ in range(200):
        a, b = sorted(rng.uniform(-1.5, 1.5, size=2))
        for f in fs:
            assert L.or_gauss_kronrod15(a, b, f, None) == ref.gauss_kronrod15(a, b, f, None)
    g = O.FXY(lambda x, y, d: math.cos(x) * (1 + x * y - y * y))
    for _ in range(200):
        p = rng.uniform(-1.5, 4.7, size=6)
        assert L.or_cubature7(*p, g, None) == ref.cubature7(*p, g, None)
