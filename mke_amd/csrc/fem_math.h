// Element arithmetic of the sphere FEM path, shared by the gfx950 assembly
// kernels and the host-side boundary-matrix code.  Every expression keeps the
// reference's operation order; translation units that include this header
// are compiled with -ffp-contract=off so that results are bit-identical on
// host and device.
//
//   cos             : correctly rounded (cr_trig.h), since the mass entries
//                     cancel ~1e7-fold and a 1-ulp change of cos moves them
//                     ~1e-10 relative; the reference calls the C library's
//                     cos (src/util.cpp:194-215).
//   ipow            : repeated multiplication (the reference's ipow comes from
//                     its solver library, absent from its tree).
//   GK15            : src/quadrature.c:49-128 (15-digit node/weight tables)
//   trapezoids      : src/util.cpp:177-251
//   integrate1      : src/polynom.cpp:188-315
//   P1 basis        : src/mesh.cpp:563-607
//   element values  : src/slaplace.cpp:75-197, examples/barvortex_impl.h:57-80
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "cr_trig.h"

#define PHG_HD __host__ __device__ inline

namespace phg {

// ---------------------------------------------------------------- cos ------
// The correctly rounded cos of cr_trig.h (Ziv's method): the value every
// correctly rounding C library returns for the reference's cos call.
PHG_HD double fem_cos(double x) { return cr_cos(x); }

// ----------------------------------------------------------------- GK15 ----
// Kronrod nodes/weights of src/quadrature.c:64-96, in the reference's order.
#define PHG_K15_NODES                                                          \
    {+0.949107912342759, -0.949107912342759, +0.741531185599394,               \
     -0.741531185599394, +0.405845151377397, -0.405845151377397,               \
     0.000000000000000,  +0.991455371120813, -0.991455371120813,               \
     +0.864864423359769, -0.864864423359769, +0.586087235467691,               \
     -0.586087235467691, +0.207784955007898, -0.207784955007898}
#define PHG_K15_WEIGHTS                                                        \
    {0.063092092629979, 0.063092092629979, 0.140653259715525,                  \
     0.140653259715525, 0.190350578064785, 0.190350578064785,                  \
     0.209482141084728, 0.022935322010529, 0.022935322010529,                  \
     0.104790010322250, 0.104790010322250, 0.169004726639267,                  \
     0.169004726639267, 0.204432940075298, 0.204432940075298}

// --------------------------------------------------------------- geometry --
// A triangle in one chart, sorted as integrate1 sorts it (x1 <= x3 <= x2),
// with the three bounding lines y = k x + b.
struct TrapGeom {
    double x1, x2, x3, y3;
    double k1, b1, k2, b2, k3, b3;
    bool has1, has2;  // |x1 - x3| > 1e-15, |x3 - x2| > 1e-15
    bool positive;    // sign rule of polynom.cpp:311-314
};

PHG_HD TrapGeom trap_geom(double x1, double x2, double x3, double y1, double y2,
                          double y3)
{
    double t;
    if (x1 <= x2 && x2 <= x3) {
        t = x2; x2 = x3; x3 = t;
        t = y2; y2 = y3; y3 = t;
    } else if (x2 <= x1 && x1 <= x3) {
        t = x1; x1 = x2; x2 = t;
        t = x2; x2 = x3; x3 = t;
        t = y1; y1 = y2; y2 = t;
        t = y2; y2 = y3; y3 = t;
    } else if (x2 <= x3 && x3 <= x1) {
        t = x1; x1 = x2; x2 = t;
        t = y1; y1 = y2; y2 = t;
    } else if (x3 <= x1 && x1 <= x2) {
        t = x1; x1 = x3; x3 = t;
        t = y1; y1 = y3; y3 = t;
    } else if (x3 <= x2 && x2 <= x1) {
        t = x1; x1 = x3; x3 = t;
        t = x2; x2 = x3; x3 = t;
        t = y1; y1 = y3; y3 = t;
        t = y2; y2 = y3; y3 = t;
    }
    TrapGeom g;
    g.x1 = x1; g.x2 = x2; g.x3 = x3; g.y3 = y3;
    g.k1 = (y2 - y1) / (x2 - x1);
    g.b1 = (y1 * x2 - x1 * y2) / (x2 - x1);
    g.has1 = fabs(x1 - x3) > 1e-15;
    g.has2 = fabs(x3 - x2) > 1e-15;
    g.k3 = g.b3 = g.k2 = g.b2 = 0.0;
    if (g.has1) {
        g.k3 = (y3 - y1) / (x3 - x1);
        g.b3 = (y1 * x3 - x1 * y3) / (x3 - x1);
    }
    if (g.has2) {
        g.k2 = (y2 - y3) / (x2 - x3);
        g.b2 = (y3 * x2 - x3 * y2) / (x2 - x3);
    }
    if (y3 >= y1 && y3 >= y2) g.positive = true;
    else if (y3 <= y1 && y3 <= y2) g.positive = false;
    else if (y3 >= g.k1 * x3 + g.b1) g.positive = true;
    else g.positive = false;
    return g;
}

// All trapezoid integrals one P1 element needs, for one trapezoid
// [xa, xb] between the lines (kl, bl) (lower, "k1") and (ku, bu) (upper):
//   icos[k][n] = trapezoid_integral_cos(k, n, kl, bl, ku, bu, xa, xb), k,n<=2
//   isec[k]    = trapezoid_integral_1_cos(k, 0, kl, bl, ku, bu, xa, xb)
// Each value equals the reference's separate GK15 call bit for bit: the node
// abscissae, cos(node) and every per-monomial expression are evaluated the
// same way, only shared between the monomials.
PHG_HD void trap_tables(double kl, double bl, double ku, double bu, double xa,
                        double xb, double icos[3][3], double isec[3])
{
    const double nodes[15] = PHG_K15_NODES;
    const double weights[15] = PHG_K15_WEIGHTS;
    const double p1 = 0.5 * (xb - xa);
    const double p2 = 0.5 * (xb + xa);
    double ac[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    double as[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 15; ++i) {
        const double x = p1 * nodes[i] + p2;
        const double c = fem_cos(x);
        const double up = ku * x + bu;
        const double lo = kl * x + bl;
        // ipow(v, 1..3) by repeated multiplication
        const double u1 = 1.0 * up, u2 = u1 * up, u3 = u2 * up;
        const double l1 = 1.0 * lo, l2 = l1 * lo, l3 = l2 * lo;
        double pt[3];
        pt[0] = (u1 - l1) / 1.0;
        pt[1] = (u2 - l2) / 2.0;
        pt[2] = (u3 - l3) / 3.0;
        const double xp[3] = {1.0, 1.0 * x, (1.0 * x) * x};
        const double w = weights[i];
        for (int k = 0; k < 3; ++k) {
            for (int n = 0; n < 3; ++n) ac[k][n] += w * (xp[k] * pt[n] * c);
            as[k] += w * (xp[k] * pt[0] / c);
        }
    }
    for (int k = 0; k < 3; ++k) {
        for (int n = 0; n < 3; ++n) icos[k][n] = ac[k][n] * p1;
        isec[k] = as[k] * p1;
    }
}

// Integral tables of one triangle in one chart.
struct ElemTables {
    double c1[3][3], c2[3][3];  // cos weight, trapezoids 1 and 2
    double s1[3], s2[3];        // 1/cos weight, n = 0
    bool has1, has2, positive;
};

PHG_HD void elem_tables(const TrapGeom& g, ElemTables& t)
{
    t.has1 = g.has1;
    t.has2 = g.has2;
    t.positive = g.positive;
    if (g.has1) trap_tables(g.k1, g.b1, g.k3, g.b3, g.x1, g.x3, t.c1, t.s1);
    if (g.has2) trap_tables(g.k1, g.b1, g.k2, g.b2, g.x3, g.x2, t.c2, t.s2);
}

// ---------------------------------------------------------------- P1 basis --
// Polynom(1,1) coefficients [k00, k01, k10, k11] (koef[i*2+j] = x^i y^j).
struct P1Poly {
    double k[4];
};

// L(p1, p2) = (X - p1.x)(p2.y - p1.y) - (Y - p1.y)(p2.x - p1.x), mesh.cpp:578-580
PHG_HD P1Poly basis_line(double p1x, double p1y, double p2x, double p2y)
{
    double dy = p2y - p1y;
    double a0 = (0.0 - p1x) * dy, a1 = 1.0 * dy;  // (X - p1.x) * dy: [k00, k10]
    double dx = p2x - p1x;
    double b0 = (0.0 - p1y) * dx, b1 = 1.0 * dx;  // (Y - p1.y) * dx: [k00, k01]
    P1Poly r;
    r.k[0] = a0 - b0;    // k(0,0)
    r.k[1] = 0.0 - b1;   // k(0,1)
    r.k[2] = a1 - 0.0;   // k(1,0)
    r.k[3] = 0.0 - 0.0;  // k(1,1)
    return r;
}

PHG_HD double p1_apply(const P1Poly& p, double x, double y)
{
    double r = 0.0;  // polynom.cpp:63-89 with ipow(., 0) = 1, ipow(., 1) = 1*v
    const double x1 = 1.0 * x, y1 = 1.0 * y;
    r += p.k[0] * 1.0 * 1.0;
    r += p.k[1] * 1.0 * y1;
    r += p.k[2] * x1 * 1.0;
    r += p.k[3] * x1 * y1;
    return r;
}

PHG_HD void p1_basis(const double xs[3], const double ys[3], P1Poly phi[3])
{
    phi[0] = basis_line(xs[1], ys[1], xs[2], ys[2]);
    phi[1] = basis_line(xs[0], ys[0], xs[2], ys[2]);
    phi[2] = basis_line(xs[0], ys[0], xs[1], ys[1]);
    for (int i = 0; i < 3; ++i) {
        const double k1 = 1.0 / p1_apply(phi[i], xs[i], ys[i]);
        for (int c = 0; c < 4; ++c) phi[i].k[c] *= k1;
    }
}

// ------------------------------------------------------------ element values --
// integrate1's final combination (polynom.cpp:267-314) for the 9-coefficient
// product of two P1 polynomials, cos weight.
PHG_HD double integrate_mass(const ElemTables& t, const P1Poly& a, const P1Poly& b)
{
    double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // operator* (polynom.cpp:667-696)
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int m = 0; m < 2; ++m) q[(i + k) * 3 + (j + m)] += a.k[i * 2 + j] * b.k[k * 2 + m];
    double int1 = 0.0, int2 = 0.0;
    if (t.has1)
        for (int k = 0; k < 3; ++k)
            for (int n = 0; n < 3; ++n) int1 += q[k * 3 + n] * t.c1[k][n];
    if (t.has2)
        for (int k = 0; k < 3; ++k)
            for (int n = 0; n < 3; ++n) int2 += q[k * 3 + n] * t.c2[k][n];
    return t.positive ? int1 + int2 : -int1 - int2;
}

// slaplace(a, b) (slaplace.cpp:75-105): -int cos d_x a d_x b - int sec d_y a d_y b
PHG_HD double integrate_stiff(const ElemTables& t, const P1Poly& a, const P1Poly& b)
{
    // diff(., 0) of Polynom(1,1) is Polynom(0,1) [k10, k11]; product (0,2)
    const double da0 = 1.0 * a.k[2], da1 = 1.0 * a.k[3];
    const double db0 = 1.0 * b.k[2], db1 = 1.0 * b.k[3];
    double q[3] = {0.0, 0.0, 0.0};
    q[0] += da0 * db0;
    q[1] += da0 * db1;
    q[1] += da1 * db0;
    q[2] += da1 * db1;
    double int1 = 0.0, int2 = 0.0;
    if (t.has1)
        for (int n = 0; n < 3; ++n) int1 += q[n] * t.c1[0][n];
    if (t.has2)
        for (int n = 0; n < 3; ++n) int2 += q[n] * t.c2[0][n];
    const double l1 = -(t.positive ? int1 + int2 : -int1 - int2);
    // diff(., 1) of Polynom(1,1) is Polynom(1,0) [k01, k11]; product (2,0)
    const double ea0 = 1.0 * a.k[1], ea1 = 1.0 * a.k[3];
    const double eb0 = 1.0 * b.k[1], eb1 = 1.0 * b.k[3];
    double s[3] = {0.0, 0.0, 0.0};
    s[0] += ea0 * eb0;
    s[1] += ea0 * eb1;
    s[1] += ea1 * eb0;
    s[2] += ea1 * eb1;
    double j1 = 0.0, j2 = 0.0;
    if (t.has1)
        for (int k = 0; k < 3; ++k) j1 += s[k] * t.s1[k];
    if (t.has2)
        for (int k = 0; k < 3; ++k) j2 += s[k] * t.s2[k];
    const double l2 = -(t.positive ? j1 + j2 : -j1 - j2);
    return l1 + l2;
}

// ------------------------------------------------------- 7-point triangle rule --
// src/quadrature.c:130-171 (degree-5 rule, 16-digit barycentric tables).
#define PHG_C7_A1 0.7974269853530880
#define PHG_C7_B1 0.1012865073234560
#define PHG_C7_A2 0.0597158717897742
#define PHG_C7_B2 0.4701420641051129

// Polynom::apply (polynom.cpp:63-89) of an (XD, YD) polynomial, koef[i*(YD+1)+j].
template <int XD, int YD>
PHG_HD double poly_apply(const double* q, double x, double y)
{
    double xi[XD + 1], yj[YD + 1];
    double r = 1.0;
    for (int i = 0; i <= XD; ++i) {
        xi[i] = r;
        r *= x;
    }
    r = 1.0;
    for (int j = 0; j <= YD; ++j) {
        yj[j] = r;
        r *= y;
    }
    double s = 0.0;
    for (int i = 0; i <= XD; ++i)
        for (int j = 0; j <= YD; ++j) s += q[i * (YD + 1) + j] * xi[i] * yj[j];
    return s;
}

// cubature7 of q(x, y), or of cos(x) q(x, y) when wcos (the integrand
// callbacks of src/sjacobian.cpp:50-58).
template <int XD, int YD>
PHG_HD double cubature7_poly(const double* q, const double xs[3], const double ys[3],
                             bool wcos)
{
    const double w7[7] = {0.225,
                          0.1259391805448265,
                          0.1259391805448265,
                          0.1259391805448265,
                          0.1323941527885112,
                          0.1323941527885112,
                          0.1323941527885112};
    const double c7[7][3] = {{1. / 3., 1. / 3., 1. / 3.},
                             {PHG_C7_A1, PHG_C7_B1, PHG_C7_B1},
                             {PHG_C7_B1, PHG_C7_A1, PHG_C7_B1},
                             {PHG_C7_B1, PHG_C7_B1, PHG_C7_A1},
                             {PHG_C7_A2, PHG_C7_B2, PHG_C7_B2},
                             {PHG_C7_B2, PHG_C7_A2, PHG_C7_B2},
                             {PHG_C7_B2, PHG_C7_B2, PHG_C7_A2}};
    double s = 0.0;
    for (int i = 0; i < 7; ++i) {
        const double u = c7[i][0] * xs[0] + c7[i][1] * xs[1] + c7[i][2] * xs[2];
        const double v = c7[i][0] * ys[0] + c7[i][1] * ys[1] + c7[i][2] * ys[2];
        double f = poly_apply<XD, YD>(q, u, v);
        if (wcos) f = fem_cos(u) * f;
        s += w7[i] * f;
    }
    s *= 0.5 * fabs(((xs[0] - xs[2]) * (ys[1] - ys[2]) - (ys[0] - ys[2]) * (xs[1] - xs[2])));
    return s;
}

enum ElemKind : int {
    ELEM_MASS = 0,
    ELEM_LAPLACE = 1,
    ELEM_HELMHOLTZ = 2,
    // SphereJacobian matrices (src/sjacobian.cpp:60-134), 7-point rule
    ELEM_JAC_MASS = 3,   // id_cb: cos x phi_i phi_j
    ELEM_DIFF1 = 4,      // diff_1_rp: d_x phi_j phi_i
    ELEM_DIFF2 = 5,      // diff_2_rp: d_y phi_j phi_i
    ELEM_DIFF1_COS = 6,  // diff_1_cos_rp: cos x d_x phi_j phi_i
    ELEM_DIFF2_COS = 7,  // diff_2_cos_rp: cos x d_y phi_j phi_i
    ELEM_NKINDS = 8
};

// Kinds integrated by integrate1 over trapezoids need the GK15 tables of
// elem_tables; the Jacobian kinds use only the vertex coordinates.
PHG_HD bool elem_needs_tables(int kind) { return kind < ELEM_JAC_MASS; }

PHG_HD double elem_cubature(int kind, const double xs[3], const double ys[3],
                            const P1Poly& phi_i, const P1Poly& phi_j)
{
    if (kind == ELEM_JAC_MASS) {
        double q[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // phi_i * phi_j, (2, 2)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k)
                    for (int m = 0; m < 2; ++m)
                        q[(i + k) * 3 + (j + m)] += phi_i.k[i * 2 + j] * phi_j.k[k * 2 + m];
        return cubature7_poly<2, 2>(q, xs, ys, true);
    }
    const bool wcos = kind >= ELEM_DIFF1_COS;
    double q[6] = {0, 0, 0, 0, 0, 0};
    double v = 1;  // the commented-out weight u[i] of the callbacks
    if (kind == ELEM_DIFF1 || kind == ELEM_DIFF1_COS) {
        // diff(phi_j, 0) = Polynom(0, 1) [k10, k11]; times phi_i -> (1, 2)
        const double d[2] = {1.0 * phi_j.k[2], 1.0 * phi_j.k[3]};
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int m = 0; m < 2; ++m) q[k * 3 + (j + m)] += d[j] * phi_i.k[k * 2 + m];
        return v * cubature7_poly<1, 2>(q, xs, ys, wcos);
    }
    // diff(phi_j, 1) = Polynom(1, 0) [k01, k11]; times phi_i -> (2, 1)
    const double d[2] = {1.0 * phi_j.k[1], 1.0 * phi_j.k[3]};
    for (int i = 0; i < 2; ++i)
        for (int k = 0; k < 2; ++k)
            for (int m = 0; m < 2; ++m) q[(i + k) * 2 + m] += d[i] * phi_i.k[k * 2 + m];
    return v * cubature7_poly<2, 1>(q, xs, ys, wcos);
}

// Value of one (phi_i, phi_j) pair over one triangle, as the reference's
// callbacks compute it: id_cb, slaplace_integrate_cb, bv integrate_cb and the
// SphereJacobian callbacks.  t is read only when elem_needs_tables(kind).
PHG_HD double elem_value(int kind, const ElemTables& t, const double xs[3],
                         const double ys[3], const P1Poly& phi_i, const P1Poly& phi_j,
                         double c_mass, double c_lap)
{
    if (kind == ELEM_MASS) return integrate_mass(t, phi_i, phi_j);
    if (kind == ELEM_LAPLACE) return integrate_stiff(t, phi_j, phi_i);
    if (kind >= ELEM_JAC_MASS) return elem_cubature(kind, xs, ys, phi_i, phi_j);
    double pt1 = integrate_mass(t, phi_i, phi_j);
    pt1 *= c_mass;
    double pt2 = integrate_stiff(t, phi_j, phi_i);
    pt2 *= c_lap;
    return pt1 + pt2;
}

}  // namespace phg
