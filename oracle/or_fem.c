/*
 * CPU oracle: quadrature, trapezoid integrals, Polynom arithmetic, P1 basis
 * and row-wise FE assembly.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every routine keeps the reference's floating-point operation order; build
 * with -ffp-contract=off so that no multiply-add is fused behind its back.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ------------------------------------------------------------------------ */
/* cos / sin: the correctly rounded routine the device kernels use           */
/* (mke_amd/csrc/cr_trig.h: Ziv's strategy, double-double fast path and a    */
/* 256-bit fixed-point slow path).  The reference calls the C library's cos  */
/* (src/util.cpp:194-215); the correctly rounded value is what any correct   */
/* library must return wherever it rounds correctly.                         */
/* ------------------------------------------------------------------------ */
#include "cr_trig.h"

static int g_libm_cos = 0;
void or_use_libm_cos(int on) { g_libm_cos = on; }

/* optional record of the distinct cos arguments (tests collect the arguments
 * an assembly evaluates): an open-addressing set of bit patterns, cap a power
 * of two, filled to at most 3/4 (further arguments are counted as lost) */
static uint64_t *g_rec = NULL;
static long g_rec_cap = 0, g_rec_n = 0, g_rec_lost = 0;
#define OR_REC_EMPTY 0xffffffffffffffffull
void or_record_cos_args(double *buf, long cap)
{
    g_rec = (uint64_t *)buf;
    g_rec_cap = cap;
    g_rec_n = 0;
    g_rec_lost = 0;
    for (long i = 0; g_rec && i < cap; ++i) g_rec[i] = OR_REC_EMPTY;
}
long or_recorded_cos_args(void) { return g_rec_lost ? -g_rec_lost : g_rec_n; }

static void record_arg(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    uint64_t h = (u ^ (u >> 29)) * 0x9e3779b97f4a7c15ull;
    long i = (long)((h >> 17) & (uint64_t)(g_rec_cap - 1));
    for (;;) {
        uint64_t cur = __atomic_load_n(&g_rec[i], __ATOMIC_RELAXED);
        if (cur == u) return;
        if (cur == OR_REC_EMPTY) {
            if (__atomic_load_n(&g_rec_n, __ATOMIC_RELAXED) >= g_rec_cap / 4 * 3) {
                __atomic_fetch_add(&g_rec_lost, 1, __ATOMIC_RELAXED);
                return;
            }
            if (__atomic_compare_exchange_n(&g_rec[i], &cur, u, 0, __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED)) {
                __atomic_fetch_add(&g_rec_n, 1, __ATOMIC_RELAXED);
                return;
            }
            if (cur == u) return;
        }
        i = (i + 1) & (g_rec_cap - 1);
    }
}

/* The reference calls cos once per GK15 node for every monomial of every
 * element integral, so the same argument comes back many times in a row; a
 * small per-thread cache of (argument bits, value) pairs returns the value
 * already computed (the same bits: the routine is a function of x). */
#define OR_COS_CACHE 256
static __thread uint64_t t_cache_key[OR_COS_CACHE];
static __thread double t_cache_val[OR_COS_CACHE];
static __thread int t_cache_init = 0;

double or_cos(double x)
{
    if (g_rec) record_arg(x);
    if (g_libm_cos) return cos(x);
    uint64_t u;
    memcpy(&u, &x, 8);
    if (!t_cache_init) {
        for (int i = 0; i < OR_COS_CACHE; ++i) t_cache_key[i] = OR_REC_EMPTY;
        t_cache_init = 1;
    }
    const int slot = (int)(((u ^ (u >> 31)) * 0x9e3779b97f4a7c15ull) >> 56);
    if (t_cache_key[slot] == u) return t_cache_val[slot];
    const double c = cr_cos(x);
    t_cache_key[slot] = u;
    t_cache_val[slot] = c;
    return c;
}

double or_sin(double x)
{
    if (g_libm_cos) return sin(x);
    return cr_sin(x);
}

/* test hooks: the rounding test at another bound, the slow path alone, and
 * the double-double value of the fast path */
double or_cr_cos_bound(double x, double err) { return cr_cos_bound(x, err); }
double or_cr_sin_bound(double x, double err) { return cr_sin_bound(x, err); }
double or_cr_slow(double x, int odd) { return crt_slow(x, odd); }
void or_cr_fast_dd(double x, int odd, double *hi, double *lo)
{
    int q;
    const crt_dd r = crt_reduce(x, &q);
    const int use_sin = odd ? !(q & 1) : (q & 1);
    const crt_dd v = use_sin ? crt_poly_sin(r) : crt_poly_cos(r);
    const int neg = odd ? (q == 2 || q == 3) : (q == 1 || q == 2);
    *hi = neg ? -v.hi : v.hi;
    *lo = neg ? -v.lo : v.lo;
}

double or_ipow(double x, int p)
{
    double r = 1.0;
    for (int i = 0; i < p; ++i) r *= x;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Quadrature: src/quadrature.c:49-171.  The node/weight tables are data    */
/* that must match the reference digit for digit (15-digit truncations).    */
/* ------------------------------------------------------------------------ */
static const double K15N[15][2] = {
    {+0.949107912342759, 0.063092092629979},
    {-0.949107912342759, 0.063092092629979},
    {+0.741531185599394, 0.140653259715525},
    {-0.741531185599394, 0.140653259715525},
    {+0.405845151377397, 0.190350578064785},
    {-0.405845151377397, 0.190350578064785},
    { 0.000000000000000, 0.209482141084728},
    {+0.991455371120813, 0.022935322010529},
    {-0.991455371120813, 0.022935322010529},
    {+0.864864423359769, 0.104790010322250},
    {-0.864864423359769, 0.104790010322250},
    {+0.586087235467691, 0.169004726639267},
    {-0.586087235467691, 0.169004726639267},
    {+0.207784955007898, 0.204432940075298},
    {-0.207784955007898, 0.204432940075298},
};

/* quadrature.c:100-128: nodes 0..6 then 7..14, k15 *= p1 at the end */
double or_gauss_kronrod15(double a, double b, or_fx_t f, void *d)
{
    double k15 = 0.0;
    double p1 = 0.5 * (b - a);
    double p2 = 0.5 * (b + a);
    for (int i = 0; i < 15; ++i) k15 += K15N[i][1] * f(p1 * K15N[i][0] + p2, d);
    return k15 * p1;
}

#define CW_A1 0.7974269853530880
#define CW_B1 0.1012865073234560
#define CW_A2 0.0597158717897742
#define CW_B2 0.4701420641051129
static const double W7[7] = {0.225, 0.1259391805448265, 0.1259391805448265,
                             0.1259391805448265, 0.1323941527885112,
                             0.1323941527885112, 0.1323941527885112};
static const double C7[7][3] = {
    {1. / 3., 1. / 3., 1. / 3.}, {CW_A1, CW_B1, CW_B1}, {CW_B1, CW_A1, CW_B1},
    {CW_B1, CW_B1, CW_A1},       {CW_A2, CW_B2, CW_B2}, {CW_B2, CW_A2, CW_B2},
    {CW_B2, CW_B2, CW_A2}};

/* quadrature.c:151-171 */
double or_cubature7(double x1, double y1, double x2, double y2, double x3,
                    double y3, or_fxy_t f, void *d)
{
    double s = 0.0;
    for (int i = 0; i < 7; ++i) {
        double a = C7[i][0], b = C7[i][1], c = C7[i][2];
        double u = a * x1 + b * x2 + c * x3;
        double v = a * y1 + b * y2 + c * y3;
        s += W7[i] * f(u, v, d);
    }
    s *= 0.5 * fabs(((x1 - x3) * (y2 - y3) - (y1 - y3) * (x2 - x3)));
    return s;
}

/* ------------------------------------------------------------------------ */
/* Trapezoid integrals: src/util.cpp:97-251                                 */
/* ------------------------------------------------------------------------ */
static double binom(int n, int k)  /* cnk_6.h rows, k <= n: exact integers */
{
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
    return r;
}

/* util.cpp:106-160 */
static double ssum(double x, int k, int n, double k1, double b1, double k2,
                   double b2)
{
    double sum1 = 0, sum2 = 0;
    if (k2 != 0 && b2 != 0) {
        for (int i = 0; i <= n + 1; i++)
            sum1 += binom(n + 1, i) * or_ipow(k2, n + 1 - i) *
                    or_ipow(x, n + k + 2 - i) * or_ipow(b2, i) /
                    (double)(n + k + 2 - i);
    } else if (k2 != 0 && b2 == 0) {
        sum1 = or_ipow(k2, n + 1) * or_ipow(x, k + n + 2) / (double)(k + n + 2);
    } else if (k2 == 0 && b2 != 0) {
        sum1 = or_ipow(b2, n + 1) * or_ipow(x, k + 1) / (double)(k + 1);
    } else {
        sum1 = 0;
    }
    if (k1 != 0 && b1 != 0) {
        for (int i = 0; i <= n + 1; i++)
            sum2 += binom(n + 1, i) * or_ipow(k1, n + 1 - i) *
                    or_ipow(x, n + k + 2 - i) * or_ipow(b1, i) /
                    (double)(n + k + 2 - i);
    } else if (k1 != 0 && b1 == 0) {
        sum2 = or_ipow(k1, n + 1) * or_ipow(x, k + n + 2) / (double)(k + n + 2);
    } else if (k1 == 0 && b1 != 0) {
        sum2 = or_ipow(b1, n + 1) * or_ipow(x, k + 1) / (double)(k + 1);
    } else {
        sum2 = 0;
    }
    return sum1 - sum2;
}

/* util.cpp:166-175 */
double or_trapezoid_integral(int k, int n, double k1, double b1, double k2,
                             double b2, double x1, double x3)
{
    return (ssum(x3, k, n, k1, b1, k2, b2) - ssum(x1, k, n, k1, b1, k2, b2)) /
           (double)(n + 1);
}

typedef struct { int k, n; double k1, b1, k2, b2; } fcos_data;

/* util.cpp:194-215: x^k [(k2x+b2)^{n+1} - (k1x+b1)^{n+1}]/(n+1) w(x) */
static double pt1_of(double x, const fcos_data *d)
{
    return (or_ipow((d->k2 * x + d->b2), d->n + 1) -
            or_ipow((d->k1 * x + d->b1), d->n + 1)) / (double)(d->n + 1);
}
static double f_cos(double x, void *v)
{
    const fcos_data *d = (const fcos_data *)v;
    double pt1 = pt1_of(x, d);
    return or_ipow(x, d->k) * pt1 * or_cos(x);
}
static double f_sin(double x, void *v)
{
    const fcos_data *d = (const fcos_data *)v;
    double pt1 = pt1_of(x, d);
    return or_ipow(x, d->k) * pt1 * or_sin(x);
}
static double f_1_cos(double x, void *v)
{
    const fcos_data *d = (const fcos_data *)v;
    double pt1 = pt1_of(x, d);
    return or_ipow(x, d->k) * pt1 / or_cos(x);
}

double or_trapezoid_integral_cos(int k, int n, double k1, double b1, double k2,
                                 double b2, double x1, double x3)
{
    fcos_data d = {k, n, k1, b1, k2, b2};
    return or_gauss_kronrod15(x1, x3, f_cos, &d);
}
double or_trapezoid_integral_sin(int k, int n, double k1, double b1, double k2,
                                 double b2, double x1, double x3)
{
    fcos_data d = {k, n, k1, b1, k2, b2};
    return or_gauss_kronrod15(x1, x3, f_sin, &d);
}
double or_trapezoid_integral_1_cos(int k, int n, double k1, double b1,
                                   double k2, double b2, double x1, double x3)
{
    fcos_data d = {k, n, k1, b1, k2, b2};
    return or_gauss_kronrod15(x1, x3, f_1_cos, &d);
}

/* ------------------------------------------------------------------------ */
/* Polynom: src/polynom.h:81-300, src/polynom.cpp:63-184,667-729             */
/* ------------------------------------------------------------------------ */
static void poly_init(or_poly *p, int xd, int yd)
{
    p->x_deg = xd;
    p->y_deg = yd;
    memset(p->koef, 0, sizeof(p->koef));
}
static double poly_k(const or_poly *p, int i, int j)  /* polynom.h:127-131 */
{
    if (i > p->x_deg || j > p->y_deg) return 0;
    return p->koef[i * (p->y_deg + 1) + j];
}

double or_poly_apply(const or_poly *a, double x, double y) /* polynom.cpp:63-89 */
{
    double xi[8], yj[8], r = 0.0;
    for (int i = 0; i <= a->x_deg; i++) xi[i] = or_ipow(x, i);
    for (int j = 0; j <= a->y_deg; j++) yj[j] = or_ipow(y, j);
    for (int i = 0; i <= a->x_deg; i++)
        for (int j = 0; j <= a->y_deg; j++)
            r += a->koef[i * (a->y_deg + 1) + j] * xi[i] * yj[j];
    return r;
}

void or_poly_diff(const or_poly *p, int d, or_poly *r) /* polynom.cpp:147-184 */
{
    int xd = p->x_deg, yd = p->y_deg, deg = yd;
    if (d == 0) {
        xd -= 1;
        poly_init(r, xd, yd);
        for (int j = 0; j <= yd; ++j)
            for (int i = 0; i <= xd; ++i)
                r->koef[i * (yd + 1) + j] =
                    (double)(i + 1) * p->koef[(i + 1) * (deg + 1) + j];
    } else {
        yd -= 1;
        poly_init(r, xd, yd);
        for (int i = 0; i <= xd; ++i)
            for (int j = 0; j <= yd; ++j)
                r->koef[i * (yd + 1) + j] =
                    (double)(j + 1) * p->koef[i * (deg + 1) + j + 1];
    }
}

void or_poly_mul(const or_poly *p1, const or_poly *p2, or_poly *p) /* :667-696 */
{
    poly_init(p, p1->x_deg + p2->x_deg, p1->y_deg + p2->y_deg);
    int deg = p->y_deg;
    for (int i = 0; i <= p1->x_deg; ++i)
        for (int j = 0; j <= p1->y_deg; ++j)
            for (int k = 0; k <= p2->x_deg; ++k)
                for (int m = 0; m <= p2->y_deg; ++m)
                    p->koef[(i + k) * (deg + 1) + (j + m)] +=
                        p1->koef[i * (p1->y_deg + 1) + j] *
                        p2->koef[k * (p2->y_deg + 1) + m];
}

static void poly_sub(const or_poly *p1, const or_poly *p2, or_poly *r) /* :698-713 */
{
    int xd = p1->x_deg > p2->x_deg ? p1->x_deg : p2->x_deg;
    int yd = p1->y_deg > p2->y_deg ? p1->y_deg : p2->y_deg;
    poly_init(r, xd, yd);
    for (int i = 0; i <= xd; ++i)
        for (int j = 0; j <= yd; ++j)
            r->koef[i * (yd + 1) + j] = poly_k(p1, i, j) - poly_k(p2, i, j);
}

/* mesh.cpp:578-580: L(p1, p2) = (X - p1.x)(p2.y - p1.y) - (Y - p1.y)(p2.x - p1.x) */
static void basis_line(double p1x, double p1y, double p2x, double p2y,
                       or_poly *out)
{
    or_poly a, b;
    poly_init(&a, 1, 0);               /* P2X */
    a.koef[1] = 1.0;
    a.koef[0] -= p1x;                  /* polynom.h:266-271 */
    double dy = p2y - p1y;
    a.koef[0] *= dy;                   /* polynom.h:281-290 */
    a.koef[1] *= dy;
    poly_init(&b, 0, 1);               /* P2Y */
    b.koef[1] = 1.0;
    b.koef[0] -= p1y;
    double dx = p2x - p1x;
    b.koef[0] *= dx;
    b.koef[1] *= dx;
    poly_sub(&a, &b, out);
}

/* mesh.cpp:582-607 (prepare_basis1) */
void or_basis_p1(const double xs[3], const double ys[3], or_poly out[3])
{
    basis_line(xs[1], ys[1], xs[2], ys[2], &out[0]);
    basis_line(xs[0], ys[0], xs[2], ys[2], &out[1]);
    basis_line(xs[0], ys[0], xs[1], ys[1], &out[2]);
    for (int i = 0; i < 3; ++i) {
        double k = or_poly_apply(&out[i], xs[i], ys[i]);
        double k1 = 1.0 / k;           /* polynom.h:136-144 */
        int sz = (out[i].x_deg + 1) * (out[i].y_deg + 1);
        for (int c = 0; c < sz; ++c) out[i].koef[c] *= k1;
    }
}

/* ------------------------------------------------------------------------ */
/* integrate1: src/polynom.cpp:188-315                                        */
/* ------------------------------------------------------------------------ */
typedef double (*t_int)(int, int, double, double, double, double, double,
                        double);

static double integrate1(const or_poly *p, const double xs[3],
                         const double ys[3], t_int trapezoid)
{
    double k1, b1, k2, b2, k3, b3, t;
    double x1 = xs[0], x2 = xs[1], x3 = xs[2];
    double y1 = ys[0], y2 = ys[1], y3 = ys[2];
    double int1 = 0.0, int2 = 0.0;

    /* polynom.cpp:205-266: afterwards x1 <= x3 <= x2 */
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

    k1 = (y2 - y1) / (x2 - x1);
    b1 = (y1 * x2 - x1 * y2) / (x2 - x1);

    if (fabs(x1 - x3) > 1e-15) {
        k3 = (y3 - y1) / (x3 - x1);
        b3 = (y1 * x3 - x1 * y3) / (x3 - x1);
        for (int k = 0; k <= p->x_deg; ++k)
            for (int n = 0; n <= p->y_deg; ++n)
                int1 += p->koef[k * (p->y_deg + 1) + n] *
                        trapezoid(k, n, k1, b1, k3, b3, x1, x3);
    } else {
        int1 = 0;
    }

    if (fabs(x3 - x2) > 1e-15) {
        k2 = (y2 - y3) / (x2 - x3);
        b2 = (y3 * x2 - x3 * y2) / (x2 - x3);
        for (int k = 0; k <= p->x_deg; ++k)
            for (int n = 0; n <= p->y_deg; ++n)
                int2 += p->koef[k * (p->y_deg + 1) + n] *
                        trapezoid(k, n, k1, b1, k2, b2, x3, x2);
    } else {
        int2 = 0;
    }

    if (y3 >= y1 && y3 >= y2) return int1 + int2;
    else if (y3 <= y1 && y3 <= y2) return -int1 - int2;
    else if (y3 >= k1 * x3 + b1) return int1 + int2;
    else return -int1 - int2;
}

double or_integrate(const or_poly *p, const double xs[3], const double ys[3],
                    int weight)
{
    switch (weight) {
    case 0: return integrate1(p, xs, ys, or_trapezoid_integral);
    case 1: return integrate1(p, xs, ys, or_trapezoid_integral_cos);
    case 2: return integrate1(p, xs, ys, or_trapezoid_integral_sin);
    default: return integrate1(p, xs, ys, or_trapezoid_integral_1_cos);
    }
}

/* ------------------------------------------------------------------------ */
/* Element callbacks                                                         */
/* ------------------------------------------------------------------------ */
/* slaplace.cpp:75-105: slaplace(a, b) = laplace1(a, b) + laplace2(a, b) */
static double slaplace(const or_poly *a, const or_poly *b, const double xs[3],
                       const double ys[3])
{
    or_poly da, db, poly1;
    or_poly_diff(a, 0, &da);
    or_poly_diff(b, 0, &db);
    or_poly_mul(&da, &db, &poly1);
    double l1 = -or_integrate(&poly1, xs, ys, 1);
    or_poly_diff(a, 1, &da);
    or_poly_diff(b, 1, &db);
    or_poly_mul(&da, &db, &poly1);
    double l2 = -or_integrate(&poly1, xs, ys, 3);
    return l1 + l2;
}

static double jac_cos_func(double x, double y, void *d)  /* sjacobian.cpp:50-53 */
{
    return or_cos(x) * or_poly_apply((const or_poly *)d, x, y);
}

static double jac_func(double x, double y, void *d)      /* sjacobian.cpp:55-58 */
{
    return or_poly_apply((const or_poly *)d, x, y);
}

double or_element(int kind, const or_poly *phi_i, const or_poly *phi_j,
                  const double xs[3], const double ys[3], double c_mass,
                  double c_lap)
{
    or_poly poly;
    switch (kind) {
    case 0:   /* id_cb: slaplace.cpp:185-197 */
        or_poly_mul(phi_i, phi_j, &poly);
        return or_integrate(&poly, xs, ys, 1);
    case 1:   /* slaplace_integrate_cb: slaplace.cpp:171-183 */
        return slaplace(phi_j, phi_i, xs, ys);
    case 2: { /* bv_private::integrate_cb: barvortex_impl.h:57-80 */
        or_poly_mul(phi_i, phi_j, &poly);
        double pt1 = or_integrate(&poly, xs, ys, 1);
        pt1 *= c_mass;
        double pt2 = slaplace(phi_j, phi_i, xs, ys);
        pt2 *= c_lap;
        return pt1 + pt2;
    }
    case 3:   /* SphereJacobian id_cb: sjacobian.cpp:60-74 */
        or_poly_mul(phi_i, phi_j, &poly);
        return or_cubature7(xs[0], ys[0], xs[1], ys[1], xs[2], ys[2],
                            jac_cos_func, &poly);
    default: {
        /* diff_1_rp, diff_2_rp, diff_1_cos_rp, diff_2_cos_rp
         * (sjacobian.cpp:76-134): poly = diff(phi_j, d) * phi_i, times
         * v = 1, integrated by cubature7 with or without cos x */
        or_poly dj;
        const int d = (kind == 4 || kind == 6) ? 0 : 1;
        const int wcos = (kind >= 6);
        or_poly_diff(phi_j, d, &dj);
        or_poly_mul(&dj, phi_i, &poly);
        double v = 1;
        return v * or_cubature7(xs[0], ys[0], xs[1], ys[1], xs[2], ys[2],
                                wcos ? jac_cos_func : jac_func, &poly);
    }
    }
}

/* ------------------------------------------------------------------------ */
/* Assembly: src/generators.cpp:57-212                                       */
/* ------------------------------------------------------------------------ */
static void tri_coords(const or_mesh *m, int t, int zone, double xs[3],
                       double ys[3])
{
    for (int a = 0; a < 3; ++a) {
        int p = m->tri[3 * t + a];
        xs[a] = m->coords[((long)p * m->n_zones + zone) * 2 + 0];
        ys[a] = m->coords[((long)p * m->n_zones + zone) * 2 + 1];
    }
}

static int local_index(const or_mesh *m, int t, int p)  /* mesh.cpp:550-561 */
{
    for (int a = 0; a < 3; ++a)
        if (m->tri[3 * t + a] == p) return a;
    abort();
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* columns of row i: sorted unique p2io of inner vertices around inner[i] */
static int row_columns(const or_mesh *m, int i, int *cols)
{
    int p = m->inner[i], cnt = 0, tmp[64];
    for (int e = m->adj_ptr[p]; e < m->adj_ptr[p + 1]; ++e) {
        int t = m->adj[e];
        for (int a = 0; a < 3; ++a) {
            int q = m->tri[3 * t + a];
            if (!m->is_boundary[q]) tmp[cnt++] = m->p2io[q];
        }
    }
    qsort(tmp, cnt, sizeof(int), cmp_int);
    int u = 0;
    for (int c = 0; c < cnt; ++c)
        if (u == 0 || tmp[c] != tmp[u - 1]) tmp[u++] = tmp[c];
    if (cols) memcpy(cols, tmp, u * sizeof(int));
    return u;
}

long or_csr_structure(const or_mesh *m, int *row_ptr, int *col_idx)
{
    long nnz = 0;
    int cols[64];
    for (int i = 0; i < m->n_inner; ++i) {
        row_ptr[i] = (int)nnz;
        int c = row_columns(m, i, cols);
        if (col_idx) memcpy(col_idx + nnz, cols, c * sizeof(int));
        nnz += c;
    }
    row_ptr[m->n_inner] = (int)nnz;
    return nnz;
}

static long find_pos(const int *row_ptr, const int *col_idx, int i, int j)
{
    for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
        if (col_idx[e] == j) return e;
    abort();
}

int or_assemble(const or_mesh *m, int kind, double c_mass, double c_lap,
                const int *row_ptr, const int *col_idx, double *vals)
{
    int rs = m->n_inner;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rs; ++i) {
        for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) vals[e] = 0.0;
        int p = m->inner[i];
        int zone = m->tri_zone[m->adj[m->adj_ptr[p]]];   /* generators.cpp:76 */
        for (int e = m->adj_ptr[p]; e < m->adj_ptr[p + 1]; ++e) {
            int t = m->adj[e];
            double xs[3], ys[3];
            or_poly phik[3];
            tri_coords(m, t, zone, xs, ys);
            or_basis_p1(xs, ys, phik);
            const or_poly *phi_i = &phik[local_index(m, t, p)];
            for (int i0 = 0; i0 < 3; ++i0) {
                int p2 = m->tri[3 * t + i0];
                if (m->is_boundary[p2]) continue;
                int j = m->p2io[p2];
                double a = or_element(kind, phi_i, &phik[i0], xs, ys, c_mass,
                                      c_lap);
                vals[find_pos(row_ptr, col_idx, i, j)] += a;
            }
        }
    }
    return 0;
}

/* Baroclin::Baroclin's A_ (examples/baroclin.cpp:69-181): generate_matrixv
 * with integrate_cb, whose 9 elements per (row i, triangle, column j) are
 * written here exactly as the reference writes them, from a = integrate_cos
 * (phi_i phi_j) and b = slaplace(phi_j, phi_i).  The 4n x 4n matrix has
 * nine n x n blocks with the inner stencil; vals9[k * nnz + e] is entry e of
 * block k in the order (1,1) (1,2) (2,1) (2,2) (2,4) (3,1) (3,3) (4,2) (4,4)
 * of the reference's push_back calls.  p = {tau, sigma, mu, sigma1, mu1,
 * alpha, theta, backward}: backward != 0 gives Ab_ instead, from
 * integrate_backward_cb (examples/baroclin.cpp:184-244), same block order. */
int or_assemble_baroclin(const or_mesh *m, const double *p, const int *row_ptr,
                         const int *col_idx, double *vals9)
{
    const double tau = p[0], sigma = p[1], mu = p[2], sigma1 = p[3], mu1 = p[4],
                 alpha = p[5], theta = p[6];
    const int backward = p[7] != 0.0;
    const long nnz = row_ptr[m->n_inner];
    int rs = m->n_inner;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rs; ++i) {
        for (int k = 0; k < 9; ++k)
            for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e) vals9[k * nnz + e] = 0.0;
        int pt = m->inner[i];
        int zone = m->tri_zone[m->adj[m->adj_ptr[pt]]];
        for (int e = m->adj_ptr[pt]; e < m->adj_ptr[pt + 1]; ++e) {
            int t = m->adj[e];
            double xs[3], ys[3];
            or_poly phik[3];
            tri_coords(m, t, zone, xs, ys);
            or_basis_p1(xs, ys, phik);
            const or_poly *phi_i = &phik[local_index(m, t, pt)];
            for (int i0 = 0; i0 < 3; ++i0) {
                int p2 = m->tri[3 * t + i0];
                if (m->is_boundary[p2]) continue;
                int j = m->p2io[p2];
                double a = or_element(0, phi_i, &phik[i0], xs, ys, 0.0, 0.0);
                double b = or_element(1, phi_i, &phik[i0], xs, ys, 0.0, 0.0);
                double v[9];
                if (!backward) {
                    v[0] = a * (1.0 / tau + 0.5 * theta * sigma) - theta * mu * b;
                    v[1] = -0.5 * a * sigma * theta;
                    v[2] = 0.5 * theta * sigma * a;
                    v[3] = a * (1.0 / tau + 0.5 * theta * sigma + alpha * alpha * theta * mu1)
                           - theta * mu * b;
                    v[4] = a * (-alpha * alpha / tau - alpha * alpha * theta * sigma1);
                } else {
                    v[0] = a * (1.0 / tau - 0.5 * (1.0 - theta) * sigma) + (1.0 - theta) * mu * b;
                    v[1] = 0.5 * a * sigma * (1.0 - theta);
                    v[2] = -0.5 * (1.0 - theta) * sigma * a;
                    v[3] = a * (1.0 / tau - 0.5 * (1.0 - theta) * sigma
                                - alpha * alpha * (1.0 - theta) * mu1)
                           + (1.0 - theta) * mu * b;
                    v[4] = a * (-alpha * alpha / tau + alpha * alpha * (1.0 - theta) * sigma1);
                }
                v[5] = a;
                v[6] = -b;
                v[7] = a;
                v[8] = -b;
                long pos = find_pos(row_ptr, col_idx, i, j);
                for (int k = 0; k < 9; ++k) vals9[k * nnz + pos] += v[k];
            }
        }
    }
    return 0;
}

/* generate_right_part (src/generators.h:245-287) with the interior branch of
 * Baroclin's right_part_cb (examples/baroclin.cpp:329-336): b[i] sums
 * F[j] * integrate_cos(phi_i phi_j) over the triangles around inner[i] (in
 * adjacency order) and their inner vertices (in triangle order).  The
 * boundary branch adds -bnd * element, zero for zero boundary values. */
int or_right_part_mass(const or_mesh *m, const double *F, double *b)
{
    int rs = m->n_inner;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < rs; ++i) {
        int pt = m->inner[i];
        int zone = m->tri_zone[m->adj[m->adj_ptr[pt]]];
        double acc = 0.0;
        for (int e = m->adj_ptr[pt]; e < m->adj_ptr[pt + 1]; ++e) {
            int t = m->adj[e];
            double xs[3], ys[3];
            or_poly phik[3];
            tri_coords(m, t, zone, xs, ys);
            or_basis_p1(xs, ys, phik);
            const or_poly *phi_i = &phik[local_index(m, t, pt)];
            for (int i0 = 0; i0 < 3; ++i0) {
                int p2 = m->tri[3 * t + i0];
                if (m->is_boundary[p2]) continue;
                double a = or_element(0, phi_i, &phik[i0], xs, ys, 0.0, 0.0);
                acc += F[m->p2io[p2]] * a;
            }
        }
        b[i] = acc;
    }
    return 0;
}

/* generate_full_matrix (src/generators.h:200-234): all points x all points,
 * row p in the zone of its first triangle, no boundary skipped.  Structure:
 * row p holds every point of every triangle around p, ascending. */
long or_full_structure(const or_mesh *m, int *row_ptr, int *col_idx)
{
    long nnz = 0;
    row_ptr[0] = 0;
    for (int p = 0; p < m->n_points; ++p) {
        int tmp[64], nt = 0;
        for (int e = m->adj_ptr[p]; e < m->adj_ptr[p + 1]; ++e)
            for (int a = 0; a < 3; ++a) {
                int q = m->tri[3 * m->adj[e] + a], dup = 0;
                for (int c = 0; c < nt; ++c) dup |= tmp[c] == q;
                if (!dup && nt < 64) tmp[nt++] = q;
            }
        for (int a = 1; a < nt; ++a)          /* insertion sort */
            for (int b = a; b > 0 && tmp[b - 1] > tmp[b]; --b) {
                int t = tmp[b]; tmp[b] = tmp[b - 1]; tmp[b - 1] = t;
            }
        if (col_idx)
            for (int c = 0; c < nt; ++c) col_idx[nnz + c] = tmp[c];
        nnz += nt;
        row_ptr[p + 1] = (int)nnz;
    }
    return nnz;
}

int or_assemble_full(const or_mesh *m, int kind, double c_mass, double c_lap,
                     const int *row_ptr, const int *col_idx, double *vals)
{
    for (int p = 0; p < m->n_points; ++p) {
        for (int e = row_ptr[p]; e < row_ptr[p + 1]; ++e) vals[e] = 0.0;
        int zone = m->tri_zone[m->adj[m->adj_ptr[p]]];   /* generators.h:213 */
        for (int e = m->adj_ptr[p]; e < m->adj_ptr[p + 1]; ++e) {
            int t = m->adj[e];
            double xs[3], ys[3];
            or_poly phik[3];
            tri_coords(m, t, zone, xs, ys);
            or_basis_p1(xs, ys, phik);
            const or_poly *phi_i = &phik[local_index(m, t, p)];
            for (int i0 = 0; i0 < 3; ++i0) {
                int p2 = m->tri[3 * t + i0];
                double a = or_element(kind, phi_i, &phik[i0], xs, ys, c_mass, c_lap);
                vals[find_pos(row_ptr, col_idx, p, p2)] += a;
            }
        }
    }
    return 0;
}

/* boundary matrix rows = inner points touching an outer point */
long or_boundary_structure(const or_mesh *m, int *row_ptr, int *col_idx)
{
    int rs = m->n_inner;
    int *cnt = (int *)calloc(rs + 1, sizeof(int));
    int **cols = (int **)calloc(rs, sizeof(int *));
    for (int j = 0; j < m->n_outer; ++j) {
        int p2 = m->outer[j];
        for (int e = m->adj_ptr[p2]; e < m->adj_ptr[p2 + 1]; ++e) {
            int t = m->adj[e];
            for (int a = 0; a < 3; ++a) {
                int p = m->tri[3 * t + a];
                if (m->is_boundary[p]) continue;
                int i = m->p2io[p];
                int dup = 0;
                for (int c = 0; c < cnt[i]; ++c) dup |= cols[i][c] == j;
                if (dup) continue;
                cols[i] = (int *)realloc(cols[i], (cnt[i] + 1) * sizeof(int));
                cols[i][cnt[i]++] = j;
            }
        }
    }
    long nnz = 0;
    for (int i = 0; i < rs; ++i) {
        row_ptr[i] = (int)nnz;
        if (cnt[i]) qsort(cols[i], cnt[i], sizeof(int), cmp_int);
        if (col_idx && cnt[i]) memcpy(col_idx + nnz, cols[i], cnt[i] * sizeof(int));
        nnz += cnt[i];
        free(cols[i]);
    }
    row_ptr[rs] = (int)nnz;
    free(cols);
    free(cnt);
    return nnz;
}

/* generators.cpp:141-186: zone of the outer point, phi_i = inner basis */
int or_assemble_boundary(const or_mesh *m, int kind, double c_mass,
                         double c_lap, const int *row_ptr, const int *col_idx,
                         double *vals)
{
    for (int e = 0; e < row_ptr[m->n_inner]; ++e) vals[e] = 0.0;
    for (int j = 0; j < m->n_outer; ++j) {
        int p2 = m->outer[j];
        int zone = m->tri_zone[m->adj[m->adj_ptr[p2]]];
        for (int e = m->adj_ptr[p2]; e < m->adj_ptr[p2 + 1]; ++e) {
            int t = m->adj[e];
            double xs[3], ys[3];
            or_poly phik[3];
            tri_coords(m, t, zone, xs, ys);
            or_basis_p1(xs, ys, phik);
            const or_poly *phi_j = &phik[local_index(m, t, p2)];
            for (int i0 = 0; i0 < 3; ++i0) {
                int p = m->tri[3 * t + i0];
                if (m->is_boundary[p]) continue;
                int i = m->p2io[p];
                double a = or_element(kind, &phik[i0], phi_j, xs, ys, c_mass,
                                      c_lap);
                vals[find_pos(row_ptr, col_idx, i, j)] += a;
            }
        }
    }
    return 0;
}
