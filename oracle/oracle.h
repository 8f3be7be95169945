/*
 * CPU oracle for the Phelm sphere FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (mke_amd/) never links, loads or calls it.
 *
 * It restates, in plain C and in the reference's operation order, the
 * reference's CPU FP64 arithmetic for:
 *   - GK15 and the 7-point triangle rule          src/quadrature.c:49-171
 *   - trapezoid integrals (exact, cos, sin, 1/cos) src/util.cpp:97-251
 *   - Polynom ops and integrate1                  src/polynom.cpp:63-340,667-729
 *   - P1 basis in a chart                         src/mesh.cpp:563-607
 *   - row-wise assembly                           src/generators.cpp:57-212
 *   - mass / stiffness / Helmholtz callbacks      src/slaplace.cpp:75-197,
 *                                                 examples/barvortex_impl.h:57-80
 *   - CSR SpMV and GMRES(m) with MGS Arnoldi and Givens rotations: the
 *     reference's solver library (linal) is not in the tree, so this GMRES is
 *     a textbook restatement (Saad & Schultz 1986) and is "parity unpinned".
 *
 * Two functions the reference takes from outside its tree are fixed here:
 *   ipow(x, p)  - linal's pow.h is absent: repeated multiplication r *= x,
 *                 p times, from r = 1 ("parity unpinned" vs linal).
 *   cos(x)      - the reference calls the C library's cos (src/util.cpp:
 *                 194-215), which is not bit-reproducible across library
 *                 builds; the oracle and the device kernels share one
 *                 correctly rounded cos/sin (mke_amd/csrc/cr_trig.h, Ziv's
 *                 method), pinned against 200-bit mpmath values.
 *                 or_use_libm_cos(1) switches to the C library's cos to
 *                 measure the spread between the two.
 */
#ifndef PHELM_ORACLE_H
#define PHELM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef double (*or_fx_t)(double x, void *data);
typedef double (*or_fxy_t)(double x, double y, void *data);

/* --- scalar math --------------------------------------------------------- */
double or_cos(double x);
double or_sin(double x);
double or_ipow(double x, int p);
void   or_use_libm_cos(int on);
/* test hooks of the correctly rounded routine: the set of distinct cos
 * arguments evaluated while buf (cap entries, a power of two) is installed
 * (NULL removes it); the count, or minus the arguments lost to a full set */
void   or_record_cos_args(double *buf, long cap);
long   or_recorded_cos_args(void);
double or_cr_cos_bound(double x, double err);
double or_cr_sin_bound(double x, double err);
double or_cr_slow(double x, int odd);
void   or_cr_fast_dd(double x, int odd, double *hi, double *lo);

/* --- quadrature (src/quadrature.c) --------------------------------------- */
double or_gauss_kronrod15(double a, double b, or_fx_t f, void *d);
double or_cubature7(double x1, double y1, double x2, double y2, double x3,
                    double y3, or_fxy_t f, void *d);

/* --- trapezoid integrals (src/util.cpp) ----------------------------------- */
double or_trapezoid_integral(int k, int n, double k1, double b1, double k2,
                             double b2, double x1, double x3);
double or_trapezoid_integral_cos(int k, int n, double k1, double b1, double k2,
                                 double b2, double x1, double x3);
double or_trapezoid_integral_sin(int k, int n, double k1, double b1, double k2,
                                 double b2, double x1, double x3);
double or_trapezoid_integral_1_cos(int k, int n, double k1, double b1,
                                   double k2, double b2, double x1, double x3);

/* --- polynomials in (x, y) (src/polynom.h:81-148) ------------------------- */
#define OR_POLY_MAX 16
typedef struct {
    int x_deg, y_deg;
    double koef[OR_POLY_MAX];   /* koef[i * (y_deg + 1) + j] is the x^i y^j coefficient */
} or_poly;

/* weight: 0 = none, 1 = cos x, 2 = sin x, 3 = 1/cos x  (src/polynom.cpp:322-340) */
double or_integrate(const or_poly *p, const double xs[3], const double ys[3],
                    int weight);
/* P1 basis of a triangle with chart coordinates xs, ys (src/mesh.cpp:582-607) */
void or_basis_p1(const double xs[3], const double ys[3], or_poly out[3]);
/* Polynom ops exported for the known-answer tests */
void or_poly_mul(const or_poly *a, const or_poly *b, or_poly *out);
void or_poly_diff(const or_poly *a, int d, or_poly *out);
double or_poly_apply(const or_poly *a, double x, double y);

/* --- element callbacks ----------------------------------------------------- */
/* kind: 0 = mass  (id_cb,                src/slaplace.cpp:185-197)
 *       1 = stiffness (slaplace_integrate_cb, src/slaplace.cpp:171-183)
 *       2 = Helmholtz c_mass*mass + c_lap*stiffness per triangle
 *                                         (examples/barvortex_impl.h:57-80)
 *       3 = Jacobian mass, cubature7 (src/sjacobian.cpp:60-74)
 *       4 = diff_1_rp, 5 = diff_2_rp, 6 = diff_1_cos_rp, 7 = diff_2_cos_rp
 *           (src/sjacobian.cpp:76-134)                                    */
double or_element(int kind, const or_poly *phi_i, const or_poly *phi_j,
                  const double xs[3], const double ys[3], double c_mass,
                  double c_lap);

/* --- mesh view (arrays as Mesh::load leaves them, src/mesh.cpp:323-519) --- */
typedef struct {
    int n_points, n_triangles, n_zones;
    const double *coords;   /* [n_points][n_zones][2] chart (x=lat, y=lon) */
    const int *tri;         /* [n_triangles][3] 0-based point ids            */
    const int *tri_zone;    /* [n_triangles] 0-based zone                   */
    const int *adj_ptr;     /* [n_points + 1] point -> triangles, tri-id order */
    const int *adj;
    const int *is_boundary; /* [n_points]                                   */
    const int *p2io;        /* [n_points]                                   */
    const int *inner;       /* [n_inner]                                    */
    const int *outer;       /* [n_outer]                                    */
    int n_inner, n_outer;
} or_mesh;

/* CSR structure of the inner x inner operator: row i holds p2io of every
 * inner vertex of every triangle around inner[i]; columns ascending.
 * row_ptr has n_inner + 1 entries; col_idx must hold nnz (ask with col_idx=0). */
long or_csr_structure(const or_mesh *m, int *row_ptr, int *col_idx);
/* Values in the reference order: zone of row i = zone of adj[p][0], adds in
 * adjacency order (src/generators.cpp:57-114).  Uses OpenMP over rows. */
int or_assemble(const or_mesh *m, int kind, double c_mass, double c_lap,
                const int *row_ptr, const int *col_idx, double *vals);
/* Baroclin's A_ (or Ab_ when backward != 0) as nine blocks
 * (examples/baroclin.cpp:69-244); p = {tau, sigma, mu, sigma1, mu1, alpha,
 * theta, backward}; vals9: 9 x nnz. */
int or_assemble_baroclin(const or_mesh *m, const double *p, const int *row_ptr,
                         const int *col_idx, double *vals9);
/* generate_right_part with the interior branch of Baroclin's right_part_cb
 * (examples/baroclin.cpp:329-336): b = element-order sums of F_j * mass. */
int or_right_part_mass(const or_mesh *m, const double *F, double *b);
/* generate_full_matrix (src/generators.h:200-234): every point x every
 * point, no boundary skipped (SphereNorm's matrix, src/norm.h:71-74) */
long or_full_structure(const or_mesh *m, int *row_ptr, int *col_idx);
int or_assemble_full(const or_mesh *m, int kind, double c_mass, double c_lap,
                     const int *row_ptr, const int *col_idx, double *vals);
/* inner x outer boundary matrix (src/generators.cpp:141-186) as CSR */
long or_boundary_structure(const or_mesh *m, int *row_ptr, int *col_idx);
int or_assemble_boundary(const or_mesh *m, int kind, double c_mass,
                         double c_lap, const int *row_ptr, const int *col_idx,
                         double *vals);

/* --- solve ----------------------------------------------------------------- */
void or_spmv(int n, const int *row_ptr, const int *col_idx, const double *vals,
             const double *x, double *y);
typedef struct {
    int iterations;   /* Arnoldi steps */
    int restarts;     /* completed cycles */
    double rel_residual; /* true ||b - A x|| / ||b|| at exit */
    int converged;
} or_gmres_info;
/* GMRES(m), x0 taken from x, MGS Arnoldi, Givens rotations; stops when the
 * true FP64 residual ||b-Ax||/||b|| <= rtol or after max_iters steps. */
int or_gmres(int n, const int *row_ptr, const int *col_idx, const double *vals,
             const double *b, double *x, int restart, double rtol,
             int max_iters, or_gmres_info *info);
int or_num_threads(void);
void or_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
