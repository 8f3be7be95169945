/*
 * phelm_gpu.h - C ABI of the MI355X (gfx950) sphere-FEM solve path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = ok, < 0 = error, text in phg_last_error()).  Vector arguments named
 * d_* are device pointers owned by the caller (the reference passes device
 * arrays under its GPU build, examples/schafe_impl.h:134-149); calls are
 * ordered on the stream given (NULL = the default stream).
 *
 * Each entry point names the reference interface it replaces.  The
 * reference's solver library (Solver<T>, vec_* helpers) is external and not
 * in its tree; its call sites are cited instead.
 */
#ifndef PHELM_GPU_H
#define PHELM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phg_mesh_s *phg_mesh_t;
typedef struct phg_matrix_s *phg_matrix_t;
typedef void *phg_stream_t; /* a hipStream_t */

enum {
    PHG_OK = 0,
    PHG_ERR_ARG = -1,      /* bad argument                               */
    PHG_ERR_FORMAT = -2,   /* mesh text rejected (Mesh::load "bad file") */
    PHG_ERR_HIP = -3,      /* HIP runtime error                          */
    PHG_ERR_NOMEM = -4,    /* allocation failed / over admission budget  */
    PHG_ERR_NOCONV = 1     /* solver hit max_iters (not an error: result valid) */
};

/* value precisions of a matrix / solve */
enum { PHG_F64 = 0, PHG_F32 = 1, PHG_BF16 = 2 };
/* element kinds: callbacks of src/slaplace.cpp:171-197 and
 * examples/barvortex_impl.h:57-80 (GK15 trapezoid integrals), and the
 * SphereJacobian callbacks of src/sjacobian.cpp:60-134 (7-point rule):
 * JAC_MASS = id_cb (cos x phi_i phi_j), DIFF1/DIFF2 = diff_1_rp/diff_2_rp
 * (d_x phi_j phi_i, d_y phi_j phi_i), DIFF1_COS/DIFF2_COS = the same times
 * cos x (diff_1_cos_rp/diff_2_cos_rp) */
enum {
    PHG_MASS = 0, PHG_LAPLACE = 1, PHG_HELMHOLTZ = 2, PHG_JAC_MASS = 3,
    PHG_DIFF1 = 4, PHG_DIFF2 = 5, PHG_DIFF1_COS = 6, PHG_DIFF2_COS = 7
};
/* orthogonalisation of the Arnoldi step: modified Gram-Schmidt, classical
 * Gram-Schmidt with one re-orthogonalisation pass, or one classical pass
 * (single pass: block solver only) */
enum { PHG_MGS = 0, PHG_CGS2 = 1, PHG_CGS1 = 2 };

/* ---- library ---------------------------------------------------------- */
/* Version string and last error of the calling thread. */
const char *phg_version(void);
const char *phg_last_error(void);
/* Selects the HIP device for later calls of this thread (linal_init). */
int phg_init(int device);

/* ---- mesh (src/mesh.cpp:323-519 Mesh::load; utils/icosahedron.cpp:481-624) */
/* Text of `mke_mesh_sphere --type full --coord local --iter level`; the
 * buffer is malloc'ed by the library, free it with phg_free(). */
int phg_mesh_sphere_text(int level, char **text, size_t *len);
void phg_free(void *p);
/* Mesh::load of a P1 mesh text (src/mesh.cpp:323-519). */
int phg_mesh_load(const char *text, size_t len, phg_mesh_t *out);
/* phg_mesh_sphere_text + phg_mesh_load without the caller seeing the text. */
int phg_mesh_sphere(int level, phg_mesh_t *out);
/* Binary image of a loaded mesh (all Mesh::load arrays, with a checksum):
 * no reference counterpart; it replaces re-running mke_mesh_sphere and
 * Mesh::load (utils/icosahedron.cpp:481-624, src/mesh.cpp:323-519) at
 * start-up.  load_file returns PHG_ERR_FORMAT for a damaged or foreign file. */
int phg_mesh_save(phg_mesh_t m, const char *path);
int phg_mesh_load_file(const char *path, phg_mesh_t *out);
/* The level's sphere from dir/sphere_L<level>.phgmesh, generated and written
 * there when absent or unreadable. */
int phg_mesh_sphere_cached(int level, const char *dir, phg_mesh_t *out);
/* print_function (src/phelm.cpp:248-309) of host values h_u (n_points):
 * x y z f u v per point (%.16f; u, v the zone-0 chart coordinates; x, y, z
 * = u, v, 0, or with sphere_xyz the unit-sphere map cos u cos v, cos u sin v,
 * sin u), then the 1-based triangles.  Free *text with phg_free(). */
int phg_print_function(phg_mesh_t m, const double *h_u, int sphere_xyz, char **text,
                       size_t *len);
/* Sizes: m.ps.size(), m.tr.size(), inner_size, outer_size, zone count. */
int phg_mesh_info(phg_mesh_t m, int *n_points, int *n_triangles,
                  int *n_inner, int *n_outer, int *n_zones);
/* Host copies of the loaded arrays (any pointer may be NULL):
 * pr[V*3], coords[V*n_zones*2] (chart x = latitude, y = longitude),
 * tri[T*3] (0-based), tri_zone[T] (0-based), adj_ptr[V+1], adj[3T]
 * (triangle-id order), is_boundary[V], inner[n_inner], outer[n_outer],
 * p2io[V] (mesh.h:354-427). */
int phg_mesh_arrays(phg_mesh_t m, double *pr, double *coords, int *tri,
                    int *tri_zone, int *adj_ptr, int *adj, int *is_boundary,
                    int *inner, int *outer, int *p2io);
void phg_mesh_destroy(phg_mesh_t m);
/* CSR pattern of the inner x inner operator that phg_assemble fills: row i
 * holds p2io of every inner vertex of every triangle around inner[i],
 * columns ascending (the Solver::add pattern of src/generators.cpp:71-110).
 * row_ptr has n_inner + 1 entries; pass col_idx = NULL to get *nnz first.
 * Host only. */
int phg_mesh_csr_structure(phg_mesh_t m, int *row_ptr, int *col_idx, long *nnz);

/* ---- matrices (Solver<T> of the external solver library) --------------- */
/* Solver<T>(n) + add(i, j, a) + first use: an n_rows x n_cols matrix built
 * from host triplets; duplicates accumulate in the order given, like
 * Solver::add (src/generators.cpp:49-55, src/impl/generators_impl.h:38-72). */
int phg_matrix_from_triplets(int n_rows, int n_cols, long n, const int *rows,
                             const int *cols, const double *vals,
                             phg_matrix_t *out);
/* Same from a host CSR (columns ascending in each row). */
int phg_matrix_from_csr(int n_rows, int n_cols, const int *row_ptr,
                        const int *col_idx, const double *vals,
                        phg_matrix_t *out);
/* generate_matrix(A, m, cb) (src/generators.cpp:116-139) on the GPU for the
 * inner x inner operator.  kind selects cb: PHG_MASS = id_cb
 * (src/slaplace.cpp:185-197), PHG_LAPLACE = slaplace_integrate_cb
 * (src/slaplace.cpp:171-183), PHG_HELMHOLTZ = c_mass * mass + c_lap *
 * stiffness per triangle (examples/barvortex_impl.h:57-80 with c_mass =
 * 1/tau + sigma*theta, c_lap = -theta*mu; examples/schafe.cpp:67-89 with
 * theta = 1/2).  CSR structure is built on the host, values on the device,
 * added in the reference's row-wise order.  cos is correctly rounded, so the
 * values equal those of a reference build whose C library cos rounds
 * correctly at every argument used; a library cos that is one ulp off moves
 * mass entries (which cancel ~1e7-fold) by up to ~1e-9 relative at level 7,
 * Helmholtz entries (M - 1e-2 L) by under 1e-12. */
int phg_assemble(phg_mesh_t m, int kind, double c_mass, double c_lap,
                 phg_stream_t stream, phg_matrix_t *out);
/* generate_boundary_matrix(A, m, cb) (src/generators.cpp:188-212): the
 * inner x outer matrix, same kinds. */
int phg_assemble_boundary(phg_mesh_t m, int kind, double c_mass, double c_lap,
                          phg_matrix_t *out);
/* generate_full_matrix(A, m, cb) (src/generators.h:200-234): every point x
 * every point, rows in the zone of their first triangle, nothing skipped;
 * with PHG_MASS it is SphereNorm's matrix (src/norm.h:71-74,
 * sphere_scalar_cb src/phelm.cpp:330-335). */
int phg_assemble_full(phg_mesh_t m, int kind, double c_mass, double c_lap,
                      phg_stream_t stream, phg_matrix_t *out);
/* Shape and host copies of the CSR (pointers may be NULL). */
int phg_matrix_info(phg_matrix_t A, int *n_rows, int *n_cols, long *nnz);
int phg_matrix_csr(phg_matrix_t A, int *row_ptr, int *col_idx, double *vals);
void phg_matrix_destroy(phg_matrix_t A);

/* ---- products (Solver::mult_vector, 68 call sites, e.g.
 *      src/impl/slaplace_impl.h:109,168-169) ------------------------------- */
/* y = A x, FP64, device vectors of n_cols / n_rows entries. */
int phg_spmv(phg_matrix_t A, const double *d_x, double *d_y, phg_stream_t s);
/* y = A x with FP32 vectors and values stored as prec (PHG_F32 / PHG_BF16). */
int phg_spmv_f32(phg_matrix_t A, int prec, const float *d_x, float *d_y,
                 phg_stream_t s);
/* Y = A X for nrhs right-hand sides, row-major n x nrhs (ld = nrhs), FP32. */
int phg_spmm_f32(phg_matrix_t A, int prec, int nrhs, const float *d_X,
                 float *d_Y, phg_stream_t s);

/* ---- BLAS-1 (vec_scalar2, vec_sum1, vec_mult_scalar, vec_diff: e.g.
 *      src/mesh.h:550-591, examples/barvortex_impl.h:234-270) ------------- */
int phg_dot(int n, const double *d_x, const double *d_y, double *result,
            phg_stream_t s);   /* result: host pointer, synchronises */
int phg_nrm2(int n, const double *d_x, double *result, phg_stream_t s);
/* Allocation-free dot product.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (work captured into a HIP
 * graph, tight serving loops, harnesses that forbid hipMalloc).  *d_result
 * (a device pointer) = x . y, ordered on the stream, with no host
 * synchronisation and no hipMalloc/hipFree: the block partials live in a
 * 32 KiB device array of the library (one per GPU, held until the process
 * ends), which limits n to PHG_DOT_NOALLOC_MAX (larger n: PHG_ERR_ARG).  All
 * calls on one GPU share that array, so calls that overlap in time on one GPU
 * race; order them on one stream.  Deterministic; the summation order
 * differs from phg_dot's, whose partials are added on the host. */
#define PHG_DOT_NOALLOC_MAX 4194304
int phg_dot_noalloc(int n, const double *d_x, const double *d_y, double *d_result,
                    phg_stream_t s);
/* r = k1 * a + k2 * b  (vec_sum1) */
int phg_axpby(int n, double k1, const double *d_a, double k2,
              const double *d_b, double *d_r, phg_stream_t s);

/* ---- gather / scatter (src/proj.cpp:49-160, src/gpu/phelm.cu:45-157) ----- */
/* p[i] = u[inner[i]]  (u2p) */
int phg_u2p(phg_mesh_t m, const double *d_u, double *d_p, phg_stream_t s);
/* u[i] = boundary ? (bnd ? bnd[p2io[i]] : 0) : p[p2io[i]]  (p2u) */
int phg_p2u(phg_mesh_t m, const double *d_p, const double *d_bnd, double *d_u,
            phg_stream_t s);
/* b[k] = u[outer[k]]  (proj_bnd) */
int phg_proj_bnd(phg_mesh_t m, const double *d_u, double *d_b, phg_stream_t s);
/* u[i] = bnd ? bnd[p2io[i]] : 0 at the boundary points only  (set_bnd) */
int phg_set_bnd(phg_mesh_t m, const double *d_bnd, double *d_u, phg_stream_t s);

/* ---- GMRES (Solver::solve: src/mesh.h:740-746, src/impl/slaplace_impl.h:171) */
typedef struct {
    int restart;     /* m of GMRES(m), default 30                          */
    int max_iters;   /* Arnoldi steps in total (inner steps for "mixed")   */
    double rtol;     /* stop at ||b - A x||_2 / ||b||_2 <= rtol, FP64       */
    int ortho;       /* PHG_MGS (reference form) or PHG_CGS2               */
    int precision;   /* PHG_F64: plain form.  PHG_F32 / PHG_BF16: inner
                        Krylov in FP32 (values stored as precision) under
                        FP64 iterative refinement                          */
    double inner_rtol; /* relative reduction per refinement step (mixed)   */
    int use_x0;      /* 0: start from x = 0; 1: start from the given x      */
    int graph;       /* 1: replay each restart cycle as one captured hipGraph */
    int diag_scale;  /* 1: right-precondition with the inverse diagonal, x = D^-1 y
                        (the batched FP64 solves without another preconditioner;
                        the stopping test stays on ||b - A x|| / ||b||)    */
    int persistent;  /* phg_gmres with PHG_MGS: 1 (default) runs each restart
                        cycle as one cooperative kernel whose workgroups
                        exchange the dot partials through device memory; 0
                        launches a kernel per MGS pass (captured in a
                        hipGraph).  Both give the same bits and, measured at
                        level 7, the same speed; 1 falls back to 0 where
                        the device cannot keep every workgroup resident   */
} phg_gmres_opts;

typedef struct {
    int iterations;       /* Arnoldi steps executed                       */
    int restarts;         /* completed restart cycles                     */
    int refinements;      /* outer refinement steps (mixed forms)         */
    double rel_residual;  /* true FP64 ||b - A x|| / ||b|| at exit          */
    int converged;
    int cycle_launches;   /* persistent MGS cycles: kernels launched, and  */
    double cycle_ms;      /* their summed duration (HIP events recorded on
                             the solve's stream around each launch); 0
                             for the other forms                        */
} phg_gmres_info;

void phg_gmres_default_opts(phg_gmres_opts *o);
/* Solves A x = b; d_b and d_x are device vectors of n_rows entries. */
int phg_gmres(phg_matrix_t A, const double *d_b, double *d_x,
              const phg_gmres_opts *opts, phg_gmres_info *info,
              phg_stream_t s);

/* ---- batched: 64 right-hand sides, one matrix (no single reference call:
 *      it batches the Solver::solve of examples/barvortex_impl.h:296 over
 *      ensemble members sharing one mesh) ------------------------------- */
typedef struct {
    int iterations;          /* block Arnoldi steps                        */
    int cycles;              /* restart cycles (= refinement steps)        */
    double max_rel_residual; /* max over columns of the true FP64 residual */
    int converged;
} phg_block_info;
/* Block GMRES(restart <= 32) for A X = B with nrhs = 64 columns.  B and X
 * are device arrays of n_rows x 64 FP64, row-major (entry (i, q) at
 * i * 64 + q).  Krylov basis in FP32 (matrix values FP32, or BF16 with
 * opts->precision = PHG_BF16), block Gram-Schmidt on MFMA, FP64 Gram
 * matrices, Hessenberg reduction and refinement residual; stops when every
 * column has ||b - A x|| / ||b|| <= rtol. */
int phg_block_gmres(phg_matrix_t A, int nrhs, const double *d_B, double *d_X,
                    const phg_gmres_opts *opts, phg_block_info *info,
                    phg_stream_t s);

/* ---- batched independent solves: nrhs (1..64) systems A x_q = b_q of one
 *      matrix, each with its own Krylov space and stopping test, sharing the
 *      matrix reads (one SpMM per Arnoldi step).  Replaces repeated
 *      Solver::solve calls on one matrix: the 16 idt_.solve calls (8 distinct
 *      right-hand sides) of SphereJacobian::calc2 (src/sjacobian.cpp:136-362)
 *      and the implicit-step solves of ensemble members on one mesh
 *      (examples/barvortex_impl.h:296-298). --------------------------------- */
/* d_B, d_X: n_rows x nrhs FP64, row-major (entry (i, q) at i * nrhs + q).
 * opts->restart <= 32, opts->ortho PHG_CGS2 or PHG_CGS1, opts->precision
 * PHG_F64 or mixed (PHG_F32 / PHG_BF16 inner Krylov, FP64 refinement).
 * A column's result does not depend on the other columns of the call.
 * infos: nrhs host entries (may be NULL).  Returns PHG_ERR_NOCONV unless
 * every column converged. */
int phg_gmres_batch(phg_matrix_t A, int nrhs, const double *d_B, double *d_X,
                    const phg_gmres_opts *opts, phg_gmres_info *infos,
                    phg_stream_t s);
/* Launch shape of the batched solves (no reference counterpart: the CPU
 * solve has none).  The row kernels give each thread rows_per_thread rows
 * (1, 2, 4, 8 or 16) of one column; the value is kept per process for each
 * (n_rows, nrhs rounded up to a power of two, precision) shape and fixes the
 * grouping of every dot product's block partials, so results are bitwise
 * reproducible for a fixed value.  Untuned shapes take the largest value
 * (at most 8) that still gives >= 1024 workgroups; preconditioned solves
 * (phg_gmres_batch_mg) always take that default, whatever is tuned or pinned
 * for the plain form.  phg_gmres_batch_tune times one
 * restart cycle (opts->restart steps, opts->ortho and opts->precision; NULL:
 * defaults with CGS2) of nrhs hash-valued right-hand sides on A for each value,
 * keeps the fastest for the shape and returns it in *rows_per_thread (may be
 * NULL); ms (5 entries, may be NULL) receives the cycle times for 1, 2, 4, 8
 * and 16.
 * It allocates scratch vectors; do not call it while solves of the same
 * shape run on other streams. */
int phg_gmres_batch_tune(phg_matrix_t A, int nrhs, const phg_gmres_opts *opts,
                         phg_stream_t s, int *rows_per_thread, double *ms);
/* The current value for a shape (>= 1), or a negative error code. */
int phg_gmres_batch_shape(int n_rows, int nrhs, int precision);
/* Pins a shape's value (e.g. one a previous process tuned); 0 returns the
 * shape to the untuned default. */
int phg_gmres_batch_set_shape(int n_rows, int nrhs, int precision, int rows_per_thread);

/* ---- multigrid-preconditioned solves on the icosahedral sphere ----------
 * No single reference call: it makes the solves of SphereLaplace::solve
 * (src/impl/slaplace_impl.h:84-124, the pole-pinned Laplace stiffness, as
 * called once per fixed-point iteration by BarVortex::calc,
 * examples/barvortex_impl.h:298) converge in a level-independent number of
 * GMRES steps.  V-cycles over the nested meshes of levels 2..L of
 * mke_mesh_sphere; every level's operator is assembled on its own mesh. */
typedef struct phg_mg_s *phg_mg_t;
/* fine: a full sphere mesh of mke_mesh_sphere (any level); kind/c_mass/
 * c_lap as phg_assemble (PHG_MASS, PHG_LAPLACE or PHG_HELMHOLTZ).  The mesh
 * must outlive the object. */
int phg_mg_create(phg_mesh_t fine, int kind, double c_mass, double c_lap,
                  phg_stream_t s, phg_mg_t *out);
void phg_mg_destroy(phg_mg_t M);
/* The fine-level operator (owned by M; bit-identical to phg_assemble). */
int phg_mg_matrix(phg_mg_t M, phg_matrix_t *out);
int phg_mg_levels(phg_mg_t M);
/* phg_gmres_batch with M's fine operator, right-preconditioned by one
 * V-cycle per Arnoldi step (flexible form: the preconditioned vectors are
 * kept for the update).  opts->precision PHG_F64, or PHG_F32 / PHG_BF16:
 * Krylov vectors and V-cycle in FP32 (level operators FP32) under FP64
 * refinement.  Calls sharing one M must not overlap in time (its level
 * vectors are shared). */
int phg_gmres_batch_mg(phg_mg_t M, int nrhs, const double *d_B, double *d_X,
                       const phg_gmres_opts *opts, phg_gmres_info *infos,
                       phg_stream_t s);

/* ---- SphereJacobian (src/sjacobian.h:65-149) ---------------------------- */
typedef struct phg_jacobian_s *phg_jacobian_t;
/* SphereJacobian::SphereJacobian (src/sjacobian.cpp:415-446): assembles
 * idt_ (PHG_JAC_MASS), diff1_, diff2_, diff1_cos_, diff2_cos_ and their
 * boundary matrices on the GPU.  The mesh must outlive the object. */
int phg_jacobian_create(phg_mesh_t m, phg_stream_t s, phg_jacobian_t *out);
void phg_jacobian_destroy(phg_jacobian_t J);
/* The matrices (owned by J, do not destroy): which = 0 idt_, 1..4 diff1_,
 * diff2_, diff1_cos_, diff2_cos_, 5..8 their boundary (_rp_) matrices. */
int phg_jacobian_matrix(phg_jacobian_t J, int which, phg_matrix_t *out);
/* SphereJacobian::calc2 (src/sjacobian.cpp:348-362) for npairs (1..8)
 * pairs at once: u_k = d_u + k * n_points, v_k = d_v + k * n_points (full
 * mesh vectors), J(u_k, v_k) at the inner points to d_ans + k * n_inner.
 * The 8 * npairs mass solves run as one phg_gmres_batch with opts (NULL:
 * CGS1, restart 8, rtol 1e-13, diagonal scaling); infos (8 * npairs entries)
 * may be NULL. */
int phg_jacobian_calc2(phg_jacobian_t J, int npairs, const double *d_u,
                       const double *d_v, double *d_ans,
                       const phg_gmres_opts *opts, phg_gmres_info *infos,
                       phg_stream_t s);
/* SphereJacobian::calc2t (src/sjacobian.cpp:404-408): -calc2. */
int phg_jacobian_calc2t(phg_jacobian_t J, int npairs, const double *d_u,
                        const double *d_v, double *d_ans,
                        const phg_gmres_opts *opts, phg_gmres_info *infos,
                        phg_stream_t s);

/* ---- BarVortex implicit step (examples/barvortex.h:62-155,
 *      examples/barvortex_impl.h:108-310) --------------------------------- */
/* BarVortexConf without the callbacks: the Coriolis/orography term lh is
 * passed as a mesh function (proj(lh, m, conf.coriolis),
 * barvortex_impl.h:121-122), and the right-part callback conf.rp as the
 * mesh function of its values at the step's t (phg_barvortex_calc_rp,
 * phg_barvortex_load_rp). */
typedef struct {
    double tau, sigma, mu, k1, k2, theta;
} phg_bv_conf;
typedef struct {
    int max_iters;         /* fixed-point passes, reference: 1000              */
    double tol;            /* stop at ||u_n1 - u_n|| / ||u|| < tol, ref. 1e-14  */
    double rtol_mass;      /* mass solves of L() and of the Jacobian           */
    double rtol_helmholtz; /* A_ solve                                        */
    double rtol_poisson;   /* laplace_ solve (multigrid-preconditioned); its
                              FP64 residual floor is ~1e-13..1e-12 at levels
                              7-8 (stiffness rows sum to zero)               */
    int precision;         /* PHG_F64, or PHG_F32 / PHG_BF16: mass and Helmholtz
                              solves in FP32 Krylov with FP64 refinement      */
    int max_solver_iters;  /* cap of every inner solve                        */
    double stall_floor;    /* also end the step when the change is below this
                              and no longer contracting (rel >= half the value
                              two passes before; passes alternate):
                              the iterate then moves by rounding and solve-
                              tolerance noise only; 0 keeps the reference's
                              tol-only test                                   */
} phg_bv_opts;
typedef struct {
    int iterations;        /* fixed-point passes of the last step             */
    double rel_change;     /* ||u_n1 - u_n|| / ||u|| of the last pass         */
    int converged;         /* rel_change < tol, or stalled below stall_floor  */
    int stalled;           /* ended by the stall_floor rule                   */
    int diverged;          /* ended because the change grew past the state
                              (rel_change > 1 after 3 passes), 1000x above
                              its smallest value, or became NaN: the fixed
                              point does not contract, e.g. for grid-scale
                              vorticity at large tau |v| / h; u1 = u         */
    int solver_iterations; /* GMRES steps of the member's solves (all kinds)  */
    int it_mass, it_jacobian, it_helmholtz, it_poisson; /* the same by solve */
} phg_bv_info;
typedef struct phg_barvortex_s *phg_barvortex_t;

void phg_bv_default_opts(phg_bv_opts *o);
/* Device bytes an object with `slots` slots needs on a mesh of n_points
 * (estimate used for admission). */
size_t phg_barvortex_estimate_bytes(int n_points, int slots);
/* BarVortex(m, conf) for up to `slots` (1, 2, 4 or 8) members at once:
 * assembles A_, SphereLaplace's and SphereJacobian's matrices, SphereNorm's
 * matrix and the Poisson V-cycle hierarchy; d_lh: n_points values (device).
 * The boundary (pole) values of u and w are zero, as the reference drivers
 * pass them (examples/test_barvortex.cpp:445). */
int phg_barvortex_create(phg_mesh_t m, const phg_bv_conf *conf, const double *d_lh,
                         int slots, phg_stream_t s, phg_barvortex_t *out);
void phg_barvortex_destroy(phg_barvortex_t B);
int phg_barvortex_slots(phg_barvortex_t B);
/* BarVortex::calc(u1, u, 0, 0, t) for nmembers (<= slots) members: d_u and
 * d_u1 hold nmembers mesh functions of n_points values each, one after the
 * other.  infos: nmembers host entries (may be NULL). */
int phg_barvortex_calc(phg_barvortex_t B, int nmembers, const double *d_u, double *d_u1,
                       const phg_bv_opts *opts, phg_bv_info *infos, phg_stream_t s);
/* The same with the right part conf.rp (barvortex_impl.h:233-244): d_rp
 * holds, member after member, rp(x, y, t, mu, sigma) at every mesh point
 * (x, y: zone-0 chart coordinates, MeshPoint::x()/y()) for the step's t; it
 * is added to the fixed right side FC.  d_rp NULL = no right part. */
int phg_barvortex_calc_rp(phg_barvortex_t B, int nmembers, const double *d_u,
                          const double *d_rp, double *d_u1, const phg_bv_opts *opts,
                          phg_bv_info *infos, phg_stream_t s);
/* BarVortex::calc(u1, u, bnd_u, bnd_w, t) (barvortex_impl.h:171-310) with
 * the boundary values: d_bnd_u and d_bnd_w hold, member after member, the
 * new-time values of psi and of its Laplacian at the mesh's boundary points
 * (outer order, n_outer each); either may be NULL (zeros).  w = L(u) takes
 * bnd_w as its boundary values, bnd_ bnd_w is ADDED to the Helmholtz right
 * part as the reference does (:287-293), and the Poisson solve takes bnd_u
 * (SphereLaplace::solve, src/impl/slaplace_impl.h:84-130). */
int phg_barvortex_calc_bnd(phg_barvortex_t B, int nmembers, const double *d_u,
                           const double *d_rp, const double *d_bnd_u, const double *d_bnd_w,
                           double *d_u1, const phg_bv_opts *opts, phg_bv_info *infos,
                           phg_stream_t s);
/* The same step, slot by slot, for a caller that admits and retires members
 * while others run (continuous batching): load a free slot with u, then call
 * iterate (starts loaded slots, runs one fixed-point pass of the running
 * ones; state[k] = 0 free, 1 running, 2 done), and store a done slot's u1,
 * which frees it.  A member's result does not depend on the other slots. */
int phg_barvortex_load(phg_barvortex_t B, int slot, const double *d_u, phg_stream_t s);
/* load with the slot's right-part values (as phg_barvortex_calc_rp); NULL
 * d_rp = phg_barvortex_load. */
int phg_barvortex_load_rp(phg_barvortex_t B, int slot, const double *d_u,
                          const double *d_rp, phg_stream_t s);
/* After load or load_rp: the slot's boundary values (as
 * phg_barvortex_calc_bnd, n_outer each, NULL = zeros).  A later load of the
 * slot clears them. */
int phg_barvortex_load_bnd(phg_barvortex_t B, int slot, const double *d_bnd_u,
                           const double *d_bnd_w, phg_stream_t s);
int phg_barvortex_iterate(phg_barvortex_t B, const phg_bv_opts *opts, int *state,
                          phg_bv_info *infos, phg_stream_t s);
int phg_barvortex_store(phg_barvortex_t B, int slot, double *d_u1, phg_stream_t s);
/* Linearised and adjoint steps with zero boundary values (the _bnd forms
 * below take them), for nmembers
 * (<= slots) members at once; no step may be in flight on B.  Every solve
 * starts from zero; infos (may be NULL) report the GMRES steps by solve and
 * converged = all of the member's solves converged.
 *   calc_L   (barvortex_impl.h:326-400), L_step:  u1 = the step linearised
 *            about z, with the Jacobian terms explicit:
 *            laplace^-1 idt A_^-1 idt [ (1/tau - (1-theta) sigma) L(u)
 *            + (1-theta) mu L(L(u)) - (k2 J(u, lh) + k1 J(z, L(u)) + k1 J(u, L(z))) ]
 *   calc_L_1 (:402-467), L_1_step: its inverse form with Ab_
 *            (integrate_backward_cb, :83-103); as in the reference, the
 *            terms are ADDED to d_u1's content (L_1_step passes zeros)
 *   calc_LT  (:469-535), LT_step: the adjoint form, with J^T = calc1t.
 * d_u/d_v, d_z, d_u1/d_v1: nmembers mesh functions each, one after the
 * other. */
int phg_barvortex_calc_L(phg_barvortex_t B, int nmembers, const double *d_u,
                         const double *d_z, double *d_u1, const phg_bv_opts *opts,
                         phg_bv_info *infos, phg_stream_t s);
int phg_barvortex_calc_L_1(phg_barvortex_t B, int nmembers, const double *d_u,
                           const double *d_z, double *d_u1, const phg_bv_opts *opts,
                           phg_bv_info *infos, phg_stream_t s);
int phg_barvortex_calc_LT(phg_barvortex_t B, int nmembers, const double *d_v,
                          const double *d_z, double *d_v1, const phg_bv_opts *opts,
                          phg_bv_info *infos, phg_stream_t s);
/* The same steps with the members' boundary (pole) values of psi (d_bnd_u)
 * and of its Laplacian (d_bnd_w): outer_size values per member, one member
 * after the other; either may be NULL for zero.  As the reference uses
 * them (barvortex_impl.h:326-535): bnd_w gives the boundary values of
 * L(z) and L(u), and of the A_ (Ab_) solve's result, to whose right part
 * bnd_ bnd_w (bndb_ bnd_w for calc_L_1) is ADDED; bnd_u is the Poisson
 * solve's boundary value; calc_LT also takes bnd_w for both L(v1) and
 * ends with set_bnd(v1, bnd_u). */
int phg_barvortex_calc_L_bnd(phg_barvortex_t B, int nmembers, const double *d_u,
                             const double *d_z, const double *d_bnd_u, const double *d_bnd_w,
                             double *d_u1, const phg_bv_opts *opts, phg_bv_info *infos,
                             phg_stream_t s);
int phg_barvortex_calc_L_1_bnd(phg_barvortex_t B, int nmembers, const double *d_u,
                               const double *d_z, const double *d_bnd_u,
                               const double *d_bnd_w, double *d_u1, const phg_bv_opts *opts,
                               phg_bv_info *infos, phg_stream_t s);
int phg_barvortex_calc_LT_bnd(phg_barvortex_t B, int nmembers, const double *d_v,
                              const double *d_z, const double *d_bnd_u, const double *d_bnd_w,
                              double *d_v1, const phg_bv_opts *opts, phg_bv_info *infos,
                              phg_stream_t s);

/* ---- Baroclin implicit step (examples/baroclin.h:60-292,
 *      examples/baroclin.cpp:69-181, 445-606) ------------------------------ */
/* Baroclin's parameters without the callbacks: the Coriolis/orography term
 * is a mesh function (proj(lh, m, coriolis), baroclin.cpp:81-82) and the
 * forcing callbacks f, g are passed per call as mesh functions. */
typedef struct {
    double tau, sigma, mu, sigma1, mu1, alpha;
} phg_bc_conf;
typedef struct {
    int max_iters;         /* fixed-point passes, reference: 20                */
    double tol;            /* stop at max(dist(u1_n1, u1_n), dist(u2_n1, u2_n))
                              < tol (absolute, SphereNorm), reference: 1e-8   */
    double rtol_mass;      /* mass solves of L() and of the Jacobian           */
    double rtol_system;    /* the 4n x 4n solve of A_                          */
    int max_solver_iters;  /* cap of every inner solve                        */
    int precondition;      /* 1: block lower-triangular preconditioner of the
                              4n system (diagonal A11, A22; one Poisson V-cycle
                              per Laplace block); 0: plain GMRES, as the
                              reference's solve                              */
} phg_bc_opts;
typedef struct {
    int iterations;        /* fixed-point passes                               */
    double distance;       /* the last pass's max distance                     */
    int converged;         /* distance < tol                                   */
    int it_mass, it_jacobian, it_system; /* GMRES steps by solve              */
} phg_bc_info;
typedef struct phg_baroclin_s *phg_baroclin_t;

void phg_bc_default_opts(phg_bc_opts *o);
/* Baroclin::Baroclin (baroclin.cpp:69-103) for up to `slots` (1, 2, 4, 8)
 * members at once: A_ as nine GPU-assembled blocks (bit-identical to the
 * reference's integrate_cb values) stacked into the 4n x 4n CSR,
 * SphereLaplace's, SphereJacobian's and SphereNorm's matrices and the Poisson
 * V-cycle hierarchy.  d_lh: n_points values (device).  alpha != 0. */
int phg_baroclin_create(phg_mesh_t m, const phg_bc_conf *conf, const double *d_lh,
                        int slots, phg_stream_t s, phg_baroclin_t *out);
void phg_baroclin_destroy(phg_baroclin_t B);
int phg_baroclin_slots(phg_baroclin_t B);
/* A_ (owned by B). */
int phg_baroclin_matrix(phg_baroclin_t B, phg_matrix_t *out);
/* Baroclin::calc(u11, u21, u1, u2, bnd = 0, t) (baroclin.cpp:445-606) for
 * nmembers (<= slots) members: d_u1, d_u2, d_u11, d_u21 hold nmembers mesh
 * functions of n_points values each, one after the other; d_f, d_g (n_points
 * values each, may be NULL) are the forcing callbacks f_ and g_ evaluated at
 * the mesh points for this step's t, shared by the members.  A member's
 * result does not depend on the other members.  infos: nmembers host entries
 * (may be NULL). */
int phg_baroclin_calc(phg_baroclin_t B, int nmembers, const double *d_u1,
                      const double *d_u2, const double *d_f, const double *d_g,
                      double *d_u11, double *d_u21, const phg_bc_opts *opts,
                      phg_bc_info *infos, phg_stream_t s);
/* The same step with the boundary (pole) values: d_bnd holds, member after
 * member, n_outer values (outer order) that are the boundary values of w1,
 * w2, u1 and u2 alike (baroclin.cpp:494-498, 541-544, 587-590); right_part_cb
 * subtracts the boundary columns of the blocks times them (:259-323).  NULL
 * = zeros. */
int phg_baroclin_calc_bnd(phg_baroclin_t B, int nmembers, const double *d_u1,
                          const double *d_u2, const double *d_f, const double *d_g,
                          const double *d_bnd, double *d_u11, double *d_u21,
                          const phg_bc_opts *opts, phg_bc_info *infos, phg_stream_t s);
/* The linearised steps about (z1, z2), zero boundary values, for nmembers
 * (<= slots) members; all arrays hold nmembers mesh functions one after the
 * other.  The Jacobian arguments are the reference code's (two of them,
 * J(L z1, avg w2) and J(L z2, avg u1), differ from its comments).
 *   calc_L   (examples/baroclin.cpp:607-802, L_step): one pass with A_ and the
 *            explicit right side w/tau - 0.5 (1-theta) sigma (w1 -+ w2) + ...;
 *   calc_L_1 (:804-986, L_1_step): up to 10 passes with Ab_
 *            (integrate_backward_cb, :184-244, assembled on the first call)
 *            until max(dist) < 1e-8.
 * infos: passes, last distance, converged (every solve converged and, for
 * calc_L_1, the distance below 1e-8).  The reference's calc_LT has an
 * empty body (:988-993), so its LT_step (:1009-1014) leaves the outputs as
 * they were: there is no device counterpart, and the C++ class's LT_step
 * does the same. */
int phg_baroclin_calc_L(phg_baroclin_t B, int nmembers, const double *d_u1,
                        const double *d_u2, const double *d_z1, const double *d_z2,
                        double *d_u11, double *d_u21, const phg_bc_opts *opts,
                        phg_bc_info *infos, phg_stream_t s);
int phg_baroclin_calc_L_1(phg_baroclin_t B, int nmembers, const double *d_u1,
                          const double *d_u2, const double *d_z1, const double *d_z2,
                          double *d_u11, double *d_u21, const phg_bc_opts *opts,
                          phg_bc_info *infos, phg_stream_t s);
/* calc_L and calc_L_1 with the boundary (pole) values (Baroclin::calc_L /
 * calc_L_1 with a non-NULL bnd, examples/baroclin.cpp:620-802, 804-986): d_bnd
 * holds, member after member, n_outer values (outer order) used as the
 * reference uses its one bnd array: the boundary values of every Laplacian
 * (u, w and z: :675-682, :859-866), of the right part's boundary branch for
 * w1, w2, u1 and u2 (A_'s elements for calc_L, :708-711 with right_part_cb;
 * Ab_'s for calc_L_1, :892-895 with right_part_backward_cb, :341-436) and of
 * the four p2u (:783-786, :967-970).  NULL = zeros, the plain entries'
 * results. */
int phg_baroclin_calc_L_bnd(phg_baroclin_t B, int nmembers, const double *d_u1,
                            const double *d_u2, const double *d_z1, const double *d_z2,
                            const double *d_bnd, double *d_u11, double *d_u21,
                            const phg_bc_opts *opts, phg_bc_info *infos, phg_stream_t s);
int phg_baroclin_calc_L_1_bnd(phg_baroclin_t B, int nmembers, const double *d_u1,
                              const double *d_u2, const double *d_z1, const double *d_z2,
                              const double *d_bnd, double *d_u11, double *d_u21,
                              const phg_bc_opts *opts, phg_bc_info *infos, phg_stream_t s);

/* ---- serving: implicit-step requests of ensemble members on several mesh
 *      levels, continuously batched (no counterpart in the reference, which
 *      steps one BarVortex at a time, examples/test_barvortex.cpp:441-470;
 *      every request is one BarVortex::calc of its member) --------------- */
#define PHG_SERVE_MAXLEVELS 8
typedef struct {
    int n_levels;
    int levels[PHG_SERVE_MAXLEVELS];  /* mke_mesh_sphere levels served        */
    int slots[PHG_SERVE_MAXLEVELS];   /* slots per level (1, 2, 4, 8); 0: the
                                         largest that fits the budget        */
    size_t hbm_budget;                /* device bytes for engines + members   */
    double coriolis;                  /* lh = coriolis * sin(latitude)        */
    phg_bv_conf conf;
    phg_bv_opts opts;
    int verbose;                      /* 1: a stderr line per finished step  */
    int substeps[PHG_SERVE_MAXLEVELS]; /* per level: a request runs this many
                                         BarVortex::calc steps of tau/k
                                         (0 or 1: one step of tau); for levels
                                         where the fixed point does not
                                         contract at tau                      */
} phg_serve_config;
typedef struct {
    int completed;
    long passes;            /* fixed-point passes run, all levels           */
    double wall_s, throughput;            /* steps/s over the run           */
    double p50_latency_s, p99_latency_s, mean_latency_s;  /* finish - arrival */
    double p50_service_s;   /* finish - admission                           */
    double mean_occupancy;  /* members per fixed-point pass                 */
    size_t device_bytes;    /* estimated engine + member-state bytes        */
} phg_serve_stats;
typedef struct phg_server_s *phg_server_t;

/* Levels 5-8, budget 200 GiB, BarVortexConf of the ensemble workload in units
 * of 1/Earth's angular velocity: tau = 600 s, sigma 1.6e-2, mu 8e-5,
 * k1 = k2 = 1, theta 1/2, lh = 2 sin(latitude); default phg_bv_opts with at
 * most 100 fixed-point passes per step. */
void phg_serve_default_config(phg_serve_config *c);
int phg_serve_create(const phg_serve_config *cfg, phg_server_t *out);
void phg_serve_destroy(phg_server_t S);
int phg_serve_slots(phg_server_t S, int level);
/* A member on `level` with initial stream function h_u (host, n_points
 * values); its state stays on the device.  PHG_ERR_NOMEM over budget. */
int phg_serve_add_member(phg_server_t S, int level, const double *h_u, int *member_id);
int phg_serve_member_state(phg_server_t S, int member, double *h_u);
/* Serves n requests (member[r], released at arrival_s[r] seconds after the
 * call starts, wall clock) and returns when all are done; start_s/finish_s
 * (host, n entries) receive admission and completion times, infos (may be
 * NULL) each step's fixed-point info. */
int phg_serve_run(phg_server_t S, int n, const int *member, const double *arrival_s,
                  double *start_s, double *finish_s, phg_bv_info *infos,
                  phg_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PHELM_GPU_H */
