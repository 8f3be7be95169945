/*
 * CPU oracle: CSR SpMV and restarted GMRES(m) in FP64.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).
 *
 * The reference calls Solver<T>::mult_vector / solve from its solver library
 * (call sites src/impl/slaplace_impl.h:109,168-171, src/mesh.h:740-746), which
 * is not in the tree.  This is a textbook GMRES (Saad & Schultz 1986):
 * Arnoldi with modified Gram-Schmidt, Givens rotations on the Hessenberg
 * matrix, x0 from the caller, restart after m steps, stop on the true FP64
 * residual ||b - A x||_2 / ||b||_2 <= rtol.  Parity of the iterates is
 * unpinned; tests pin it by residuals and by a dense/scipy solve.
 * OpenMP parallelises SpMV and the BLAS-1 loops; this is also the timed CPU
 * baseline.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void or_spmv(int n, const int *row_ptr, const int *col_idx, const double *vals,
             const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            s += vals[e] * x[col_idx[e]];
        y[i] = s;
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void residual(int n, const int *rp, const int *ci, const double *v,
                     const double *b, const double *x, double *r)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) s += v[e] * x[ci[e]];
        r[i] = b[i] - s;
    }
}

int or_gmres(int n, const int *rp, const int *ci, const double *vals,
             const double *b, double *x, int m, double rtol, int max_iters,
             or_gmres_info *info)
{
    double *V = (double *)malloc((size_t)(m + 1) * n * sizeof(double));
    double *w = (double *)malloc((size_t)n * sizeof(double));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double *)calloc(m, sizeof(double));
    double *sn = (double *)calloc(m, sizeof(double));
    double *g = (double *)calloc(m + 1, sizeof(double));
    double *y = (double *)calloc(m, sizeof(double));
    if (!V || !w || !H || !cs || !sn || !g || !y) return -1;
#define HH(i, j) H[(size_t)(i) * m + (j)]

    double bnorm = sqrt(dot(n, b, b));
    if (bnorm == 0.0) bnorm = 1.0;
    int iters = 0, cycles = 0, converged = 0;
    double rel = 0.0;

    for (;;) {
        residual(n, rp, ci, vals, b, x, V);
        double beta = sqrt(dot(n, V, V));
        rel = beta / bnorm;
        if (rel <= rtol) { converged = 1; break; }
        if (iters >= max_iters) break;
        double inv = 1.0 / beta;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n; ++k) V[k] *= inv;
        memset(g, 0, (m + 1) * sizeof(double));
        g[0] = beta;
        int j, steps = 0;
        for (j = 0; j < m && iters < max_iters; ++j) {
            double *vj = V + (size_t)j * n;
            or_spmv(n, rp, ci, vals, vj, w);
            for (int i = 0; i <= j; ++i) {          /* modified Gram-Schmidt */
                const double *vi = V + (size_t)i * n;
                double h = dot(n, w, vi);
                HH(i, j) = h;
#pragma omp parallel for schedule(static)
                for (int k = 0; k < n; ++k) w[k] -= h * vi[k];
            }
            double hn = sqrt(dot(n, w, w));
            HH(j + 1, j) = hn;
            double *vn = V + (size_t)(j + 1) * n;
            double hinv = 1.0 / hn;
#pragma omp parallel for schedule(static)
            for (int k = 0; k < n; ++k) vn[k] = w[k] * hinv;
            for (int i = 0; i < j; ++i) {           /* previous rotations */
                double t = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                HH(i, j) = t;
            }
            double a = HH(j, j), c = HH(j + 1, j);
            double r = hypot(a, c);
            cs[j] = a / r;
            sn[j] = c / r;
            HH(j, j) = r;
            HH(j + 1, j) = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++iters;
            ++steps;
            if (fabs(g[j + 1]) / bnorm <= rtol) { ++j; break; }
        }
        int kdim = steps;
        for (int i = kdim - 1; i >= 0; --i) {       /* back substitution */
            double s = g[i];
            for (int l = i + 1; l < kdim; ++l) s -= HH(i, l) * y[l];
            y[i] = s / HH(i, i);
        }
#pragma omp parallel for schedule(static)
        for (int k = 0; k < n; ++k) {
            double s = 0.0;
            for (int i = 0; i < kdim; ++i) s += V[(size_t)i * n + k] * y[i];
            x[k] += s;
        }
        ++cycles;
    }
#undef HH
    if (info) {
        info->iterations = iters;
        info->restarts = cycles;
        info->rel_residual = rel;
        info->converged = converged;
    }
    free(V); free(w); free(H); free(cs); free(sn); free(g); free(y);
    return converged ? 0 : 1;
}
