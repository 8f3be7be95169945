// SphereJacobian (src/sjacobian.h:65-149, src/sjacobian.cpp:136-468) on
// gfx950.
//
// The reference's calc2 runs calc2_1..calc2_4, each 4 x (SpMV with a
// derivative matrix + SpMV with its boundary matrix + mass solve with idt_),
// 16 solves in all, with only 8 distinct right-hand sides: idt^-1 (D u_in +
// D_rp u_bnd) for the four derivative operators applied to u and to v.  Here
// one call evaluates the Jacobian of up to 8 (u, v) pairs:
//   1. gather the inner and boundary values of every u_k, v_k;
//   2. one pass over the shared stencil applies the four derivative matrices
//      (values interleaved per entry, one column read) and their boundary
//      matrices to every gathered vector: 8 right-hand sides per pair;
//   3. one batched GMRES with idt_ solves all 8 * npairs systems, each with
//      its own Krylov space and stopping test, sharing the matrix reads;
//   4. an elementwise pass forms calc2 = 0.25 * (((c1 + c2) + c3) + c4) in
//      the reference's order (c_k = pt1 - pt2, pt = solve * solve).
// Column q = 8 k + 4 c + op of the solve holds pair k, vector c (0 = u,
// 1 = v), operator op (0 diff1, 1 diff2, 2 diff1_cos, 3 diff2_cos).
#include <cstring>

#include "ell.h"
#include "phg_internal.h"

using namespace phg;

constexpr int JK_MAXPAIRS = 8;

struct phg_jacobian_s {
    phg_mesh_s* mesh = nullptr;
    int n = 0, os = 0, width = 0;
    phg_matrix_s* idt = nullptr;        // JAC_MASS (id_cb)
    phg_matrix_s* D[4] = {};            // diff1_, diff2_, diff1_cos_, diff2_cos_
    phg_matrix_s* Db[4] = {};           // diff*_rp_ (inner x outer)
    DevBuf<int> col;                    // slot-major ELL columns of the shared stencil
    DevBuf<double> val4;                // [slot][row][4]
    DevBuf<int> brp, bci;               // boundary CSR (shared pattern)
    DevBuf<double> bval4;               // [entry][4]
    DevBuf<double> Uin, Ubnd, B, X;     // work, sized for JK_MAXPAIRS pairs
    ~phg_jacobian_s()
    {
        for (auto* m : D) delete m;
        for (auto* m : Db) delete m;
        delete idt;
    }
};

namespace {

__global__ void jk_interleave(int n, int width, const double* __restrict__ v0,
                              const double* __restrict__ v1, const double* __restrict__ v2,
                              const double* __restrict__ v3, double* __restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * width) return;
    out[e * 4 + 0] = v0[e];
    out[e * 4 + 1] = v1[e];
    out[e * 4 + 2] = v2[e];
    out[e * 4 + 3] = v3[e];
}

// Uin[j][2k + c] = (c ? v_k : u_k)[inner[j]], Ubnd[o][2k + c] likewise at
// outer[o]; point p of pair k is read at p * su + k * sk.
__global__ void jk_gather(int n, int os, int np, size_t su, size_t sk,
                          const int* __restrict__ inner, const int* __restrict__ outer,
                          const double* __restrict__ u, const double* __restrict__ v,
                          double* __restrict__ Uin, double* __restrict__ Ubnd)
{
    const int P = 2 * np;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot_in = (size_t)n * P;
    if (e < tot_in) {
        const int j = (int)(e / P), p = (int)(e % P);
        const double* src = (p & 1) ? v : u;
        Uin[e] = src[(size_t)inner[j] * su + (size_t)(p >> 1) * sk];
    } else if (e < tot_in + (size_t)os * P) {
        const size_t f = e - tot_in;
        const int o = (int)(f / P), p = (int)(f % P);
        const double* src = (p & 1) ? v : u;
        Ubnd[f] = src[(size_t)outer[o] * su + (size_t)(p >> 1) * sk];
    }
}

// B[i][8k + 4c + op] = (D_op,rp u_bnd)_i + (D_op u_in)_i, slot order, no fma
// (vec_sum(rp, rp, tmp) of sjacobian.cpp).  Thread per (row, gathered vector).
__global__ void __launch_bounds__(256)
jk_rhs(int n, int width, int np, const int* __restrict__ col, const double* __restrict__ val4,
       const int* __restrict__ brp, const int* __restrict__ bci,
       const double* __restrict__ bval4, const double* __restrict__ Uin,
       const double* __restrict__ Ubnd, double* __restrict__ B)
{
    const int P = 2 * np;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * P) return;
    const int row = (int)(e / P), p = (int)(e % P);
    int c[ELL_MAXW];
    double x[ELL_MAXW];
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) c[k] = k < width ? col[(size_t)k * n + row] : 0;
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) x[k] = k < width ? Uin[(size_t)c[k] * P + p] : 0.0;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) {
        if (k < width) {
            const double* v = val4 + ((size_t)k * n + row) * 4;
#pragma unroll
            for (int op = 0; op < 4; ++op) s[op] += v[op] * x[k];
        }
    }
    double b[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = brp[row]; q < brp[row + 1]; ++q) {
        const double ub = Ubnd[(size_t)bci[q] * P + p];
#pragma unroll
        for (int op = 0; op < 4; ++op) b[op] += bval4[(size_t)q * 4 + op] * ub;
    }
    const int k = p >> 1, cc = p & 1;
    double* out = B + (size_t)row * (8 * np) + 8 * k + 4 * cc;
#pragma unroll
    for (int op = 0; op < 4; ++op) out[op] = b[op] + s[op];
}

// ans_k[i] = 0.25 * (((c1 + c2) + c3) + c4) * sign (sjacobian.cpp:136-362),
// stored at i * ai + k * ak.
__global__ void jk_combine(int n, int np, const double* __restrict__ X, double sign,
                           double* __restrict__ ans, size_t ai, size_t ak)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * np) return;
    const int k = (int)(e / n), i = (int)(e % n);
    const double* x = X + (size_t)i * (8 * np) + 8 * k;
    const double u1 = x[0], u2 = x[1], u1c = x[2], u2c = x[3];
    const double v1 = x[4], v2 = x[5], v1c = x[6], v2c = x[7];
    const double c1 = u2c * v1 - u1c * v2;
    const double c2 = u2 * v1c - u1 * v2c;
    const double c3 = u2c * v1 - u1 * v2c;
    const double c4 = u2 * v1c - u1c * v2;
    double a = c1;
    a = a + c2;
    a = a + c3;
    a = a + c4;
    a = a * 0.25;
    if (sign < 0.0) a = a * -1.0;
    ans[(size_t)i * ai + (size_t)k * ak] = a;
}

// Zeroes the mass-solve solutions of the pairs k with bit k of `bits` set
// (X is n x 8 np row-major, pair k in columns 8k..8k+7).
__global__ void jk_clear(int n, int np, unsigned bits, double* __restrict__ X)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * 8 * np) return;
    const int k = (int)(e % (8 * np)) >> 3;
    if (bits >> k & 1u) X[e] = 0.0;
}

bool same_structure(const phg_matrix_s* a, const phg_matrix_s* b)
{
    return a->n_rows == b->n_rows && a->n_cols == b->n_cols && a->h_rp == b->h_rp &&
           a->h_ci == b->h_ci;
}

}  // namespace

namespace phg {

// Zeroes the stored mass-solve solutions of the pairs in `bits` (bit k =
// pair k of an np-pair call), so that a later call with opts->use_x0 starts
// those pairs from zero and the others from their previous solutions.
int jacobian_clear_pairs(phg_jacobian_s* J, int np, unsigned bits, hipStream_t s)
{
    if (!J || np < 1 || np > JK_MAXPAIRS) {
        set_error("jacobian_clear_pairs: bad argument");
        return PHG_ERR_ARG;
    }
    const size_t t = (size_t)J->n * 8 * np;
    if (t && bits)
        hipLaunchKernelGGL(jk_clear, dim3((unsigned)((t + 255) / 256)), dim3(256), 0, s, J->n, np,
                           bits, J->X.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// calc2 (sign 1) or calc2t (sign -1) of np pairs; point p of pair k is read
// at p * su + k * sk of u and v, J(u_k, v_k) at inner point i is stored at
// i * ai + k * ak of ans.  With opts->use_x0 the 8 np mass solves start
// from the solutions of the previous call (same np), which the object keeps.
int jacobian_calc2_strided(phg_jacobian_s* J, int np, const double* d_u, const double* d_v,
                           size_t su, size_t sk, double* d_ans, size_t ai, size_t ak,
                           double sign, const phg_gmres_opts* opts, phg_gmres_info* infos,
                           hipStream_t s)
{
    if (!J || !d_u || !d_v || !d_ans || np < 1 || np > JK_MAXPAIRS) {
        set_error("phg_jacobian_calc2: bad argument (1 <= npairs <= 8)");
        return PHG_ERR_ARG;
    }
    const int n = J->n, os = J->os, P = 2 * np;
    phg_gmres_opts o;
    if (opts) o = *opts;
    else {
        phg_gmres_default_opts(&o);
        o.ortho = PHG_CGS1;
        o.restart = MASS_RESTART;
        o.rtol = 1e-13;
        o.diag_scale = 1;
    }
    const size_t gat = (size_t)(n + os) * P;
    hipLaunchKernelGGL(jk_gather, dim3((unsigned)((gat + 255) / 256)), dim3(256), 0, s, n, os, np,
                       su, sk, J->mesh->d_inner.p, J->mesh->d_outer.p, d_u, d_v, J->Uin.p,
                       J->Ubnd.p);
    const size_t tr = (size_t)n * P;
    hipLaunchKernelGGL(jk_rhs, dim3((unsigned)((tr + 255) / 256)), dim3(256), 0, s, n, J->width,
                       np, J->col.p, J->val4.p, J->brp.p, J->bci.p, J->bval4.p, J->Uin.p,
                       J->Ubnd.p, J->B.p);
    PHG_HIP(hipGetLastError());
    int rc = gmres_batch_solve(J->idt, 8 * np, J->B.p, J->X.p, &o, infos, s);
    if (rc < 0) return rc;
    const size_t tc = (size_t)n * np;
    hipLaunchKernelGGL(jk_combine, dim3((unsigned)((tc + 255) / 256)), dim3(256), 0, s, n, np,
                       J->X.p, sign, d_ans, ai, ak);
    PHG_HIP(hipGetLastError());
    return rc;
}

}  // namespace phg

extern "C" {

int phg_jacobian_create(phg_mesh_t m, phg_stream_t stream, phg_jacobian_t* out)
{
    if (!m || !out) {
        set_error("phg_jacobian_create: null argument");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<phg_jacobian_s> J(new phg_jacobian_s());
    J->mesh = m;
    int rc = assemble_device(m, PHG_JAC_MASS, 1.0, 0.0, s, &J->idt);
    if (rc) return rc;
    for (int op = 0; op < 4; ++op) {
        rc = assemble_device(m, PHG_DIFF1 + op, 1.0, 0.0, s, &J->D[op]);
        if (rc) return rc;
        rc = assemble_boundary_device(m, PHG_DIFF1 + op, 1.0, 0.0, s, &J->Db[op]);
        if (rc) return rc;
        rc = ensure_ell64(J->D[op], s);
        if (rc) return rc;
    }
    for (int op = 1; op < 4; ++op)
        if (!same_structure(J->D[0], J->D[op]) || !same_structure(J->Db[0], J->Db[op])) {
            set_error("phg_jacobian_create: derivative matrices differ in structure");
            return PHG_ERR_ARG;
        }
    const int n = J->D[0]->n_rows;
    J->n = n;
    J->os = (int)m->h.outer.size();
    J->width = J->D[0]->ell64.width;
    if (J->width > ELL_MAXW) {
        set_error("phg_jacobian_create: rows longer than 8 entries");
        return PHG_ERR_ARG;
    }
    const size_t ne = (size_t)n * J->width;
    PHG_HIP(J->col.alloc(ne));
    PHG_HIP(hipMemcpyAsync(J->col.p, J->D[0]->ell64.col.p, ne * sizeof(int),
                           hipMemcpyDeviceToDevice, s));
    PHG_HIP(J->val4.alloc(ne * 4));
    if (ne)
        hipLaunchKernelGGL(jk_interleave, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, n,
                           J->width, J->D[0]->ell64.val.p, J->D[1]->ell64.val.p,
                           J->D[2]->ell64.val.p, J->D[3]->ell64.val.p, J->val4.p);
    const size_t nb = J->Db[0]->h_ci.size();
    PHG_HIP(J->brp.upload(J->Db[0]->h_rp.data(), n + 1, s));
    PHG_HIP(J->bci.upload(J->Db[0]->h_ci.data(), nb, s));
    PHG_HIP(J->bval4.alloc(std::max<size_t>(nb, 1) * 4));
    if (nb)
        hipLaunchKernelGGL(jk_interleave, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s,
                           (int)nb, 1, J->Db[0]->val.p, J->Db[1]->val.p, J->Db[2]->val.p,
                           J->Db[3]->val.p, J->bval4.p);
    PHG_HIP(hipGetLastError());
    rc = ensure_mesh_device(m);
    if (rc) return rc;
    PHG_HIP(J->Uin.alloc((size_t)n * 2 * JK_MAXPAIRS));
    PHG_HIP(J->Ubnd.alloc(std::max<size_t>((size_t)J->os * 2 * JK_MAXPAIRS, 1)));
    PHG_HIP(J->B.alloc((size_t)n * 8 * JK_MAXPAIRS));
    PHG_HIP(J->X.alloc((size_t)n * 8 * JK_MAXPAIRS));
    PHG_HIP(hipStreamSynchronize(s));
    *out = J.release();
    return PHG_OK;
}

void phg_jacobian_destroy(phg_jacobian_t J) { delete J; }

int phg_jacobian_matrix(phg_jacobian_t J, int which, phg_matrix_t* out)
{
    if (!J || !out || which < 0 || which > 8) {
        set_error("phg_jacobian_matrix: bad argument");
        return PHG_ERR_ARG;
    }
    *out = which == 0 ? J->idt : which <= 4 ? J->D[which - 1] : J->Db[which - 5];
    return PHG_OK;
}

int phg_jacobian_calc2(phg_jacobian_t J, int npairs, const double* d_u, const double* d_v,
                       double* d_ans, const phg_gmres_opts* opts, phg_gmres_info* infos,
                       phg_stream_t s)
{
    const size_t V = J ? J->mesh->h.is_boundary.size() : 0, n = J ? J->n : 0;
    return jacobian_calc2_strided(J, npairs, d_u, d_v, 1, V, d_ans, 1, n, 1.0, opts, infos,
                                  (hipStream_t)s);
}

int phg_jacobian_calc2t(phg_jacobian_t J, int npairs, const double* d_u, const double* d_v,
                        double* d_ans, const phg_gmres_opts* opts, phg_gmres_info* infos,
                        phg_stream_t s)
{
    const size_t V = J ? J->mesh->h.is_boundary.size() : 0, n = J ? J->n : 0;
    return jacobian_calc2_strided(J, npairs, d_u, d_v, 1, V, d_ans, 1, n, -1.0, opts, infos,
                                  (hipStream_t)s);
}

}  // extern "C"
