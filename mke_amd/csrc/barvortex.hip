// BarVortex::calc (examples/barvortex_impl.h:171-310) device resident, for S
// slots (ensemble members on one mesh) at once.
//
// One implicit step of the barotropic vorticity equation: with w = L(u)
// (SphereLaplace::calc1) and FC = w/tau + mu(1-theta) L(w) - sigma(1-theta) w,
// iterate from u_n = w_n = 0
//     tmp1 = k1 theta w_n + k1 (1-theta) w + k2 lh
//     tmp2 = theta u_n + (1-theta) u
//     F    = FC_inner - J(tmp2, tmp1)                     (SphereJacobian::calc2)
//     w_n  = p2u(A^-1 (idt F))                            (phelm::solve, Helmholtz A)
//     u_n1 = p2u(laplace^-1 (idt u2p(w_n)))               (SphereLaplace::solve)
//     stop when ||u_n1 - u_n|| / ||u|| < tol (SphereNorm, src/norm.h:58-120)
// as the reference does, with the boundary (pole) values zero as its drivers
// pass them (examples/test_barvortex.cpp:445).
//
// Layout: every per-member vector is V x S (or n x S) row-major, slot fastest,
// so each solve is one batched GMRES over the S slots, each slot with its own
// Krylov space and stopping test.  A slot that is empty, not yet started or
// finished has zero columns in every solve, which the batched solver skips;
// a slot's arithmetic never depends on the other slots, so a member gives the
// same bits alone or in any mix.  The Poisson solve is preconditioned by the
// multigrid V-cycle; the Helmholtz, Jacobian and L() mass solves are plain
// batched GMRES (mass-dominated, well conditioned).  The Helmholtz and
// Poisson solves start from the current iterate (w_n, u_n): once the iterate
// stops moving they return it unchanged and the step ends.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ell.h"
#include "phg_internal.h"
#include "slot_kernels.h"

namespace phg {
int jacobian_calc2_strided(phg_jacobian_s* J, int np, const double* d_u, const double* d_v,
                           size_t su, size_t sk, double* d_ans, size_t ai, size_t ak,
                           double sign, const phg_gmres_opts* opts, phg_gmres_info* infos,
                           hipStream_t s);
int jacobian_clear_pairs(phg_jacobian_s* J, int np, unsigned bits, hipStream_t s);
}  // namespace phg

using namespace phg;


struct phg_barvortex_s {
    phg_mesh_s* mesh = nullptr;
    int V = 0, n = 0, os = 0, S = 0;
    phg_bv_conf conf{};
    phg_matrix_s* idt = nullptr;    // SphereLaplace::idt_ (MASS)
    phg_matrix_s* bnd3 = nullptr;   // SphereLaplace::bnd3_ (slaplace_integrate_cb, boundary)
    phg_matrix_s* A = nullptr;      // BarVortex::A_ (Helmholtz)
    phg_matrix_s* Ab = nullptr;     // BarVortex::Ab_ (backward; built on first calc_L_1)
    phg_matrix_s* bndA = nullptr;   // BarVortex::bnd_ (integrate_cb, boundary; first load_bnd)
    phg_matrix_s* bndAb = nullptr;  // BarVortex::bndb_ (integrate_backward_cb, boundary)
    phg_matrix_s* bnd2 = nullptr;   // SphereLaplace::bnd2_ (laplace_bnd2_cb; first load_bnd)
    phg_matrix_s* norm = nullptr;   // SphereNorm::NORM_ (full mass, V x V)
    phg_jacobian_t J = nullptr;
    phg_mg_t mg = nullptr;          // Poisson: laplace_ and its V-cycle
    DevBuf<double> lh;              // V
    DevBuf<double> U, W, FC, Un, Wn, Un1, T1, T2;  // V x S
    DevBuf<double> JAC, F, RP, X, Y, Z, Q;          // n x S
    DevBuf<double> Ub;                              // outer x S
    DevBuf<double> RPF;                             // V x S right parts (conf.rp values)
    DevBuf<double> BU, BW;                          // outer x S boundary values of psi, lap psi
    DevBuf<double> part, nrm;       // reductions
    DevBuf<int> dmask;
    int G = 0;
    std::vector<int> status, iters;                 // per slot
    std::vector<int> it_mass, it_jac, it_helm, it_pois;  // GMRES steps per slot and solve
    std::vector<double> nr0, rel, rel_prev, rel_prev2, rel_min;
    std::vector<int> stalled, diverged;
    std::vector<int> has_rp;                        // slot loaded with a right part
    std::vector<int> has_bnd;                       // slot loaded with boundary values
    std::vector<int> solve_fail;                    // a solve of the slot did not converge
    ~phg_barvortex_s()
    {
        delete idt;
        delete bnd3;
        delete A;
        delete Ab;
        delete bndA;
        delete bndAb;
        delete bnd2;
        delete norm;
        if (J) phg_jacobian_destroy(J);
        if (mg) phg_mg_destroy(mg);
    }
};

namespace {

bool any_bnd(const phg_barvortex_s* B, const std::vector<int>& mask)
{
    for (int k = 0; k < B->S; ++k)
        if (mask[k] && B->has_bnd[k]) return true;
    return false;
}

int upload_mask(phg_barvortex_s* B, const std::vector<int>& m, hipStream_t s)
{
    PHG_HIP(hipMemcpyAsync(B->dmask.p, m.data(), sizeof(int) * B->S, hipMemcpyHostToDevice, s));
    return PHG_OK;
}

void spmm(phg_barvortex_s* B, const phg_matrix_s* M, const double* X, double* Y, int accumulate,
          hipStream_t s)
{
    hipLaunchKernelGGL(bv_spmm, g256((size_t)M->n_rows * B->S), dim3(BV_BS), 0, s, M->n_rows,
                       B->S, M->rp.p, M->ci.p, M->val.p, X, Y, accumulate, B->dmask.p);
}

// out = sqrt(d^T NORM d) per slot, d = a - b; copied to host (synchronises).
int norms(phg_barvortex_s* B, const double* a, const double* b, std::vector<double>& out,
          hipStream_t s)
{
    const int S = B->S;
    hipLaunchKernelGGL(bv_normsq, dim3(B->G), dim3(BV_BS), 0, s, B->V, S, B->norm->rp.p,
                       B->norm->ci.p, B->norm->val.p, a, b, B->part.p, B->G, B->dmask.p);
    hipLaunchKernelGGL(bv_sum_cols, dim3(1), dim3(64 * BV_MAXS), 0, s, S, B->part.p, B->G,
                       B->nrm.p);
    PHG_HIP(hipGetLastError());
    out.assign(S, 0.0);
    PHG_HIP(hipMemcpyAsync(out.data(), B->nrm.p, sizeof(double) * S, hipMemcpyDeviceToHost, s));
    PHG_HIP(hipStreamSynchronize(s));
    return PHG_OK;
}

phg_gmres_opts solve_opts(double rtol, int precision, int ortho, int use_x0, int max_iters)
{
    phg_gmres_opts o;
    phg_gmres_default_opts(&o);
    o.restart = 30;
    o.rtol = rtol;
    o.ortho = ortho;
    o.precision = precision;
    o.inner_rtol = 1e-4;
    o.use_x0 = use_x0;
    o.max_iters = max_iters;
    return o;
}

int count_iters(const std::vector<phg_gmres_info>& inf, int S, std::vector<int>& acc, int stride)
{
    for (int s = 0; s < S; ++s)
        for (int k = 0; k < stride; ++k) acc[s] += inf[(size_t)s * stride + k].iterations;
    return 0;
}


enum { SLOT_FREE = 0, SLOT_NEW = 1, SLOT_RUN = 2, SLOT_DONE = 3 };

void zero_cols(phg_barvortex_s* B, size_t rows, double* x, hipStream_t s)
{
    hipLaunchKernelGGL(bv_zero_cols, g256(rows * B->S), dim3(BV_BS), 0, s, rows, B->S, x,
                       B->dmask.p);
}

int batched(phg_barvortex_s* B, phg_matrix_s* A, const double* rhs, double* x, double rtol,
            int precision, int ortho, int use_x0, phg_mg_s* pc, int max_iters,
            std::vector<int>& acc, hipStream_t s, int diag_scale = 0, int restart = 30)
{
    phg_gmres_opts o = solve_opts(rtol, precision, ortho, use_x0, max_iters);
    o.diag_scale = diag_scale;
    o.restart = restart;
    std::vector<phg_gmres_info> inf(B->S);
    const int rc = gmres_batch_solve(A, B->S, rhs, x, &o, inf.data(), s, pc);
    if (rc < 0) return rc;
    count_iters(inf, B->S, acc, 1);
    for (int k = 0; k < B->S; ++k)
        if (!inf[k].converged) B->solve_fail[k] = 1;
    return PHG_OK;
}

// SphereLaplace::calc1 (src/impl/slaplace_impl.h:153-192) for the masked
// slots: out = p2u(idt^-1 (laplace u_in + bnd3 u_bnd), bnd) (bnd: outer x S,
// or null for 0).
int calc1(phg_barvortex_s* B, const double* u, double* out, const phg_bv_opts& o,
          std::vector<int>& acc, hipStream_t s, const double* bnd = nullptr)
{
    const int S = B->S, n = B->n;
    phg_matrix_t lap = nullptr;
    phg_mg_matrix(B->mg, &lap);
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       u, B->Y.p, B->dmask.p);
    if (B->os)
        hipLaunchKernelGGL(bv_outer, g256((size_t)B->os * S), dim3(BV_BS), 0, s, B->os, S,
                           B->mesh->d_outer.p, u, B->Ub.p, B->dmask.p);
    spmm(B, lap, B->Y.p, B->Q.p, 0, s);
    if (B->os) spmm(B, B->bnd3, B->Ub.p, B->Q.p, 1, s);
    zero_cols(B, n, B->Q.p, s);
    PHG_HIP(hipGetLastError());
    int rc = batched(B, B->idt, B->Q.p, B->X.p, o.rtol_mass, o.precision, PHG_CGS1, 0, nullptr,
                     o.max_solver_iters, acc, s, 1, MASS_RESTART);
    if (rc) return rc;
    if (bnd)
        hipLaunchKernelGGL(bv_p2u_bnd, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->X.p, bnd, out, B->dmask.p);
    else
        hipLaunchKernelGGL(bv_p2u, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->X.p, out, B->dmask.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// Start of BarVortex::calc for the slots in `mask` (barvortex_impl.h:175-245).
int init_slots(phg_barvortex_s* B, const std::vector<int>& mask, const phg_bv_opts& o,
               hipStream_t s)
{
    const int S = B->S;
    const size_t V = B->V;
    const phg_bv_conf& c = B->conf;
    int rc = upload_mask(B, mask, s);
    if (rc) return rc;
    unsigned fresh = 0;
    for (int k = 0; k < S; ++k)
        if (mask[k]) fresh |= 1u << k;
    rc = jacobian_clear_pairs(B->J, S, fresh, s);               // Jacobian solves from zero
    if (rc) return rc;
    const double* bw = any_bnd(B, mask) ? B->BW.p : nullptr;   // 0 for slots without
    rc = calc1(B, B->U.p, B->W.p, o, B->it_mass, s, bw);        // w = L(u), bnd_w
    if (rc) return rc;
    rc = calc1(B, B->W.p, B->T1.p, o, B->it_mass, s);           // dw = L(w)
    if (rc) return rc;
    // FC = w / tau + mu (1 - theta) dw;  FC = FC - sigma (1 - theta) w
    hipLaunchKernelGGL(bv_axpby, g256(V * S), dim3(BV_BS), 0, s, V, S, 1.0 / c.tau, B->W.p,
                       c.mu * (1.0 - c.theta), B->T1.p, B->FC.p, B->dmask.p, 0);
    hipLaunchKernelGGL(bv_axpby, g256(V * S), dim3(BV_BS), 0, s, V, S, 1.0, B->FC.p,
                       -c.sigma * (1.0 - c.theta), B->W.p, B->FC.p, B->dmask.p, 0);
    hipLaunchKernelGGL(bv_set_cols, g256(V * S), dim3(BV_BS), 0, s, V, S, 0.0, B->Un.p,
                       B->dmask.p);
    hipLaunchKernelGGL(bv_set_cols, g256(V * S), dim3(BV_BS), 0, s, V, S, 0.0, B->Wn.p,
                       B->dmask.p);
    PHG_HIP(hipGetLastError());
    std::vector<double> nr;
    rc = norms(B, B->U.p, nullptr, nr, s);                       // nr0 = ||u||
    if (rc) return rc;
    // FC += rp(x, y, t, mu, sigma) for the slots loaded with a right part
    // (barvortex_impl.h:233-244)
    std::vector<int> forced(S, 0);
    bool any_forced = false;
    for (int k = 0; k < S; ++k) {
        forced[k] = mask[k] && B->has_rp[k];
        any_forced = any_forced || forced[k];
    }
    if (any_forced) {
        rc = upload_mask(B, forced, s);
        if (rc) return rc;
        hipLaunchKernelGGL(bv_axpby, g256(V * S), dim3(BV_BS), 0, s, V, S, 1.0, B->FC.p, 1.0,
                           B->RPF.p, B->FC.p, B->dmask.p, 0);
        PHG_HIP(hipGetLastError());
    }
    for (int k = 0; k < S; ++k)
        if (mask[k]) {
            B->nr0[k] = nr[k];
            B->iters[k] = 0;
            B->rel[k] = 1.0;
            B->rel_prev[k] = 1.0;
            B->rel_prev2[k] = 1.0;
            B->rel_min[k] = 1.0;
            B->stalled[k] = 0;
            B->diverged[k] = 0;
        }
    return PHG_OK;
}

// One pass of the fixed point (barvortex_impl.h:250-305) for the slots in mask.
int iterate_slots(phg_barvortex_s* B, const std::vector<int>& mask, const phg_bv_opts& o,
                  std::vector<double>& nr, hipStream_t s)
{
    const int S = B->S, n = B->n;
    const size_t V = B->V;
    const phg_bv_conf& c = B->conf;
    int rc = upload_mask(B, mask, s);
    if (rc) return rc;
    // tmp1 = k1 theta w_n + k1 (1 - theta) w + k2 lh;  tmp2 = theta u_n + (1 - theta) u
    hipLaunchKernelGGL(bv_axpby, g256(V * S), dim3(BV_BS), 0, s, V, S, c.k1 * c.theta, B->Wn.p,
                       c.k1 * (1.0 - c.theta), B->W.p, B->T1.p, B->dmask.p, 1);
    hipLaunchKernelGGL(bv_add_shared, g256(V * S), dim3(BV_BS), 0, s, B->V, S, B->T1.p, c.k2,
                       B->lh.p, B->T1.p, B->dmask.p);
    hipLaunchKernelGGL(bv_axpby, g256(V * S), dim3(BV_BS), 0, s, V, S, c.theta, B->Un.p,
                       1.0 - c.theta, B->U.p, B->T2.p, B->dmask.p, 1);
    PHG_HIP(hipGetLastError());
    // jac = J(tmp2, tmp1): 8 mass solves per slot, one batched solve.  Each
    // slot's solves start from its solutions of the previous pass (cleared
    // when the slot starts a step); idle slots are cleared so that their
    // zero right-hand sides need no step.
    {
        unsigned idle = 0;
        for (int k = 0; k < S; ++k)
            if (!mask[k]) idle |= 1u << k;
        rc = jacobian_clear_pairs(B->J, S, idle, s);
        if (rc) return rc;
        phg_gmres_opts jo = solve_opts(o.rtol_mass, o.precision, PHG_CGS1, 1, o.max_solver_iters);
        jo.restart = MASS_RESTART;
        jo.diag_scale = 1;
        std::vector<phg_gmres_info> inf((size_t)8 * S);
        rc = jacobian_calc2_strided(B->J, S, B->T2.p, B->T1.p, S, 1, B->JAC.p, S, 1, 1.0, &jo,
                                    inf.data(), s);
        if (rc < 0) return rc;
        count_iters(inf, S, B->it_jac, 8);
    }
    // F = FC_inner - jac;  rp = idt F
    hipLaunchKernelGGL(bv_rhs, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       B->FC.p, B->JAC.p, B->F.p, B->dmask.p);
    spmm(B, B->idt, B->F.p, B->RP.p, 0, s);
    // Boundary values (barvortex_impl.h:287-293): rp += bnd_ bnd_w.  The
    // reference adds the block that a Dirichlet elimination would subtract;
    // this follows it.  Slots without boundary values hold zeros.
    const bool bnd = any_bnd(B, mask);
    if (bnd) spmm(B, B->bndA, B->BW.p, B->RP.p, 1, s);
    zero_cols(B, n, B->RP.p, s);
    // w_n = p2u(A^-1 rp, bnd_w), starting from the current w_n
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       B->Wn.p, B->X.p, B->dmask.p);
    PHG_HIP(hipGetLastError());
    rc = batched(B, B->A, B->RP.p, B->X.p, o.rtol_helmholtz, o.precision, PHG_CGS2, 1, nullptr,
                 o.max_solver_iters, B->it_helm, s);
    if (rc) return rc;
    if (bnd)
        hipLaunchKernelGGL(bv_p2u_bnd, g256(V * S), dim3(BV_BS), 0, s, B->V, S, B->mesh->d_isb.p,
                           B->mesh->d_p2io.p, B->X.p, B->BW.p, B->Wn.p, B->dmask.p);
    else
        hipLaunchKernelGGL(bv_p2u, g256(V * S), dim3(BV_BS), 0, s, B->V, S, B->mesh->d_isb.p,
                           B->mesh->d_p2io.p, B->X.p, B->Wn.p, B->dmask.p);
    // u_n1 = p2u(laplace^-1 (idt u2p(w_n) + bnd2 bnd_u), bnd_u)
    // (SphereLaplace::solve, slaplace_impl.h:84-130), starting from the current u_n
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       B->Wn.p, B->Y.p, B->dmask.p);
    spmm(B, B->idt, B->Y.p, B->Q.p, 0, s);
    if (bnd) spmm(B, B->bnd2, B->BU.p, B->Q.p, 1, s);
    zero_cols(B, n, B->Q.p, s);
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       B->Un.p, B->Z.p, B->dmask.p);
    PHG_HIP(hipGetLastError());
    phg_matrix_t lap = nullptr;
    phg_mg_matrix(B->mg, &lap);
    rc = batched(B, lap, B->Q.p, B->Z.p, o.rtol_poisson, PHG_F64, PHG_CGS2, 1, B->mg,
                 o.max_solver_iters, B->it_pois, s);
    if (rc) return rc;
    if (bnd)
        hipLaunchKernelGGL(bv_p2u_bnd, g256(V * S), dim3(BV_BS), 0, s, B->V, S, B->mesh->d_isb.p,
                           B->mesh->d_p2io.p, B->Z.p, B->BU.p, B->Un1.p, B->dmask.p);
    else
        hipLaunchKernelGGL(bv_p2u, g256(V * S), dim3(BV_BS), 0, s, B->V, S, B->mesh->d_isb.p,
                           B->mesh->d_p2io.p, B->Z.p, B->Un1.p, B->dmask.p);
    PHG_HIP(hipGetLastError());
    // nr = ||u_n1 - u_n||, then u_n = u_n1
    rc = norms(B, B->Un1.p, B->Un.p, nr, s);
    if (rc) return rc;
    hipLaunchKernelGGL(bv_copy_cols, g256(V * S), dim3(BV_BS), 0, s, V, S, B->Un1.p, B->Un.p,
                       B->dmask.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// ---- linearised and adjoint steps (barvortex_impl.h:326-571), bnd = 0 ----

// Checks that no step is in flight, selects slots 0..nm-1 and clears every
// work vector, so that the other slots' columns stay zero (zero right-hand
// sides, which the batched solver skips).
int lin_begin(phg_barvortex_s* B, int nm, hipStream_t s)
{
    const int S = B->S;
    for (int k = 0; k < S; ++k)
        if (B->status[k] != SLOT_FREE) {
            set_error("phg_barvortex linearised step: slots in use by a step");
            return PHG_ERR_ARG;
        }
    std::vector<int> mask(S, 0);
    for (int k = 0; k < nm; ++k) mask[k] = 1;
    int rc = upload_mask(B, mask, s);
    if (rc) return rc;
    const size_t V = B->V, n = B->n;
    for (auto* b : {&B->U, &B->W, &B->FC, &B->Un, &B->Wn, &B->Un1, &B->T1, &B->T2})
        PHG_HIP(hipMemsetAsync(b->p, 0, V * S * sizeof(double), s));
    for (auto* b : {&B->JAC, &B->F, &B->RP, &B->X, &B->Y, &B->Z, &B->Q})
        PHG_HIP(hipMemsetAsync(b->p, 0, n * S * sizeof(double), s));
    for (auto* v : {&B->it_mass, &B->it_jac, &B->it_helm, &B->it_pois}) v->assign(S, 0);
    B->solve_fail.assign(S, 0);
    return PHG_OK;
}

void cols_in(phg_barvortex_s* B, int nm, const double* d, double* M, hipStream_t s)
{
    for (int k = 0; k < nm; ++k)
        hipLaunchKernelGGL(bv_col_in, g256(B->V), dim3(BV_BS), 0, s, B->V, B->S, k,
                           d + (size_t)k * B->V, M);
}

void cols_out(phg_barvortex_s* B, int nm, const double* M, double* d, hipStream_t s)
{
    for (int k = 0; k < nm; ++k)
        hipLaunchKernelGGL(bv_col_out, g256(B->V), dim3(BV_BS), 0, s, B->V, B->S, k, M,
                           d + (size_t)k * B->V);
}

// out = SphereJacobian::calc1 (sign 1) or calc1t (sign -1) of (a, b) with
// bnd = 0 (src/sjacobian.cpp:456-468); the mass solves start from zero.
int jac_full(phg_barvortex_s* B, const double* a, const double* b, double sign, double* out,
             const phg_bv_opts& o, hipStream_t s)
{
    const int S = B->S;
    phg_gmres_opts jo = solve_opts(o.rtol_mass, o.precision, PHG_CGS1, 0, o.max_solver_iters);
    jo.restart = MASS_RESTART;
    jo.diag_scale = 1;
    std::vector<phg_gmres_info> inf((size_t)8 * S);
    int rc = jacobian_calc2_strided(B->J, S, a, b, S, 1, B->JAC.p, S, 1, sign, &jo, inf.data(), s);
    if (rc < 0) return rc;
    count_iters(inf, S, B->it_jac, 8);
    for (int k = 0; k < S; ++k)
        for (int j = 0; j < 8; ++j)
            if (!inf[(size_t)k * 8 + j].converged) B->solve_fail[k] = 1;
    hipLaunchKernelGGL(bv_p2u, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                       B->mesh->d_isb.p, B->mesh->d_p2io.p, B->JAC.p, out, B->dmask.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// out = phelm::solve(M, rp) with rp = idt u2p(in) (+ bndM bw) and out =
// p2u(M^-1 rp, bw) (mesh.h:719-746; barvortex_impl.h:386-397, 455-465): the
// Helmholtz-type A_ / Ab_ solves of the linearised steps, from zero; bw
// (outer x S boundary values of the Laplacian) null for zero.
int solve_full(phg_barvortex_s* B, phg_matrix_s* M, const double* in, double* out,
               const phg_bv_opts& o, hipStream_t s, const phg_matrix_s* bndM = nullptr,
               const double* bw = nullptr)
{
    const int S = B->S, n = B->n;
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       in, B->Y.p, B->dmask.p);
    spmm(B, B->idt, B->Y.p, B->RP.p, 0, s);
    if (bw) spmm(B, bndM, bw, B->RP.p, 1, s);   // added, as the reference adds it
    zero_cols(B, n, B->RP.p, s);
    PHG_HIP(hipGetLastError());
    int rc = batched(B, M, B->RP.p, B->X.p, o.rtol_helmholtz, o.precision, PHG_CGS2, 0, nullptr,
                     o.max_solver_iters, B->it_helm, s);
    if (rc) return rc;
    if (bw)
        hipLaunchKernelGGL(bv_p2u_bnd, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->X.p, bw, out, B->dmask.p);
    else
        hipLaunchKernelGGL(bv_p2u, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->X.p, out, B->dmask.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// out = SphereLaplace::solve(in, bu) (src/impl/slaplace_impl.h:84-130):
// p2u(laplace^-1 (idt u2p(in) + bnd2 bu), bu), multigrid-preconditioned, from
// zero; bu (outer x S boundary values of psi) null for zero.
int poisson_full(phg_barvortex_s* B, const double* in, double* out, const phg_bv_opts& o,
                 hipStream_t s, const double* bu = nullptr)
{
    const int S = B->S, n = B->n;
    hipLaunchKernelGGL(bv_u2p, g256((size_t)n * S), dim3(BV_BS), 0, s, n, S, B->mesh->d_inner.p,
                       in, B->Y.p, B->dmask.p);
    spmm(B, B->idt, B->Y.p, B->Q.p, 0, s);
    if (bu) spmm(B, B->bnd2, bu, B->Q.p, 1, s);
    zero_cols(B, n, B->Q.p, s);
    PHG_HIP(hipGetLastError());
    phg_matrix_t lap = nullptr;
    phg_mg_matrix(B->mg, &lap);
    int rc = batched(B, lap, B->Q.p, B->Z.p, o.rtol_poisson, PHG_F64, PHG_CGS2, 0, B->mg,
                     o.max_solver_iters, B->it_pois, s);
    if (rc) return rc;
    if (bu)
        hipLaunchKernelGGL(bv_p2u_bnd, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->Z.p, bu, out, B->dmask.p);
    else
        hipLaunchKernelGGL(bv_p2u, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, B->Z.p, out, B->dmask.p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// The boundary matrices and outer x S arrays of boundary values, on first
// use: bnd_ (integrate_cb) and bnd2_ (laplace_bnd2_cb = -(stiffness
// block)), and with backward bndb_ (integrate_backward_cb).
int ensure_bnd(phg_barvortex_s* B, bool backward, hipStream_t s)
{
    const phg_bv_conf& c = B->conf;
    if (!B->BU.p) {
        int rc = assemble_boundary_device(B->mesh, PHG_HELMHOLTZ, 1.0 / c.tau + c.sigma * c.theta,
                                          -c.theta * c.mu, s, &B->bndA);
        if (rc) return rc;
        rc = assemble_boundary_device(B->mesh, PHG_HELMHOLTZ, 0.0, -1.0, s, &B->bnd2);
        if (rc) return rc;
        for (DevBuf<double>* b : {&B->BU, &B->BW}) {
            PHG_HIP(b->alloc((size_t)B->os * B->S));
            PHG_HIP(hipMemsetAsync(b->p, 0, (size_t)B->os * B->S * sizeof(double), s));
        }
    }
    if (backward && !B->bndAb) {
        int rc = assemble_boundary_device(B->mesh, PHG_HELMHOLTZ,
                                          1.0 / c.tau - c.sigma * (1.0 - c.theta),
                                          (1.0 - c.theta) * c.mu, s, &B->bndAb);
        if (rc) return rc;
    }
    return PHG_OK;
}

// Boundary values of members 0..nm-1 for a linearised step (outer_size each;
// a null array is zero) into the BU / BW columns of slots 0..nm-1; false
// when neither is given (the steps then take null, as for zero values).
int lin_bnd_in(phg_barvortex_s* B, int nm, const double* d_bnd_u, const double* d_bnd_w,
               bool backward, bool* have, hipStream_t s)
{
    *have = false;
    if (!d_bnd_u && !d_bnd_w) return PHG_OK;
    if (B->os == 0) {
        set_error("phg_barvortex linearised step: the mesh has no boundary points");
        return PHG_ERR_ARG;
    }
    int rc = ensure_bnd(B, backward, s);
    if (rc) return rc;
    for (int k = 0; k < nm; ++k) {
        const size_t ob = (size_t)k * B->os;
        hipLaunchKernelGGL(bv_col_set, g256(B->os), dim3(BV_BS), 0, s, B->os, B->S, k,
                           d_bnd_u ? d_bnd_u + ob : nullptr, B->BU.p);
        hipLaunchKernelGGL(bv_col_set, g256(B->os), dim3(BV_BS), 0, s, B->os, B->S, k,
                           d_bnd_w ? d_bnd_w + ob : nullptr, B->BW.p);
        // a later load of the slot clears its columns (phg_barvortex_load)
        B->has_bnd[k] = 1;
    }
    PHG_HIP(hipGetLastError());
    *have = true;
    return PHG_OK;
}

void lin_sum(phg_barvortex_s* B, const double* base, const double* P1, double c1,
             const double* P2, double c2, const double* A, double ka, const double* Bv,
             double kb, const double* C, double kc, double sj, int zb, double* r, hipStream_t s)
{
    hipLaunchKernelGGL(bv_lin_sum, g256((size_t)B->V * B->S), dim3(BV_BS), 0, s, B->V, B->S, base,
                       P1, c1, P2, c2, A, ka, Bv, kb, C, kc, sj, B->mesh->d_isb.p, zb, r,
                       B->dmask.p);
}

// calc_L (which 0) or calc_L_1 (which 1) of members 0..nm-1, with the
// members' boundary values of psi and of its Laplacian (outer_size each,
// null for zero).
int lin_forward(phg_barvortex_s* B, int which, int nm, const double* d_u, const double* d_z,
                const double* d_bnd_u, const double* d_bnd_w, double* d_u1,
                const phg_bv_opts& o, hipStream_t s)
{
    const phg_bv_conf& c = B->conf;
    const int S = B->S;
    if (which == 1 && !B->Ab) {
        // Ab_ = (1/tau - sigma (1-theta)) M + (1-theta) mu L (integrate_backward_cb)
        int rc = assemble_device(B->mesh, PHG_HELMHOLTZ, 1.0 / c.tau - c.sigma * (1.0 - c.theta),
                                 (1.0 - c.theta) * c.mu, s, &B->Ab);
        if (rc) return rc;
    }
    int rc = lin_begin(B, nm, s);
    if (rc) return rc;
    bool lb = false;
    if ((rc = lin_bnd_in(B, nm, d_bnd_u, d_bnd_w, which == 1, &lb, s))) return rc;
    const double* bu = lb ? B->BU.p : nullptr;
    const double* bw = lb ? B->BW.p : nullptr;
    cols_in(B, nm, d_u, B->U.p, s);                                   // u
    cols_in(B, nm, d_z, B->T2.p, s);                                  // z
    PHG_HIP(hipGetLastError());
    if ((rc = calc1(B, B->T2.p, B->Wn.p, o, B->it_mass, s, bw))) return rc;   // z_lapl = L(z)
    if ((rc = calc1(B, B->U.p, B->W.p, o, B->it_mass, s, bw))) return rc;    // pt1 = L(u)
    hipLaunchKernelGGL(bv_bcast, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S, B->lh.p,
                       B->T1.p, B->dmask.p);
    if ((rc = jac_full(B, B->U.p, B->T1.p, 1.0, B->Un.p, o, s))) return rc;    // J(u, lh)
    if ((rc = jac_full(B, B->T2.p, B->W.p, 1.0, B->Un1.p, o, s))) return rc;   // J(z, pt1)
    if ((rc = jac_full(B, B->U.p, B->Wn.p, 1.0, B->FC.p, o, s))) return rc;    // J(u, z_lapl)
    if ((rc = calc1(B, B->W.p, B->T1.p, o, B->it_mass, s))) return rc;   // L(pt1)
    const double* base = nullptr;
    double c1, c2, sj;
    if (which == 0) {
        c1 = 1.0 / c.tau - (1.0 - c.theta) * c.sigma;
        c2 = (1.0 - c.theta) * c.mu;
        sj = -1.0;
    } else {
        c1 = 1.0 / c.tau + c.theta * c.sigma;
        c2 = -c.theta * c.mu;
        sj = 1.0;
        cols_in(B, nm, d_u1, B->T2.p, s);                             // u1 accumulates
        base = B->T2.p;
    }
    lin_sum(B, base, B->W.p, c1, B->T1.p, c2, B->Un.p, c.k2, B->Un1.p, c.k1, B->FC.p, c.k1, sj, 0,
            B->T2.p, s);
    PHG_HIP(hipGetLastError());
    if ((rc = solve_full(B, which == 0 ? B->A : B->Ab, B->T2.p, B->Wn.p, o, s,
                         which == 0 ? B->bndA : B->bndAb, bw)))
        return rc;
    if ((rc = poisson_full(B, B->Wn.p, B->Un.p, o, s, bu))) return rc;
    cols_out(B, nm, B->Un.p, d_u1, s);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

// calc_LT of members 0..nm-1, with boundary values as lin_forward.
int lin_adjoint(phg_barvortex_s* B, int nm, const double* d_v, const double* d_z,
                const double* d_bnd_u, const double* d_bnd_w, double* d_v1,
                const phg_bv_opts& o, hipStream_t s)
{
    const phg_bv_conf& c = B->conf;
    const int S = B->S;
    int rc = lin_begin(B, nm, s);
    if (rc) return rc;
    bool lb = false;
    if ((rc = lin_bnd_in(B, nm, d_bnd_u, d_bnd_w, false, &lb, s))) return rc;
    const double* bu = lb ? B->BU.p : nullptr;
    const double* bw = lb ? B->BW.p : nullptr;
    cols_in(B, nm, d_v, B->U.p, s);                                   // v
    cols_in(B, nm, d_z, B->T2.p, s);                                  // z
    PHG_HIP(hipGetLastError());
    if ((rc = calc1(B, B->T2.p, B->Wn.p, o, B->it_mass, s, bw))) return rc;   // lz = L(z)
    if ((rc = poisson_full(B, B->U.p, B->W.p, o, s, bu))) return rc;         // v1 = L^-1 v
    if ((rc = solve_full(B, B->A, B->W.p, B->W.p, o, s, B->bndA, bw))) return rc;  // A^-1
    // L(v1) with bnd_w: the reference's pt1 and its first calc1 of pt2 are
    // the same product (barvortex_impl.h:501-506)
    if ((rc = calc1(B, B->W.p, B->Un.p, o, B->it_mass, s, bw))) return rc;
    if ((rc = calc1(B, B->Un.p, B->Un1.p, o, B->it_mass, s))) return rc; // L(L(v1))
    hipLaunchKernelGGL(bv_bcast, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S, B->lh.p,
                       B->T1.p, B->dmask.p);
    if ((rc = jac_full(B, B->W.p, B->T1.p, -1.0, B->FC.p, o, s))) return rc;   // JT(v1, lh)
    if ((rc = jac_full(B, B->T2.p, B->W.p, -1.0, B->U.p, o, s))) return rc;    // JT(z, v1)
    if ((rc = calc1(B, B->U.p, B->T1.p, o, B->it_mass, s))) return rc;   // p_lapl
    if ((rc = jac_full(B, B->W.p, B->Wn.p, -1.0, B->U.p, o, s))) return rc;    // JT(v1, lz)
    lin_sum(B, nullptr, B->Un.p, 1.0 / c.tau - (1.0 - c.theta) * c.sigma, B->Un1.p,
            (1.0 - c.theta) * c.mu, B->U.p, c.k1, B->T1.p, c.k1, B->FC.p, c.k2, -1.0, 1, B->T2.p,
            s);
    if (bu)   // set_bnd(v1, bnd_u): lin_sum left zeros there
        hipLaunchKernelGGL(bv_set_bnd_cols, g256((size_t)B->V * S), dim3(BV_BS), 0, s, B->V, S,
                           B->mesh->d_isb.p, B->mesh->d_p2io.p, bu, B->T2.p, B->dmask.p);
    cols_out(B, nm, B->T2.p, d_v1, s);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int lin_finish(phg_barvortex_s* B, int nm, phg_bv_info* infos, hipStream_t s)
{
    PHG_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int k = 0; k < nm; ++k) {
        const bool ok = !B->solve_fail[k];
        all = all && ok;
        if (!infos) continue;
        phg_bv_info& f = infos[k];
        std::memset(&f, 0, sizeof(f));
        f.iterations = 1;
        f.converged = ok;
        f.it_mass = B->it_mass[k];
        f.it_jacobian = B->it_jac[k];
        f.it_helmholtz = B->it_helm[k];
        f.it_poisson = B->it_pois[k];
        f.solver_iterations = f.it_mass + f.it_jacobian + f.it_helmholtz + f.it_poisson;
    }
    return all ? PHG_OK : PHG_ERR_NOCONV;
}

}  // namespace

extern "C" {

void phg_bv_default_opts(phg_bv_opts* o)
{
    o->max_iters = 1000;
    o->tol = 1e-14;
    o->rtol_mass = 1e-13;
    o->rtol_helmholtz = 1e-12;
    o->rtol_poisson = 1e-10;
    o->precision = PHG_F64;
    o->max_solver_iters = 1000;
    o->stall_floor = 1e-8;
}

size_t phg_barvortex_estimate_bytes(int n_points, int slots)
{
    // matrices: ~7 entries per row, 12 B each (value + column), for idt, A,
    // laplace + V-cycle levels (4/3), 5 Jacobian matrices, the full mass
    // and the ELL copies; Krylov bases: Jacobian (33 x 8S), Helmholtz (31 x S),
    // Poisson (31 x S + 30 x S preconditioned), L() mass (31 x S); slot
    // vectors: 8 full + 7 inner + V-cycle levels.
    const double n = (double)n_points, S = (double)(slots < 1 ? 1 : slots);
    const int kp8 = 8 * slots <= 8 ? 8 : 8 * slots <= 16 ? 16 : 8 * slots <= 32 ? 32 : 64;
    double kp = 1;
    while (kp < S) kp *= 2;
    const double mats = n * 7 * 12 * (2 + 4.0 / 3 + 5 + 1) * 2;
    const double bases = n * 8 * (33.0 * kp8 + 31 * kp + 61 * kp + 31 * kp + 8 * kp8);
    const double vecs = n * 8 * S * (8 + 7 + 4 * 4.0 / 3);
    return (size_t)(mats + bases + vecs);
}

int phg_barvortex_create(phg_mesh_t m, const phg_bv_conf* conf, const double* d_lh, int slots,
                         phg_stream_t stream, phg_barvortex_t* out)
{
    if (!m || !conf || !d_lh || !out || slots < 1 || slots > BV_MAXS ||
        (slots & (slots - 1)) != 0) {
        set_error("phg_barvortex_create: bad argument (slots must be 1, 2, 4 or 8)");
        return PHG_ERR_ARG;
    }
    if (!(conf->tau > 0.0)) {
        set_error("phg_barvortex_create: tau must be positive");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<phg_barvortex_s> B(new phg_barvortex_s());
    B->mesh = m;
    B->conf = *conf;
    B->S = slots;
    B->V = m->h.n_points;
    B->n = (int)m->h.inner.size();
    B->os = (int)m->h.outer.size();
    const phg_bv_conf& c = *conf;
    int rc = assemble_device(m, PHG_MASS, 1.0, 0.0, s, &B->idt);
    if (rc) return rc;
    rc = assemble_boundary_device(m, PHG_LAPLACE, 1.0, 0.0, s, &B->bnd3);
    if (rc) return rc;
    // A_ = (1/tau + sigma theta) M - theta mu L  (bv_private::integrate_cb)
    rc = assemble_device(m, PHG_HELMHOLTZ, 1.0 / c.tau + c.sigma * c.theta, -c.theta * c.mu, s,
                         &B->A);
    if (rc) return rc;
    rc = assemble_device(m, PHG_MASS, 1.0, 0.0, s, &B->norm, true);
    if (rc) return rc;
    rc = phg_jacobian_create(m, stream, &B->J);
    if (rc) return rc;
    rc = phg_mg_create(m, PHG_LAPLACE, 1.0, 0.0, stream, &B->mg);
    if (rc) return rc;
    const size_t V = B->V, n = B->n, S = slots;
    PHG_HIP(B->lh.alloc(V));
    PHG_HIP(hipMemcpyAsync(B->lh.p, d_lh, V * sizeof(double), hipMemcpyDeviceToDevice, s));
    DevBuf<double>* full[] = {&B->U, &B->W, &B->FC, &B->Un, &B->Wn, &B->Un1, &B->T1, &B->T2};
    for (auto* b : full) {
        PHG_HIP(b->alloc(V * S));
        PHG_HIP(hipMemsetAsync(b->p, 0, V * S * sizeof(double), s));
    }
    DevBuf<double>* inner[] = {&B->JAC, &B->F, &B->RP, &B->X, &B->Y, &B->Z, &B->Q};
    for (auto* b : inner) {
        PHG_HIP(b->alloc(n * S));
        PHG_HIP(hipMemsetAsync(b->p, 0, n * S * sizeof(double), s));
    }
    PHG_HIP(B->Ub.alloc(std::max<size_t>(B->os, 1) * S));
    B->G = (int)((V + (BV_BS / S) - 1) / (BV_BS / S));
    PHG_HIP(B->part.alloc((size_t)B->G * S));
    PHG_HIP(B->nrm.alloc(S));
    PHG_HIP(B->dmask.alloc(S));
    B->status.assign(S, SLOT_FREE);
    B->iters.assign(S, 0);
    for (auto* v : {&B->it_mass, &B->it_jac, &B->it_helm, &B->it_pois}) v->assign(S, 0);
    B->nr0.assign(S, 0.0);
    B->rel.assign(S, 0.0);
    B->rel_prev.assign(S, 1.0);
    B->rel_prev2.assign(S, 1.0);
    B->rel_min.assign(S, 1.0);
    B->stalled.assign(S, 0);
    B->diverged.assign(S, 0);
    B->has_rp.assign(S, 0);
    B->has_bnd.assign(S, 0);
    B->solve_fail.assign(S, 0);
    PHG_HIP(hipStreamSynchronize(s));
    *out = B.release();
    return PHG_OK;
}

void phg_barvortex_destroy(phg_barvortex_t B) { delete B; }

int phg_barvortex_slots(phg_barvortex_t B) { return B ? B->S : 0; }

int phg_barvortex_load(phg_barvortex_t B, int slot, const double* d_u, phg_stream_t stream)
{
    if (!B || !d_u || slot < 0 || slot >= B->S || B->status[slot] != SLOT_FREE) {
        set_error("phg_barvortex_load: bad argument or slot not free");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(bv_col_in, g256(B->V), dim3(BV_BS), 0, s, B->V, B->S, slot, d_u, B->U.p);
    PHG_HIP(hipGetLastError());
    B->status[slot] = SLOT_NEW;
    B->has_rp[slot] = 0;
    if (B->has_bnd[slot]) {
        for (DevBuf<double>* b : {&B->BU, &B->BW})
            hipLaunchKernelGGL(bv_col_set, g256(B->os), dim3(BV_BS), 0, s, B->os, B->S, slot,
                               nullptr, b->p);
        PHG_HIP(hipGetLastError());
        B->has_bnd[slot] = 0;
    }
    for (auto* v : {&B->it_mass, &B->it_jac, &B->it_helm, &B->it_pois}) (*v)[slot] = 0;
    return PHG_OK;
}

int phg_barvortex_load_rp(phg_barvortex_t B, int slot, const double* d_u, const double* d_rp,
                          phg_stream_t stream)
{
    int rc = phg_barvortex_load(B, slot, d_u, stream);
    if (rc || !d_rp) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (!B->RPF.p) {
        PHG_HIP(B->RPF.alloc((size_t)B->V * B->S));
        PHG_HIP(hipMemsetAsync(B->RPF.p, 0, (size_t)B->V * B->S * sizeof(double), s));
    }
    hipLaunchKernelGGL(bv_col_in, g256(B->V), dim3(BV_BS), 0, s, B->V, B->S, slot, d_rp,
                       B->RPF.p);
    PHG_HIP(hipGetLastError());
    B->has_rp[slot] = 1;
    return PHG_OK;
}

int phg_barvortex_load_bnd(phg_barvortex_t B, int slot, const double* d_bnd_u,
                           const double* d_bnd_w, phg_stream_t stream)
{
    if (!B || slot < 0 || slot >= B->S || B->status[slot] != SLOT_NEW) {
        set_error("phg_barvortex_load_bnd: bad argument or slot not loaded");
        return PHG_ERR_ARG;
    }
    if (!d_bnd_u && !d_bnd_w) return PHG_OK;
    if (B->os == 0) {
        set_error("phg_barvortex_load_bnd: the mesh has no boundary points");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc = ensure_bnd(B, false, s);
    if (rc) return rc;
    hipLaunchKernelGGL(bv_col_set, g256(B->os), dim3(BV_BS), 0, s, B->os, B->S, slot, d_bnd_u,
                       B->BU.p);
    hipLaunchKernelGGL(bv_col_set, g256(B->os), dim3(BV_BS), 0, s, B->os, B->S, slot, d_bnd_w,
                       B->BW.p);
    PHG_HIP(hipGetLastError());
    B->has_bnd[slot] = 1;
    return PHG_OK;
}

int phg_barvortex_iterate(phg_barvortex_t B, const phg_bv_opts* opts, int* state,
                          phg_bv_info* infos, phg_stream_t stream)
{
    if (!B) {
        set_error("phg_barvortex_iterate: null argument");
        return PHG_ERR_ARG;
    }
    phg_bv_opts o;
    if (opts) o = *opts;
    else phg_bv_default_opts(&o);
    hipStream_t s = (hipStream_t)stream;
    const int S = B->S;
    std::vector<int> fresh(S, 0), run(S, 0);
    bool any_new = false, any_run = false;
    for (int k = 0; k < S; ++k) {
        fresh[k] = B->status[k] == SLOT_NEW;
        any_new = any_new || fresh[k];
    }
    int rc;
    if (any_new) {
        rc = init_slots(B, fresh, o, s);
        if (rc) return rc;
        for (int k = 0; k < S; ++k)
            if (fresh[k]) B->status[k] = SLOT_RUN;
    }
    for (int k = 0; k < S; ++k) {
        run[k] = B->status[k] == SLOT_RUN;
        any_run = any_run || run[k];
    }
    if (any_run) {
        std::vector<double> nr;
        rc = iterate_slots(B, run, o, nr, s);
        if (rc) return rc;
        for (int k = 0; k < S; ++k) {
            if (!run[k]) continue;
            B->iters[k] += 1;
            B->rel_prev2[k] = B->rel_prev[k];
            B->rel_prev[k] = B->rel[k];
            B->rel[k] = B->nr0[k] > 0.0 ? nr[k] / B->nr0[k] : nr[k];
            B->stalled[k] = B->rel[k] >= o.tol && B->rel[k] < o.stall_floor &&
                            B->iters[k] >= 3 && B->rel[k] >= 0.5 * B->rel_prev2[k];
            // not contracting: the change grew past the state, or 1000x above
            // its smallest value (a mode the fixed point amplifies)
            B->diverged[k] = !std::isfinite(B->rel[k]) ||
                             (B->iters[k] >= 3 && B->rel[k] > 1.0) ||
                             (B->iters[k] >= 4 && B->rel[k] > 1e3 * B->rel_min[k] &&
                              B->rel[k] > o.stall_floor);
            B->rel_min[k] = std::min(B->rel_min[k], B->rel[k]);
            if (B->rel[k] < o.tol || B->stalled[k] || B->diverged[k] ||
                B->iters[k] >= o.max_iters)
                B->status[k] = SLOT_DONE;
        }
    }
    for (int k = 0; k < S; ++k) {
        if (state) state[k] = B->status[k] == SLOT_FREE ? 0 : B->status[k] == SLOT_DONE ? 2 : 1;
        if (infos) {
            infos[k].iterations = B->iters[k];
            infos[k].rel_change = B->rel[k];
            infos[k].converged =
                B->status[k] == SLOT_DONE && !B->diverged[k] && (B->rel[k] < o.tol || B->stalled[k]);
            infos[k].stalled = B->stalled[k];
            infos[k].diverged = B->diverged[k];
            infos[k].it_mass = B->it_mass[k];
            infos[k].it_jacobian = B->it_jac[k];
            infos[k].it_helmholtz = B->it_helm[k];
            infos[k].it_poisson = B->it_pois[k];
            infos[k].solver_iterations = B->it_mass[k] + B->it_jac[k] + B->it_helm[k] + B->it_pois[k];
        }
    }
    return PHG_OK;
}

int phg_barvortex_store(phg_barvortex_t B, int slot, double* d_u1, phg_stream_t stream)
{
    if (!B || !d_u1 || slot < 0 || slot >= B->S || B->status[slot] != SLOT_DONE) {
        set_error("phg_barvortex_store: bad argument or slot not done");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    // a diverged step leaves the member's state as it was
    hipLaunchKernelGGL(bv_col_out, g256(B->V), dim3(BV_BS), 0, s, B->V, B->S, slot,
                       B->diverged[slot] ? B->U.p : B->Un.p, d_u1);
    PHG_HIP(hipGetLastError());
    B->status[slot] = SLOT_FREE;
    return PHG_OK;
}

int phg_barvortex_calc(phg_barvortex_t B, int nmembers, const double* d_u, double* d_u1,
                       const phg_bv_opts* opts, phg_bv_info* infos, phg_stream_t stream)
{
    return phg_barvortex_calc_rp(B, nmembers, d_u, nullptr, d_u1, opts, infos, stream);
}

int phg_barvortex_calc_rp(phg_barvortex_t B, int nmembers, const double* d_u, const double* d_rp,
                          double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                          phg_stream_t stream)
{
    return phg_barvortex_calc_bnd(B, nmembers, d_u, d_rp, nullptr, nullptr, d_u1, opts, infos,
                                  stream);
}

int phg_barvortex_calc_bnd(phg_barvortex_t B, int nmembers, const double* d_u,
                           const double* d_rp, const double* d_bnd_u, const double* d_bnd_w,
                           double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                           phg_stream_t stream)
{
    if (!B || !d_u || !d_u1 || nmembers < 1 || nmembers > B->S) {
        set_error("phg_barvortex_calc: bad argument (1 <= nmembers <= slots)");
        return PHG_ERR_ARG;
    }
    for (int k = 0; k < B->S; ++k)
        if (B->status[k] != SLOT_FREE) {
            set_error("phg_barvortex_calc: slots in use");
            return PHG_ERR_ARG;
        }
    int rc;
    for (int k = 0; k < nmembers; ++k) {
        rc = phg_barvortex_load_rp(B, k, d_u + (size_t)k * B->V,
                                   d_rp ? d_rp + (size_t)k * B->V : nullptr, stream);
        if (rc) return rc;
        const size_t ob = (size_t)k * B->os;
        rc = phg_barvortex_load_bnd(B, k, d_bnd_u ? d_bnd_u + ob : nullptr,
                                    d_bnd_w ? d_bnd_w + ob : nullptr, stream);
        if (rc) return rc;
    }
    std::vector<int> state(B->S);
    std::vector<phg_bv_info> inf(B->S);
    for (;;) {
        rc = phg_barvortex_iterate(B, opts, state.data(), inf.data(), stream);
        if (rc) return rc;
        bool busy = false;
        for (int k = 0; k < nmembers; ++k) busy = busy || state[k] == 1;
        if (!busy) break;
    }
    bool all = true;
    for (int k = 0; k < nmembers; ++k) {
        rc = phg_barvortex_store(B, k, d_u1 + (size_t)k * B->V, stream);
        if (rc) return rc;
        if (infos) infos[k] = inf[k];
        all = all && inf[k].converged;
    }
    PHG_HIP(hipStreamSynchronize((hipStream_t)stream));
    return all ? PHG_OK : PHG_ERR_NOCONV;
}

// The linearised steps with boundary values (null arrays: zero); the plain
// entry points are these with none.
static int lin_entry(const char* name, int which, phg_barvortex_t B, int nmembers,
                     const double* d_u, const double* d_z, const double* d_bnd_u,
                     const double* d_bnd_w, double* d_u1, const phg_bv_opts* opts,
                     phg_bv_info* infos, phg_stream_t stream)
{
    if (!B || !d_u || !d_z || !d_u1 || nmembers < 1 || nmembers > B->S) {
        set_error(std::string(name) + ": bad argument (1 <= nmembers <= slots)");
        return PHG_ERR_ARG;
    }
    phg_bv_opts o;
    if (opts) o = *opts;
    else phg_bv_default_opts(&o);
    hipStream_t s = (hipStream_t)stream;
    const int rc = which == 2 ? lin_adjoint(B, nmembers, d_u, d_z, d_bnd_u, d_bnd_w, d_u1, o, s)
                              : lin_forward(B, which, nmembers, d_u, d_z, d_bnd_u, d_bnd_w, d_u1,
                                            o, s);
    return rc ? rc : lin_finish(B, nmembers, infos, s);
}

int phg_barvortex_calc_L(phg_barvortex_t B, int nmembers, const double* d_u, const double* d_z,
                         double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                         phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_L", 0, B, nmembers, d_u, d_z, nullptr, nullptr, d_u1,
                     opts, infos, stream);
}

int phg_barvortex_calc_L_1(phg_barvortex_t B, int nmembers, const double* d_u, const double* d_z,
                           double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                           phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_L_1", 1, B, nmembers, d_u, d_z, nullptr, nullptr, d_u1,
                     opts, infos, stream);
}

int phg_barvortex_calc_LT(phg_barvortex_t B, int nmembers, const double* d_v, const double* d_z,
                          double* d_v1, const phg_bv_opts* opts, phg_bv_info* infos,
                          phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_LT", 2, B, nmembers, d_v, d_z, nullptr, nullptr, d_v1,
                     opts, infos, stream);
}

int phg_barvortex_calc_L_bnd(phg_barvortex_t B, int nmembers, const double* d_u,
                             const double* d_z, const double* d_bnd_u, const double* d_bnd_w,
                             double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                             phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_L_bnd", 0, B, nmembers, d_u, d_z, d_bnd_u, d_bnd_w,
                     d_u1, opts, infos, stream);
}

int phg_barvortex_calc_L_1_bnd(phg_barvortex_t B, int nmembers, const double* d_u,
                               const double* d_z, const double* d_bnd_u, const double* d_bnd_w,
                               double* d_u1, const phg_bv_opts* opts, phg_bv_info* infos,
                               phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_L_1_bnd", 1, B, nmembers, d_u, d_z, d_bnd_u, d_bnd_w,
                     d_u1, opts, infos, stream);
}

int phg_barvortex_calc_LT_bnd(phg_barvortex_t B, int nmembers, const double* d_v,
                              const double* d_z, const double* d_bnd_u, const double* d_bnd_w,
                              double* d_v1, const phg_bv_opts* opts, phg_bv_info* infos,
                              phg_stream_t stream)
{
    return lin_entry("phg_barvortex_calc_LT_bnd", 2, B, nmembers, d_v, d_z, d_bnd_u, d_bnd_w,
                     d_v1, opts, infos, stream);
}

}  // extern "C"
