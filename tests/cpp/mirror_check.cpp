// Drives the C++ host classes of include/phelm_gpu.hpp on one sphere mesh
// and writes their outputs as raw FP64 files for the GPU test to compare
// with the oracle and with the Python host classes:
//
//   mirror_check LEVEL OUTDIR
//
// Files (little-endian doubles): spmv_add.bin / spmv_asm.bin (Solver built
// by add() vs the GPU-assembled one, same x), solve_x.bin (Solver::solve of
// the Helmholtz system), lap_calc2.bin, lap_calc1.bin, lap_solve.bin
// (SphereLaplace), jac_calc2.bin, jac_calc2t.bin (SphereJacobian),
// bv_u1.bin (one BarVortex::calc step), bv_rp_u1.bin (one step of the
// reference-style BarVortex(m, conf) with the rp and coriolis callbacks at
// t = 0.5, zero boundary arrays passed), bv_L.bin / bv_L1.bin / bv_LT.bin
// (L_step, L_1_step of bv_L, LT_step about psi), bv_u1_bnd.bin,
// bv_L_bnd.bin, bv_LT_bnd.bin (calc, calc_L and calc_LT about psi with a pole
// value of psi of 1), bc_u11.bin / bc_u21.bin (one
// Baroclin::calc step of (u, v) with the forcing written to bc_f.bin,
// bc_g.bin and the Coriolis term bc_lh.bin; bc_L1/2.bin = L_step of (u, v)
// about (psi, u), bc_B1/2.bin = L_1_step of that, bc_LT1.bin = bc_L1 after
// LT_step; bc_Lb1/2.bin, bc_Bb1/2.bin = calc_L of (u, v) about (psi, u) and
// calc_L_1 of that with a pole value of 0.5), sc_u1.bin / sc_u2.bin (two
// SphereChafe::solve steps of the
// reference driver's analytic case from sc_u0.bin with the boundary values
// sc_b1.bin, sc_b2.bin), inputs u.bin, v.bin, x.bin, psi.bin and summary.txt
// (scalars).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "phelm_gpu.hpp"

using namespace phelm_gpu;

static void dump(const std::string& dir, const char* name, const std::vector<double>& v)
{
    const std::string p = dir + "/" + name;
    FILE* f = std::fopen(p.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(double), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", p.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static std::vector<double> from_dev(const ArrayDevice<double>& a) { return a.host(); }

// smooth forcing and Coriolis callbacks of the reference's signatures
static double bc_f(double x, double y, double t, double, double, double, double, double)
{
    return 0.1 * std::sin(x) * std::cos(y + t);
}
static double bc_g(double x, double y, double t, double, double, double, double, double)
{
    return 0.05 * std::cos(x) * std::sin(2.0 * y - t);
}
static double bc_cor(double phi, double) { return 2.0 * std::sin(phi); }
// BarVortexConf callbacks
static double bv_rp(double x, double y, double t, double mu, double sigma)
{
    return (1.0 + mu + sigma) * std::cos(x) * std::sin(2.0 * y + t);
}
static double bv_cor(double phi, double lambda)
{
    return 2.0 * std::sin(phi) + 0.1 * std::cos(2.0 * lambda) * std::sin(2.0 * phi);
}

static ArrayDevice<double> to_dev(const std::vector<double>& h)
{
    ArrayDevice<double> a(h.size());
    a.upload(h.data());
    return a;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: mirror_check LEVEL OUTDIR\n");
        return 2;
    }
    const int level = std::atoi(argv[1]);
    const std::string out = argv[2];
    try {
        check(phg_init(0), "phg_init");
        Mesh m = Mesh::sphere(level);
        const int V = m.ps_size(), n = m.inner_size();
        std::vector<double> pr((size_t)V * 3);
        check(phg_mesh_arrays(m.handle(), pr.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr),
              "phg_mesh_arrays");
        std::vector<int> inner((size_t)n);
        check(phg_mesh_arrays(m.handle(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, inner.data(), nullptr, nullptr),
              "phg_mesh_arrays");

        // smooth test fields on every point
        std::vector<double> u((size_t)V), v((size_t)V), psi((size_t)V);
        int pole = 0;
        for (int i = 0; i < V; ++i) {
            const double x = pr[3 * i], y = pr[3 * i + 1], z = pr[3 * i + 2];
            u[i] = std::sin(2.0 * x) + 0.5 * y * z;
            v[i] = std::cos(1.5 * y) * z + 0.25 * x;
            const double ph = std::asin(std::fmax(-1.0, std::fmin(1.0, z)));
            const double la = std::atan2(y, x);
            const double w = 7.848e-6 / 7.292e-5;
            psi[i] = -w * std::sin(ph) + w * std::pow(std::cos(ph), 4) * std::sin(ph) *
                                             std::cos(4.0 * la);
            if (z < pr[3 * pole + 2]) pole = i;
        }
        const double p0 = psi[pole];  // the south pole (Dirichlet point) is 0
        for (int i = 0; i < V; ++i) psi[i] -= p0;
        dump(out, "u.bin", u);
        dump(out, "v.bin", v);
        dump(out, "psi.bin", psi);
        ArrayDevice<double> du = to_dev(u), dv = to_dev(v), dpsi = to_dev(psi);

        // Solver<double>: add() in reverse row order vs the assembled matrix
        const double cm = 1.0, cl = -1e-2;
        Solver<double> Aasm = Solver<double>::assemble(m, PHG_HELMHOLTZ, cm, cl);
        long nnz = 0;
        int nr = 0, nc = 0;
        check(phg_matrix_info(Aasm.handle(), &nr, &nc, &nnz), "phg_matrix_info");
        std::vector<int> rp((size_t)n + 1), ci((size_t)nnz);
        std::vector<double> vals((size_t)nnz);
        check(phg_matrix_csr(Aasm.handle(), rp.data(), ci.data(), vals.data()), "phg_matrix_csr");
        Solver<double> Aadd(n);
        for (int i = n - 1; i >= 0; --i)
            for (int k = rp[i + 1] - 1; k >= rp[i]; --k) Aadd.add(i, ci[k], vals[k]);
        std::vector<double> xh((size_t)n);
        for (int i = 0; i < n; ++i) xh[i] = std::sin(0.001 * i) + 0.1;
        dump(out, "x.bin", xh);
        ArrayDevice<double> dx = to_dev(xh), y1((size_t)n), y2((size_t)n);
        Aadd.mult_vector(y1.data(), dx.data());
        Aasm.mult_vector(y2.data(), dx.data());
        dump(out, "spmv_add.bin", from_dev(y1));
        dump(out, "spmv_asm.bin", from_dev(y2));
        ArrayDevice<double> sol((size_t)n);
        Aadd.solve(sol.data(), y2.data());
        dump(out, "solve_x.bin", from_dev(sol));
        const phg_gmres_info si = Aadd.last_info();

        // SphereLaplace
        SphereLaplace<> L(m);
        ArrayDevice<double> c2((size_t)n), c1((size_t)V), ls((size_t)V);
        L.calc2(c2.data(), du.data());
        L.calc1(c1.data(), du.data(), nullptr);
        L.solve(ls.data(), dv.data(), nullptr);
        dump(out, "lap_calc2.bin", from_dev(c2));
        dump(out, "lap_calc1.bin", from_dev(c1));
        dump(out, "lap_solve.bin", from_dev(ls));
        const phg_gmres_info li = L.last_info();

        // SphereNorm
        SphereNorm<> N(m);
        const double nu = N.norm(du.data()), suv = N.scalar(du.data(), dv.data()),
                     duv = N.dist(du.data(), dv.data());

        // SphereJacobian
        SphereJacobian J(m);
        ArrayDevice<double> j2((size_t)n), j2t((size_t)n);
        J.calc2(j2.data(), du.data(), dv.data());
        J.calc2t(j2t.data(), du.data(), dv.data());
        dump(out, "jac_calc2.bin", from_dev(j2));
        dump(out, "jac_calc2t.bin", from_dev(j2t));

        // BarVortex: one step of the Rossby-Haurwitz state
        std::vector<double> lh((size_t)V);
        for (int i = 0; i < V; ++i) lh[i] = 2.0 * std::fmax(-1.0, std::fmin(1.0, pr[3 * i + 2]));
        ArrayDevice<double> dlh = to_dev(lh), u1((size_t)V);
        BarVortex bv(m, BarVortexConf{600.0 * 7.292e-5, 1.6e-2, 8e-5, 1.0, 1.0, 0.5}, dlh.data());
        bv.calc(u1.data(), dpsi.data(), nullptr, nullptr, 0.0);
        dump(out, "bv_u1.bin", from_dev(u1));
        const phg_bv_info bi = bv.last_info();
        ArrayDevice<double> nzb(1), zb(1);
        {
            const double one = 1.0;
            nzb.upload(&one);
        }
        // a pole value of psi of 1 at the new time (bnd_w null = 0)
        ArrayDevice<double> u1b((size_t)V);
        bv.calc(u1b.data(), dpsi.data(), nzb.data(), nullptr, 0.0);
        dump(out, "bv_u1_bnd.bin", from_dev(u1b));
        // the linearised and adjoint steps with that pole value
        {
            ArrayDevice<double> l((size_t)V), lt((size_t)V);
            bv.calc_L(l.data(), dpsi.data(), dpsi.data(), nzb.data(), nullptr, 0.0);
            bv.calc_LT(lt.data(), dpsi.data(), dpsi.data(), nzb.data(), nullptr, 0.0);
            dump(out, "bv_L_bnd.bin", from_dev(l));
            dump(out, "bv_LT_bnd.bin", from_dev(lt));
        }

        // BarVortex(m, conf) with the reference's callbacks; zero boundary
        // arrays, as examples/test_barvortex.cpp passes them
        BarVortexConf rc{600.0 * 7.292e-5, 1.6e-2, 8e-5, 1.0, 1.0, 0.5};
        rc.rp = bv_rp;
        rc.coriolis = bv_cor;
        BarVortex bvr(m, rc);
        ArrayDevice<double> ur((size_t)V);
        bvr.calc(ur.data(), dpsi.data(), zb.data(), zb.data(), 0.5);
        dump(out, "bv_rp_u1.bin", from_dev(ur));
        const phg_bv_info bri = bvr.last_info();
        ArrayDevice<double> dL((size_t)V), dL1((size_t)V), dLT((size_t)V);
        bvr.L_step(dL.data(), du.data(), dpsi.data());
        bvr.L_1_step(dL1.data(), dL.data(), dpsi.data());
        bvr.LT_step(dLT.data(), dv.data(), dpsi.data());
        dump(out, "bv_L.bin", from_dev(dL));
        dump(out, "bv_L1.bin", from_dev(dL1));
        dump(out, "bv_LT.bin", from_dev(dLT));
        {
            // the callbacks at the zone-0 chart coordinates, as the class
            // evaluates them
            std::vector<double> lhr((size_t)V);
            int VV = 0, TT = 0, ni = 0, no = 0, nz = 0;
            check(phg_mesh_info(m.handle(), &VV, &TT, &ni, &no, &nz), "phg_mesh_info");
            std::vector<double> cc((size_t)VV * nz * 2), rpv((size_t)V);
            check(phg_mesh_arrays(m.handle(), nullptr, cc.data(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, nullptr),
                  "phg_mesh_arrays");
            for (int i = 0; i < V; ++i) {
                const double x = cc[(size_t)i * nz * 2], y = cc[(size_t)i * nz * 2 + 1];
                lhr[i] = bv_cor(x, y);
                rpv[i] = bv_rp(x, y, 0.5, rc.mu, rc.sigma);
            }
            dump(out, "bv_rp_lh.bin", lhr);
            dump(out, "bv_rp_f.bin", rpv);
        }

        // Baroclin: one step of (u, v) with the callbacks above at t = 0.25
        std::vector<double> coords;
        {
            int VV = 0, TT = 0, ni = 0, no = 0, nz = 0;
            check(phg_mesh_info(m.handle(), &VV, &TT, &ni, &no, &nz), "phg_mesh_info");
            coords.resize((size_t)VV * nz * 2);
            check(phg_mesh_arrays(m.handle(), nullptr, coords.data(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, nullptr),
                  "phg_mesh_arrays");
            std::vector<double> fv((size_t)V), gv((size_t)V), lhv((size_t)V);
            for (int i = 0; i < V; ++i) {
                const double x = coords[(size_t)i * nz * 2], y = coords[(size_t)i * nz * 2 + 1];
                fv[i] = bc_f(x, y, 0.25, 0, 0, 0, 0, 0);
                gv[i] = bc_g(x, y, 0.25, 0, 0, 0, 0, 0);
                lhv[i] = bc_cor(x, y);
            }
            dump(out, "bc_f.bin", fv);
            dump(out, "bc_g.bin", gv);
            dump(out, "bc_lh.bin", lhv);
        }
        Baroclin bc(m, bc_f, bc_g, bc_cor, 1e-3, 1.6e-2, 8e-2, 1.6e-2, 8e-2, 1.0);
        ArrayDevice<double> b1((size_t)V), b2((size_t)V);
        bc.calc(b1.data(), b2.data(), du.data(), dv.data(), nullptr, 0.25);
        dump(out, "bc_u11.bin", from_dev(b1));
        dump(out, "bc_u21.bin", from_dev(b2));
        const phg_bc_info bci = bc.last_info();
        // linearised steps about (psi, u): L_step of (u, v), L_1_step of that;
        // LT_step leaves its outputs as they are (empty in the reference)
        ArrayDevice<double> l1((size_t)V), l2((size_t)V), k1((size_t)V), k2((size_t)V);
        bc.L_step(l1.data(), l2.data(), du.data(), dv.data(), dpsi.data(), du.data());
        bc.L_1_step(k1.data(), k2.data(), l1.data(), l2.data(), dpsi.data(), du.data());
        dump(out, "bc_L1.bin", from_dev(l1));
        dump(out, "bc_L2.bin", from_dev(l2));
        dump(out, "bc_B1.bin", from_dev(k1));
        dump(out, "bc_B2.bin", from_dev(k2));
        bc.LT_step(l1.data(), l2.data(), du.data(), dv.data(), dpsi.data(), du.data());
        dump(out, "bc_LT1.bin", from_dev(l1));
        // the linearised steps with a pole value of 0.5
        {
            std::vector<double> hb((size_t)std::max(1, m.outer_size()), 0.5);
            ArrayDevice<double> db(hb.size());
            db.upload(hb.data());
            bc.calc_L(l1.data(), l2.data(), du.data(), dv.data(), dpsi.data(), du.data(),
                      db.data(), 0.0);
            bc.calc_L_1(k1.data(), k2.data(), l1.data(), l2.data(), dpsi.data(), du.data(),
                        db.data(), 0.0);
            dump(out, "bc_Lb1.bin", from_dev(l1));
            dump(out, "bc_Lb2.bin", from_dev(l2));
            dump(out, "bc_Bb1.bin", from_dev(k1));
            dump(out, "bc_Bb2.bin", from_dev(k2));
        }

        // SphereChafe: two steps of the reference driver's analytic case
        // (examples/test_schafe.cpp:61-125: tau 0.01, mu 1, sigma 70, boundary
        // values of x sin(y + t) cos^4 x at the new time)
        phg_gmres_info sci{};
        {
            SphereChafe<double> sc(m, 0.01, 70.0, 1.0);
            sc.A_.opts().rtol = 1e-14;
            int nz = 0, np_ = 0, nt = 0, ni = 0, no = 0;
            check(phg_mesh_info(m.handle(), &np_, &nt, &ni, &no, &nz), "phg_mesh_info");
            std::vector<double> coords((size_t)np_ * nz * 2);
            std::vector<int> outer((size_t)std::max(no, 1));
            check(phg_mesh_arrays(m.handle(), nullptr, coords.data(), nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, outer.data(), nullptr),
                  "phg_mesh_arrays");
            auto ans = [&](int p, double t) {
                const double x = coords[(size_t)p * nz * 2], y = coords[(size_t)p * nz * 2 + 1];
                const double c = std::cos(x);
                return x * std::sin(y + t) * ((((1.0 * c) * c) * c) * c);
            };
            std::vector<double> u0((size_t)V), b1((size_t)no), b2((size_t)no);
            for (int p = 0; p < V; ++p) u0[p] = ans(p, 0.0);
            for (int k = 0; k < no; ++k) {
                b1[k] = ans(outer[k], 0.01);
                b2[k] = ans(outer[k], 0.02);
            }
            ArrayDevice<double> dU = to_dev(u0), dB1 = to_dev(b1), dB2 = to_dev(b2);
            sc.solve(dU.data(), dU.data(), dB1.data(), 0.0);
            dump(out, "sc_u1.bin", from_dev(dU));
            sc.solve(dU.data(), dU.data(), dB2.data(), 0.01);
            dump(out, "sc_u2.bin", from_dev(dU));
            sci = sc.A_.last_info();
            dump(out, "sc_u0.bin", u0);
            dump(out, "sc_b1.bin", b1);
            dump(out, "sc_b2.bin", b2);
        }

        const std::string sp = out + "/summary.txt";
        FILE* f = std::fopen(sp.c_str(), "w");
        if (!f) return 2;
        std::fprintf(f, "level %d\npoints %d\ninner %d\nnnz %ld\n", level, V, n, nnz);
        std::fprintf(f, "solve_converged %d\nsolve_iterations %d\nsolve_rel %.17g\n", si.converged,
                     si.iterations, si.rel_residual);
        std::fprintf(f, "solve_cycle_launches %d\n", si.cycle_launches);
        std::fprintf(f, "lap_solve_converged %d\nlap_solve_rel %.17g\n", li.converged,
                     li.rel_residual);
        std::fprintf(f, "norm_u %.17g\nscalar_uv %.17g\ndist_uv %.17g\n", nu, suv, duv);
        std::fprintf(f, "bv_converged %d\nbv_iterations %d\n", bi.converged, bi.iterations);
        std::fprintf(f, "bv_rp_converged %d\nbv_rp_iterations %d\n", bri.converged,
                     bri.iterations);
        std::fprintf(f, "bc_converged %d\nbc_iterations %d\n", bci.converged, bci.iterations);
        std::fprintf(f, "sc_rel_residual %.17g\n", sci.rel_residual);
        std::fclose(f);
        std::printf("mirror_check level %d ok\n", level);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mirror_check: %s\n", e.what());
        return 1;
    }
    return 0;
}
