// phelm_gpu.hpp - C++ host classes of the sphere-FEM solve path over the C
// ABI of phelm_gpu.h, with the names, argument meaning and error behaviour of
// the reference's classes:
//
//   Solver<double>   Solver<T> of the external solver library: Solver(n),
//                    add(i, j, a) (accumulates), mult_vector(y, x),
//                    solve(x, b) (call sites e.g. src/generators.cpp:49-55,
//                    src/impl/slaplace_impl.h:109,171)
//   SphereLaplace    src/slaplace.h:60-108, src/impl/slaplace_impl.h:84-192
//   SphereJacobian   src/sjacobian.h:65-149
//   SphereNorm       src/norm.h:58-120 with fast_scalar/norm/dist
//                    (src/mesh.h:550-591)
//   SphereChafe      examples/schafe.h:52-97, examples/schafe_impl.h:102-206
//   BarVortex        examples/barvortex.h:62-155 (calc only)
//   Baroclin         examples/baroclin.h:60-292 (calc only)
//
// Vector arguments are device pointers (the reference's GPU build passes
// device arrays, examples/schafe_impl.h:134-149); ArrayDevice<T> owns one.
// Errors throw phelm_gpu::Error (a std::runtime_error, as the solver
// library throws).  Header only; link with libphelm_hip.so and amdhip64.
#ifndef PHELM_GPU_HPP
#define PHELM_GPU_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "phelm_gpu.h"

namespace phelm_gpu {

class Error : public std::runtime_error {
public:
    Error(const std::string& what, int code) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline int check(int rc, const char* what)
{
    if (rc < 0) throw Error(std::string(what) + ": " + phg_last_error(), rc);
    return rc;
}

inline void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e), PHG_ERR_HIP);
}

// phelm::ArrayDevice<T>: a zero-initialised device array.
template <typename T>
class ArrayDevice {
public:
    ArrayDevice() = default;
    explicit ArrayDevice(size_t n) { resize(n); }
    ArrayDevice(const ArrayDevice&) = delete;
    ArrayDevice& operator=(const ArrayDevice&) = delete;
    ArrayDevice(ArrayDevice&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    ArrayDevice& operator=(ArrayDevice&& o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~ArrayDevice()
    {
        if (p_) (void)hipFree(p_);
    }
    void resize(size_t n)
    {
        if (p_) check_hip(hipFree(p_), "hipFree");
        p_ = nullptr;
        n_ = n;
        if (n) {
            check_hip(hipMalloc((void**)&p_, n * sizeof(T)), "hipMalloc");
            check_hip(hipMemset(p_, 0, n * sizeof(T)), "hipMemset");
        }
    }
    size_t size() const { return n_; }
    T* data() { return p_; }
    const T* data() const { return p_; }
    void upload(const T* h) { check_hip(hipMemcpy(p_, h, n_ * sizeof(T), hipMemcpyHostToDevice), "upload"); }
    void download(T* h) const
    {
        check_hip(hipMemcpy(h, p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "download");
    }
    std::vector<T> host() const
    {
        std::vector<T> v(n_);
        if (n_) download(v.data());
        return v;
    }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Mesh (src/mesh.h:354-427): Mesh::load of a text, or the icosahedral sphere
// of mke_mesh_sphere --type full --coord local --iter level.
class Mesh {
public:
    static Mesh sphere(int level)
    {
        phg_mesh_t h = nullptr;
        check(phg_mesh_sphere(level, &h), "phg_mesh_sphere");
        return Mesh(h);
    }
    static Mesh load(const std::string& text)
    {
        phg_mesh_t h = nullptr;
        check(phg_mesh_load(text.data(), text.size(), &h), "phg_mesh_load");
        return Mesh(h);
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    Mesh(Mesh&& o) noexcept { *this = std::move(o); }
    Mesh& operator=(Mesh&& o) noexcept
    {
        std::swap(h_, o.h_);
        ps_ = o.ps_, tr_ = o.tr_, inner_ = o.inner_, outer_ = o.outer_;
        return *this;
    }
    ~Mesh()
    {
        if (h_) phg_mesh_destroy(h_);
    }
    phg_mesh_t handle() const { return h_; }
    int ps_size() const { return ps_; }       // m.ps.size()
    int tr_size() const { return tr_; }       // m.tr.size()
    int inner_size() const { return inner_; } // m.inner.size()
    int outer_size() const { return outer_; } // m.outer.size()

private:
    explicit Mesh(phg_mesh_t h) : h_(h)
    {
        int z = 0;
        check(phg_mesh_info(h_, &ps_, &tr_, &inner_, &outer_, &z), "phg_mesh_info");
    }
    phg_mesh_t h_ = nullptr;
    int ps_ = 0, tr_ = 0, inner_ = 0, outer_ = 0;
};

template <typename T>
class Solver;

// Solver<double>: add() stages entries on the host (safe from several threads
// for distinct rows, as generate_matrix calls it); the first mult_vector or
// solve builds the device matrix, and a later add() rebuilds it.  solve()
// runs GMRES(30) with modified Gram-Schmidt to a relative residual of 1e-10
// from x = 0 unless opts() says otherwise; a solve that ends without
// converging leaves its best x, as the solver library's does.
template <>
class Solver<double> {
public:
    typedef double data_type;

    explicit Solver(int n) : Solver(n, n) {}
    Solver(int n_rows, int n_cols) : n_rows_(n_rows), n_cols_(n_cols), rows_((size_t)n_rows)
    {
        phg_gmres_default_opts(&opts_);
    }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;
    Solver(Solver&& o) noexcept { *this = std::move(o); }
    Solver& operator=(Solver&& o) noexcept
    {
        std::swap(h_, o.h_);
        std::swap(owned_, o.owned_);
        n_rows_ = o.n_rows_, n_cols_ = o.n_cols_;
        rows_ = std::move(o.rows_);
        opts_ = o.opts_, info_ = o.info_;
        stream_ = o.stream_;
        dirty_.store(o.dirty_.load());
        return *this;
    }
    ~Solver() { release(); }

    // Matrices assembled on the GPU (generate_matrix / _boundary_ / _full_).
    static Solver assemble(const Mesh& m, int kind, double c_mass = 1.0, double c_lap = 0.0)
    {
        phg_matrix_t h = nullptr;
        check(phg_assemble(m.handle(), kind, c_mass, c_lap, nullptr, &h), "phg_assemble");
        return Solver(h, true);
    }
    static Solver assemble_boundary(const Mesh& m, int kind, double c_mass = 1.0,
                                    double c_lap = 0.0)
    {
        phg_matrix_t h = nullptr;
        check(phg_assemble_boundary(m.handle(), kind, c_mass, c_lap, &h),
              "phg_assemble_boundary");
        return Solver(h, true);
    }
    static Solver assemble_full(const Mesh& m, int kind, double c_mass = 1.0, double c_lap = 0.0)
    {
        phg_matrix_t h = nullptr;
        check(phg_assemble_full(m.handle(), kind, c_mass, c_lap, nullptr, &h),
              "phg_assemble_full");
        return Solver(h, true);
    }
    // A view of a matrix owned elsewhere (e.g. by SphereJacobian).
    static Solver view(phg_matrix_t h) { return Solver(h, false); }

    void add(int i, int j, double a)
    {
        if (i < 0 || i >= n_rows_ || j < 0 || j >= n_cols_)
            throw Error("Solver::add: index out of range", PHG_ERR_ARG);
        rows_[(size_t)i].emplace_back(j, a);
        dirty_ = true;
    }
    void mult_vector(double* y, const double* x)
    {
        check(phg_spmv(handle(), x, y, stream_), "Solver::mult_vector");
    }
    void solve(double* x, const double* b)
    {
        check(phg_gmres(handle(), b, x, &opts_, &info_, stream_), "Solver::solve");
    }

    int n_rows() const { return n_rows_; }
    int n_cols() const { return n_cols_; }
    phg_gmres_opts& opts() { return opts_; }
    const phg_gmres_info& last_info() const { return info_; }
    void set_stream(phg_stream_t s) { stream_ = s; }
    phg_stream_t stream() const { return stream_; }
    phg_matrix_t handle()
    {
        if (dirty_ || !h_) build();
        return h_;
    }

private:
    Solver(phg_matrix_t h, bool owned) : h_(h), owned_(owned)
    {
        long nnz = 0;
        check(phg_matrix_info(h_, &n_rows_, &n_cols_, &nnz), "phg_matrix_info");
        phg_gmres_default_opts(&opts_);
        dirty_ = false;
    }
    void release()
    {
        if (h_ && owned_) phg_matrix_destroy(h_);
        h_ = nullptr;
    }
    void build()
    {
        std::vector<int> r, c;
        std::vector<double> v;
        for (size_t i = 0; i < rows_.size(); ++i)
            for (const auto& e : rows_[i]) {
                r.push_back((int)i);
                c.push_back(e.first);
                v.push_back(e.second);
            }
        phg_matrix_t h = nullptr;
        check(phg_matrix_from_triplets(n_rows_, n_cols_, (long)r.size(), r.data(), c.data(),
                                       v.data(), &h),
              "Solver: building the device matrix");
        release();
        h_ = h;
        owned_ = true;
        dirty_ = false;
    }

    phg_matrix_t h_ = nullptr;
    bool owned_ = true;
    std::atomic<bool> dirty_{true};  // add() may run on several threads
    int n_rows_ = 0, n_cols_ = 0;
    std::vector<std::vector<std::pair<int, double>>> rows_;
    phg_gmres_opts opts_{};
    phg_gmres_info info_{};
    phg_stream_t stream_ = nullptr;
};

// SphereLaplace<T, Matrix> (src/slaplace.h:60-108).  solve() runs the
// Poisson solve with the multigrid-preconditioned GMRES on laplace_'s
// operator (relative residual solve_rtol, default 1e-10).
template <typename T = double, typename Matrix = Solver<T>>
class SphereLaplace {
public:
    Matrix idt_;
    Matrix laplace_;
    Matrix bnd1_;  // laplace_bnd1_cb (mass), inner x outer
    Matrix bnd2_;  // laplace_bnd2_cb (-stiffness)
    Matrix bnd3_;  // slaplace_integrate_cb (stiffness)
    const Mesh& m_;
    double solve_rtol = 1e-10;

    explicit SphereLaplace(const Mesh& m)
        : idt_(Matrix::assemble(m, PHG_MASS)),
          laplace_(Matrix::assemble(m, PHG_LAPLACE)),
          bnd1_(Matrix::assemble_boundary(m, PHG_MASS)),
          bnd2_(Matrix::assemble_boundary(m, PHG_HELMHOLTZ, 0.0, -1.0)),
          bnd3_(Matrix::assemble_boundary(m, PHG_LAPLACE)),
          m_(m),
          in_((size_t)m.inner_size()),
          out_((size_t)m.inner_size()),
          tmp_((size_t)std::max(1, m.outer_size()))
    {
        idt_.opts().ortho = PHG_CGS2;
        idt_.opts().rtol = 1e-13;
    }
    SphereLaplace(const SphereLaplace&) = delete;
    SphereLaplace& operator=(const SphereLaplace&) = delete;
    ~SphereLaplace()
    {
        if (mg_) phg_mg_destroy(mg_);
    }

    // Ans (inner) = idt^-1 (laplace F_inner + bnd3 F_outer)  (calc2)
    void calc2(T* Ans, const T* F)
    {
        check(phg_u2p(m_.handle(), F, in_.data(), nullptr), "u2p");
        check(phg_proj_bnd(m_.handle(), F, tmp_.data(), nullptr), "proj_bnd");
        laplace_.mult_vector(out_.data(), in_.data());
        bnd3_.mult_vector(in_.data(), tmp_.data());
        check(phg_axpby(m_.inner_size(), 1.0, out_.data(), 1.0, in_.data(), out_.data(), nullptr),
              "vec_sum");
        idt_.solve(Ans, out_.data());
    }
    // calc2 then p2u with the boundary values bnd (calc1)
    void calc1(T* Ans, const T* F, const T* bnd)
    {
        calc2(out2(), F);
        check(phg_p2u(m_.handle(), out2(), bnd, Ans, nullptr), "p2u");
    }
    // laplace x = idt F_inner (+ bnd2 bnd), Ans = p2u(x, bnd)  (solve)
    void solve(T* Ans, const T* F, const T* bnd)
    {
        check(phg_u2p(m_.handle(), F, in_.data(), nullptr), "u2p");
        idt_.mult_vector(out_.data(), in_.data());
        if (bnd) {
            bnd2_.mult_vector(in_.data(), bnd);
            check(phg_axpby(m_.inner_size(), 1.0, out_.data(), 1.0, in_.data(), out_.data(),
                            nullptr),
                  "vec_sum");
        }
        if (!mg_) check(phg_mg_create(m_.handle(), PHG_LAPLACE, 1.0, 0.0, nullptr, &mg_),
                        "phg_mg_create");
        phg_gmres_opts o;
        phg_gmres_default_opts(&o);
        o.ortho = PHG_CGS2;
        o.rtol = solve_rtol;
        o.max_iters = 2000;
        check(phg_gmres_batch_mg(mg_, 1, out_.data(), in_.data(), &o, &info_, nullptr),
              "SphereLaplace::solve");
        check(phg_p2u(m_.handle(), in_.data(), bnd, Ans, nullptr), "p2u");
    }
    const phg_gmres_info& last_info() const { return info_; }

private:
    T* out2()
    {
        if (out2_.size() == 0) out2_.resize((size_t)m_.inner_size());
        return out2_.data();
    }
    ArrayDevice<T> in_, out_, tmp_, out2_;
    phg_mg_t mg_ = nullptr;
    phg_gmres_info info_{};
};

// SphereJacobian (src/sjacobian.h:65-149): the cubature-7 matrices on the
// GPU; calc2 / calc2t solve the 8 distinct mass systems as one batch.
class SphereJacobian {
public:
    explicit SphereJacobian(const Mesh& m) : m_(m), tmp_((size_t)m.inner_size())
    {
        check(phg_jacobian_create(m.handle(), nullptr, &h_), "phg_jacobian_create");
    }
    SphereJacobian(const SphereJacobian&) = delete;
    SphereJacobian& operator=(const SphereJacobian&) = delete;
    ~SphereJacobian()
    {
        if (h_) phg_jacobian_destroy(h_);
    }
    // J(u, v) at the inner points (u, v: every point)
    void calc2(double* Ans, const double* u, const double* v)
    {
        check(phg_jacobian_calc2(h_, 1, u, v, Ans, nullptr, nullptr, nullptr),
              "SphereJacobian::calc2");
    }
    void calc2t(double* Ans, const double* u, const double* v)
    {
        check(phg_jacobian_calc2t(h_, 1, u, v, Ans, nullptr, nullptr, nullptr),
              "SphereJacobian::calc2t");
    }
    // calc2 then p2u with the boundary values bnd (calc1 / calc1t)
    void calc1(double* Ans, const double* u, const double* v, const double* bnd)
    {
        calc2(tmp_.data(), u, v);
        check(phg_p2u(m_.handle(), tmp_.data(), bnd, Ans, nullptr), "p2u");
    }
    void calc1t(double* Ans, const double* u, const double* v, const double* bnd)
    {
        calc2t(tmp_.data(), u, v);
        check(phg_p2u(m_.handle(), tmp_.data(), bnd, Ans, nullptr), "p2u");
    }
    phg_jacobian_t handle() const { return h_; }

private:
    const Mesh& m_;
    phg_jacobian_t h_ = nullptr;
    ArrayDevice<double> tmp_;
};

// SphereNorm<T> (src/norm.h:58-120): NORM_ = generate_full_matrix with
// sphere_scalar_cb; scalar(u, v) = u . (NORM_ v) (fast_scalar).
template <typename T = double>
class SphereNorm {
public:
    typedef Solver<T> Matrix;
    const Mesh& m_;
    Matrix NORM_;

    explicit SphereNorm(const Mesh& m)
        : m_(m), NORM_(Matrix::assemble_full(m, PHG_MASS)), tmp_((size_t)m.ps_size())
    {
    }
    T scalar(const T* u, const T* v)
    {
        NORM_.mult_vector(tmp_.data(), v);
        double r = 0.0;
        check(phg_dot(m_.ps_size(), u, tmp_.data(), &r, nullptr), "vec_scalar2");
        return r;
    }
    T norm(const T* u) { return std::sqrt(scalar(u, u)); }
    T dist(const T* u, const T* v)
    {
        if (diff_.size() == 0) diff_.resize((size_t)m_.ps_size());
        check(phg_axpby(m_.ps_size(), 1.0, u, -1.0, v, diff_.data(), nullptr), "vec_diff");
        return norm(diff_.data());
    }

private:
    ArrayDevice<T> tmp_, diff_;
};

// SphereChafe_Private::f (examples/schafe.cpp:55-63): the forcing for which
// x sin(y + t) cos^4 x solves du/dt = mu Delta u - sigma u + f on the sphere
// (examples/test_schafe.cpp:30-34).  ipow(c, 2) is (1 c) c, as the
// reference's repeated multiplication gives.
inline double schafe_f(double x, double y, double t, double mu, double sigma)
{
    const double c2 = (1.0 * std::cos(x)) * std::cos(x);
    return c2 * (x * std::cos(y + t) * c2 + sigma * std::sin(y + t) * c2 * x +
                 9 * mu * std::sin(y + t) * std::sin(x) * std::cos(x) -
                 15 * mu * std::sin(y + t) * x + 20 * mu * std::sin(y + t) * x * c2);
}

// SphereChafe<T> (examples/schafe.h:52-97): du/dt = mu Delta u - sigma u + f,
// one Crank-Nicolson step per solve() (examples/schafe_impl.h:126-206).
// A_ = (1/tau + sigma/2) M - (mu/2) L and its boundary matrix bnd_
// (schafe_integrate_cb, examples/schafe.cpp:67-89) are GPU-assembled
// Helmholtz kinds; the right part u/tau + (mu/2) calc2(u) - (sigma/2) u + f
// (f at the inner points' zone-0 chart coordinates, evaluated on the host as
// the reference does), times idt_, plus bnd_ bnd; then A_ x = right part by
// GMRES and Ans = p2u(x, bnd).  Ans may alias X0.
template <typename T = double>
class SphereChafe {
public:
    typedef double (*rp_t)(double x, double y, double t, double mu, double sigma);
    double tau_, sigma_, mu_;
    const Mesh& m_;
    SphereLaplace<T> laplace_;
    Solver<T> A_;
    Solver<T> bnd_;

    SphereChafe(const Mesh& m, double tau, double sigma, double mu, rp_t f = schafe_f)
        : tau_(tau), sigma_(sigma), mu_(mu), m_(m), laplace_(m),
          A_(Solver<T>::assemble(m, PHG_HELMHOLTZ, 1.0 / tau + sigma * 0.5, -0.5 * mu)),
          bnd_(Solver<T>::assemble_boundary(m, PHG_HELMHOLTZ, 1.0 / tau + sigma * 0.5,
                                            -0.5 * mu)),
          f_(f), u_((size_t)m.inner_size()), du_((size_t)m.inner_size()),
          crp_((size_t)m.inner_size()), tmp_((size_t)m.inner_size()),
          x_((size_t)m.inner_size()), rp_((size_t)m.inner_size())
    {
        int V = 0, Tn = 0, ni = 0, no = 0, nz = 0;
        check(phg_mesh_info(m.handle(), &V, &Tn, &ni, &no, &nz), "phg_mesh_info");
        std::vector<double> coords((size_t)V * nz * 2);
        std::vector<int> inner((size_t)std::max(ni, 1));
        check(phg_mesh_arrays(m.handle(), nullptr, coords.data(), nullptr, nullptr, nullptr,
                              nullptr, nullptr, inner.data(), nullptr, nullptr),
              "phg_mesh_arrays");
        x_in_.resize((size_t)ni);
        y_in_.resize((size_t)ni);
        for (int i = 0; i < ni; ++i) {
            x_in_[i] = coords[(size_t)inner[i] * nz * 2];
            y_in_[i] = coords[(size_t)inner[i] * nz * 2 + 1];
        }
    }
    SphereChafe(const SphereChafe&) = delete;
    SphereChafe& operator=(const SphereChafe&) = delete;

    void solve(T* Ans, const T* X0, const T* bnd, double t)
    {
        const int rs = m_.inner_size();
        check(phg_u2p(m_.handle(), X0, u_.data(), nullptr), "u2p");
        laplace_.calc2(du_.data(), X0);
        // u/dt + mu Delta u / 2, then - sigma u / 2 (vec_sum1)
        check(phg_axpby(rs, (T)(1.0 / tau_), u_.data(), (T)(mu_ * 0.5), du_.data(), du_.data(),
                        nullptr),
              "vec_sum1");
        check(phg_axpby(rs, (T)1.0, du_.data(), (T)(-sigma_ * 0.5), u_.data(), du_.data(),
                        nullptr),
              "vec_sum1");
        // + f (vec_sum)
        std::vector<double> h((size_t)rs);
        for (int i = 0; i < rs; ++i) h[i] = f_ ? f_(x_in_[i], y_in_[i], t, mu_, sigma_) : 0.0;
        rp_.upload(h.data());
        check(phg_axpby(rs, 1.0, du_.data(), 1.0, rp_.data(), u_.data(), nullptr), "vec_sum");
        // idt_ u (+ bnd_ bnd), then the inner solve and p2u
        laplace_.idt_.mult_vector(crp_.data(), u_.data());
        if (bnd) {
            bnd_.mult_vector(tmp_.data(), bnd);
            check(phg_axpby(rs, 1.0, crp_.data(), 1.0, tmp_.data(), crp_.data(), nullptr),
                  "vec_sum");
        }
        A_.solve(x_.data(), crp_.data());
        check(phg_p2u(m_.handle(), x_.data(), bnd, Ans, nullptr), "p2u");
    }

private:
    rp_t f_;
    ArrayDevice<T> u_, du_, crp_, tmp_, x_, rp_;
    std::vector<double> x_in_, y_in_;
};

// BarVortexConf (examples/barvortex.h:62-96) with the reference's callbacks:
// rp(phi, lambda, t, mu, sigma) is the right part, coriolis(phi, lambda) the
// Coriolis + orography term l + h; either may be null.  Both are evaluated at
// the mesh points' zone-0 chart coordinates, as the reference does.
struct BarVortexConf {
    typedef double (*rp_t)(double phi, double lambda, double t, double mu, double sigma);
    typedef double (*coriolis_t)(double phi, double lambda);
    double tau, sigma, mu, k1, k2, theta;
    rp_t rp = nullptr;
    coriolis_t coriolis = nullptr;
};

// BarVortex (examples/barvortex.h:62-240): calc() is one implicit step
// (examples/barvortex_impl.h:171-310) on the device, calc_L / calc_L_1 /
// calc_LT the linearised, inverse-linearised and adjoint steps
// (:326-535) and S_step / L_step / L_1_step / LT_step their wrappers
// (:537-571).  calc, calc_L, calc_L_1 and calc_LT take the boundary (pole)
// values bnd_u / bnd_w (device pointers, outer-size each, null = zeros).
class BarVortex {
public:
    // the Coriolis term as a device mesh function (the callbacks of conf are
    // not used for it)
    BarVortex(const Mesh& m, const BarVortexConf& conf, const double* d_lh) : conf_(conf)
    {
        init(m, d_lh);
    }
    // BarVortex(m, conf) of the reference: lh = proj(conf.coriolis)
    BarVortex(const Mesh& m, const BarVortexConf& conf) : conf_(conf)
    {
        read_xy(m);
        std::vector<double> lh(x_.size(), 0.0);
        if (conf.coriolis)
            for (size_t i = 0; i < lh.size(); ++i) lh[i] = conf.coriolis(x_[i], y_[i]);
        ArrayDevice<double> dlh(lh.size());
        dlh.upload(lh.data());
        init(m, dlh.data());
    }
    BarVortex(const BarVortex&) = delete;
    BarVortex& operator=(const BarVortex&) = delete;
    ~BarVortex()
    {
        if (h_) phg_barvortex_destroy(h_);
    }
    void calc(double* Ans, const double* F, const double* bnd_u, const double* bnd_w, double t)
    {
        const double* rp = nullptr;
        if (conf_.rp) {
            const size_t V = x_.size();
            std::vector<double> h(V);
            for (size_t i = 0; i < V; ++i) h[i] = conf_.rp(x_[i], y_[i], t, conf_.mu, conf_.sigma);
            rp_.upload(h.data());
            rp = rp_.data();
        }
        check(phg_barvortex_calc_bnd(h_, 1, F, rp, bnd_u, bnd_w, Ans, &opts_, &info_, nullptr),
              "BarVortex::calc");
    }
    // bnd_u, bnd_w: device arrays of outer_size boundary values (or null)
    void calc_L(double* u1, const double* u, const double* z, const double* bnd_u,
                const double* bnd_w, double)
    {
        check(phg_barvortex_calc_L_bnd(h_, 1, u, z, bnd_u, bnd_w, u1, &opts_, &info_, nullptr),
              "BarVortex::calc_L");
    }
    // adds to u1's content, as the reference does
    void calc_L_1(double* u1, const double* u, const double* z, const double* bnd_u,
                  const double* bnd_w, double)
    {
        check(phg_barvortex_calc_L_1_bnd(h_, 1, u, z, bnd_u, bnd_w, u1, &opts_, &info_, nullptr),
              "BarVortex::calc_L_1");
    }
    void calc_LT(double* v1, const double* v, const double* z, const double* bnd_u,
                 const double* bnd_w, double)
    {
        check(phg_barvortex_calc_LT_bnd(h_, 1, v, z, bnd_u, bnd_w, v1, &opts_, &info_, nullptr),
              "BarVortex::calc_LT");
    }
    void S_step(double* Ans, const double* F) { calc(Ans, F, nullptr, nullptr, 0.0); }
    void L_step(double* Ans, const double* F, const double* z)
    {
        ArrayDevice<double> tmp(x_.size());
        calc_L(tmp.data(), F, z, nullptr, nullptr, 0.0);
        copy_dev(Ans, tmp.data());
    }
    void L_1_step(double* Ans, const double* F, const double* z)
    {
        ArrayDevice<double> tmp(x_.size());  // zeros
        calc_L_1(tmp.data(), F, z, nullptr, nullptr, 0.0);
        copy_dev(Ans, tmp.data());
    }
    void LT_step(double* Ans, const double* F, const double* z)
    {
        ArrayDevice<double> tmp(x_.size());
        calc_LT(tmp.data(), F, z, nullptr, nullptr, 0.0);
        copy_dev(Ans, tmp.data());
    }
    phg_bv_opts& opts() { return opts_; }
    const phg_bv_info& last_info() const { return info_; }

private:
    void read_xy(const Mesh& m)
    {
        int V = 0, T = 0, ni = 0, no = 0, nz = 0;
        check(phg_mesh_info(m.handle(), &V, &T, &ni, &no, &nz), "phg_mesh_info");
        std::vector<double> coords((size_t)V * nz * 2);
        check(phg_mesh_arrays(m.handle(), nullptr, coords.data(), nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr, nullptr),
              "phg_mesh_arrays");
        x_.resize((size_t)V);
        y_.resize((size_t)V);
        for (int i = 0; i < V; ++i) {
            x_[i] = coords[(size_t)i * nz * 2];
            y_[i] = coords[(size_t)i * nz * 2 + 1];
        }
    }
    void init(const Mesh& m, const double* d_lh)
    {
        if (x_.empty()) read_xy(m);
        phg_bv_conf c{conf_.tau, conf_.sigma, conf_.mu, conf_.k1, conf_.k2, conf_.theta};
        check(phg_barvortex_create(m.handle(), &c, d_lh, 1, nullptr, &h_), "BarVortex");
        phg_bv_default_opts(&opts_);
        rp_.resize(x_.size());
    }
    void copy_dev(double* dst, const double* src) const
    {
        check_hip(hipMemcpy(dst, src, x_.size() * sizeof(double), hipMemcpyDeviceToDevice),
                  "BarVortex copy");
    }

    BarVortexConf conf_;
    phg_barvortex_t h_ = nullptr;
    phg_bv_opts opts_{};
    phg_bv_info info_{};
    std::vector<double> x_, y_;
    ArrayDevice<double> rp_;
};

// Baroclin (examples/baroclin.h:60-292): the constructor takes the
// reference's callbacks; calc() evaluates f and g at the mesh points' zone-0
// chart coordinates for its t (as baroclin.cpp:520-535 does) and runs the
// step on the device.  calc takes the boundary (pole) values bnd (one
// device array of outer-size for w1, w2, u1, u2, as the reference; null =
// zeros); the linearised steps take zero boundary values, as the reference
// driver passes them (bnd null or all zero; it is not read there).
class Baroclin {
public:
    typedef double (*rp_t)(double x, double y, double t, double sigma, double mu, double sigma1,
                           double mu1, double alpha);
    typedef double (*coriolis_t)(double phi, double lambda);

    Baroclin(const Mesh& m, rp_t f, rp_t g, coriolis_t coriolis, double tau, double sigma,
             double mu, double sigma1, double mu1, double alpha)
        : m_(m), f_(f), g_(g), conf_{tau, sigma, mu, sigma1, mu1, alpha}
    {
        int V = 0, T = 0, ni = 0, no = 0, nz = 0;
        check(phg_mesh_info(m.handle(), &V, &T, &ni, &no, &nz), "phg_mesh_info");
        std::vector<double> coords((size_t)V * nz * 2);
        check(phg_mesh_arrays(m.handle(), nullptr, coords.data(), nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr, nullptr),
              "phg_mesh_arrays");
        x_.resize((size_t)V);
        y_.resize((size_t)V);
        os_ = no;
        std::vector<double> lh((size_t)V);
        for (int i = 0; i < V; ++i) {
            x_[i] = coords[(size_t)i * nz * 2];
            y_[i] = coords[(size_t)i * nz * 2 + 1];
            lh[i] = coriolis ? coriolis(x_[i], y_[i]) : 0.0;
        }
        ArrayDevice<double> dlh(lh.size());
        dlh.upload(lh.data());
        fv_.resize((size_t)V);
        gv_.resize((size_t)V);
        check(phg_baroclin_create(m.handle(), &conf_, dlh.data(), 1, nullptr, &h_),
              "Baroclin");
        phg_bc_default_opts(&opts_);
    }
    Baroclin(const Baroclin&) = delete;
    Baroclin& operator=(const Baroclin&) = delete;
    ~Baroclin()
    {
        if (h_) phg_baroclin_destroy(h_);
    }
    // (u1, u2) -> (u11, u21), device pointers of m.ps.size() values
    void calc(double* u11, double* u21, const double* u1, const double* u2, const double* bnd,
              double t)
    {
        const size_t V = x_.size();
        std::vector<double> hf(V), hg(V);
        for (size_t i = 0; i < V; ++i) {
            const Config c = conf_c();
            hf[i] = f_ ? f_(x_[i], y_[i], t, c.sigma, c.mu, c.sigma1, c.mu1, c.alpha) : 0.0;
            hg[i] = g_ ? g_(x_[i], y_[i], t, c.sigma, c.mu, c.sigma1, c.mu1, c.alpha) : 0.0;
        }
        fv_.upload(hf.data());
        gv_.upload(hg.data());
        check(phg_baroclin_calc_bnd(h_, 1, u1, u2, f_ ? fv_.data() : nullptr,
                                    g_ ? gv_.data() : nullptr, os_ ? bnd : nullptr, u11, u21,
                                    &opts_, &info_, nullptr),
              "Baroclin::calc");
    }
    // the linearised steps about (z1, z2) (baroclin.cpp:607-986); bnd as calc
    // (device, n_outer values, or null for zeros)
    void calc_L(double* u11, double* u21, const double* u1, const double* u2, const double* z1,
                const double* z2, const double* bnd, double)
    {
        check(phg_baroclin_calc_L_bnd(h_, 1, u1, u2, z1, z2, os_ ? bnd : nullptr, u11, u21,
                                      &opts_, &info_, nullptr),
              "Baroclin::calc_L");
    }
    void calc_L_1(double* u11, double* u21, const double* u1, const double* u2,
                  const double* z1, const double* z2, const double* bnd, double)
    {
        check(phg_baroclin_calc_L_1_bnd(h_, 1, u1, u2, z1, z2, os_ ? bnd : nullptr, u11, u21,
                                        &opts_, &info_, nullptr),
              "Baroclin::calc_L_1");
    }
    // the reference's calc_LT has an empty body (baroclin.cpp:988-993)
    void calc_LT(double*, double*, const double*, const double*, const double*, const double*,
                 const double*, double)
    {
    }
    void L_step(double* u11, double* u21, const double* u1, const double* u2, const double* z1,
                const double* z2)
    {
        calc_L(u11, u21, u1, u2, z1, z2, nullptr, 0.0);
    }
    void L_1_step(double* u11, double* u21, const double* u1, const double* u2,
                  const double* z1, const double* z2)
    {
        calc_L_1(u11, u21, u1, u2, z1, z2, nullptr, 0.0);
    }
    void LT_step(double* u11, double* u21, const double* u1, const double* u2, const double* z1,
                 const double* z2)
    {
        calc_LT(u11, u21, u1, u2, z1, z2, nullptr, 0.0);
    }
    phg_bc_opts& opts() { return opts_; }
    const phg_bc_info& last_info() const { return info_; }

private:
    struct Config {
        double sigma, mu, sigma1, mu1, alpha;
    };
    Config conf_c() const { return {conf_.sigma, conf_.mu, conf_.sigma1, conf_.mu1, conf_.alpha}; }
    const Mesh& m_;
    rp_t f_, g_;
    phg_bc_conf conf_;
    std::vector<double> x_, y_;
    ArrayDevice<double> fv_, gv_;
    phg_baroclin_t h_ = nullptr;
    phg_bc_opts opts_{};
    phg_bc_info info_{};
    int os_ = 0;
};

}  // namespace phelm_gpu

#endif  // PHELM_GPU_HPP
