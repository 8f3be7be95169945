// Geometric multigrid over the nested icosahedral sphere meshes, used as the
// right preconditioner of the batched GMRES for the operators whose Krylov
// convergence degrades with the level: the pole-pinned Laplace stiffness of
// the Poisson solve SphereLaplace::solve (src/impl/slaplace_impl.h:84-124,
// called once per fixed-point iteration by BarVortex::calc,
// examples/barvortex_impl.h:298).  Unpreconditioned GMRES(30) needs about
// 15,000 iterations at level 5 and stalls at level 6; preconditioned by one
// V-cycle per step it needs 8-12 iterations at levels 3-6.
//
// Hierarchy: level k+1 of mke_mesh_sphere keeps every level-k vertex (same
// coordinates) and adds the projected midpoint of every edge.  Prolongation
// copies a coarse value to its own vertex and gives a midpoint the mean of
// its edge's two end points (the pole, a Dirichlet point, contributes 0);
// restriction is its transpose.  Each level's operator is re-assembled on its
// own mesh (P1 stiffness is scale-free, so this is the Galerkin operator up to
// the sphere's curvature).  Smoother: damped Jacobi (weight 0.8), two sweeps
// before and after the coarse correction; the coarsest level (2, 161
// unknowns) is solved with its inverse.
//
// All vectors are n x KP row-major like the batched GMRES; kernels map one
// thread to one (row, column) and need no reduction, so the V-cycle is a
// fixed kernel sequence that is captured with the GMRES cycle.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include <array>

#include "ell.h"
#include "phg_internal.h"

using namespace phg;

namespace {
constexpr int MG_COARSE = 2;
constexpr double MG_OMEGA = 0.8;
}  // namespace

struct MgLevel {
    int n = 0;
    phg_mesh_s* mesh = nullptr;   // owned for the coarse levels
    bool own_mesh = false;
    phg_matrix_s* A = nullptr;    // level operator (owned)
    DevBuf<double> dinv;
    DevBuf<int> p0, p1;           // prolongation parents (coarse inner ids, -1 none)
    DevBuf<int> rrp, ridx;        // restriction: coarse row j <- fine rows
    DevBuf<double> rw;
    DevBuf<double> x0, x1, b, r;  // n x kp work vectors
    // the same in FP32 for the mixed-precision form (made on first use)
    DevBuf<float> dinvf, rwf, x0f, x1f, bf, rf;
    ~MgLevel()
    {
        delete A;
        if (own_mesh) delete mesh;
    }
};

struct phg_mg_s;
namespace phg {
// One V-cycle as the batched GMRES's right preconditioner.
struct MgPrecond : BatchPrecond {
    phg_mg_s* M = nullptr;
    int rows() const override { return mg_rows(M); }
    int reserve(int kp) override { return mg_reserve(M, kp); }
    void apply(int KP, const double* v, double* z, hipStream_t s) override { mg_apply(M, KP, v, z, s); }
    bool f32() const override { return true; }
    int reserve_f32(int kp) override { return mg_reserve_f32(M, kp); }
    void apply_f32(int KP, const float* v, float* z, hipStream_t s) override
    {
        mg_apply_f32(M, KP, v, z, s);
    }
};
}  // namespace phg

struct phg_mg_s {
    int kind = 0, L = 0, kp = 0;
    double c_mass = 1.0, c_lap = 0.0;
    std::vector<std::unique_ptr<MgLevel>> lv;  // lv[0] = finest
    DevBuf<double> cinv;                       // coarsest operator's inverse
    DevBuf<float> cinvf;                       // the same in FP32
    int nc = 0, kpf = 0;
    phg::MgPrecond pc;
};

namespace {

// The V-cycle kernels in the vector type T (double, or float for the
// mixed-precision form) over matrix values MT.
template <typename T, typename MT>
__global__ void mg_jacobi(int n, int KP, int width, const int* __restrict__ col,
                          const MT* __restrict__ val, const T* __restrict__ dinv,
                          const T* __restrict__ xin, const T* __restrict__ b, T* __restrict__ xout,
                          T w)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * KP) return;
    const int row = (int)(e / KP), q = (int)(e % KP);
    T s = T(0), x = T(0);
    if (xin) {
        int c[ELL_MAXW];
        T v[ELL_MAXW], xv[ELL_MAXW];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = (size_t)k * n + row;
            c[k] = k < width ? col[o] : 0;
            v[k] = k < width ? (T)val[o] : T(0);
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < width ? xin[(size_t)c[k] * KP + q] : T(0);
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k)
            if (k < width) s += v[k] * xv[k];
        x = xin[e];
    }
    xout[e] = x + w * dinv[row] * (b[e] - s);
}

template <typename T, typename MT>
__global__ void mg_residual(int n, int KP, int width, const int* __restrict__ col,
                            const MT* __restrict__ val, const T* __restrict__ x,
                            const T* __restrict__ b, T* __restrict__ r)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * KP) return;
    const int row = (int)(e / KP), q = (int)(e % KP);
    int c[ELL_MAXW];
    T v[ELL_MAXW], xv[ELL_MAXW];
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) {
        const size_t o = (size_t)k * n + row;
        c[k] = k < width ? col[o] : 0;
        v[k] = k < width ? (T)val[o] : T(0);
    }
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < width ? x[(size_t)c[k] * KP + q] : T(0);
    T s = T(0);
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k)
        if (k < width) s += v[k] * xv[k];
    r[e] = b[e] - s;
}

template <typename T>
__global__ void mg_restrict(int nc, int KP, const int* __restrict__ rrp,
                            const int* __restrict__ ridx, const T* __restrict__ rw,
                            const T* __restrict__ rf, T* __restrict__ bc)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nc * KP) return;
    const int j = (int)(e / KP), q = (int)(e % KP);
    T s = T(0);
    for (int t = rrp[j]; t < rrp[j + 1]; ++t) s += rw[t] * rf[(size_t)ridx[t] * KP + q];
    bc[e] = s;
}

template <typename T>
__global__ void mg_prolong_add(int nf, int KP, const int* __restrict__ p0,
                               const int* __restrict__ p1, const T* __restrict__ xc,
                               T* __restrict__ xf)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nf * KP) return;
    const int i = (int)(e / KP), q = (int)(e % KP);
    const int a = p0[i], c = p1[i];
    T v;
    if (c == -2) v = a >= 0 ? xc[(size_t)a * KP + q] : T(0);  // the vertex itself
    else v = T(0.5) * ((a >= 0 ? xc[(size_t)a * KP + q] : T(0)) + (c >= 0 ? xc[(size_t)c * KP + q] : T(0)));
    xf[e] += v;
}

__global__ void mg_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// x = inv b on the coarsest level: one workgroup per row i; thread t takes
// column q = t % KP and the j = t / KP, t / KP + 256 / KP, ... terms, and the
// 256 / KP partial sums of a column are added in a fixed order (a column's
// sum does not involve the other columns).
template <typename T>
__global__ void __launch_bounds__(256)
mg_dense(int nc, int KP, const T* __restrict__ inv, const T* __restrict__ b, T* __restrict__ x)
{
    __shared__ T part[256];
    const int i = blockIdx.x, t = threadIdx.x;
    const int q = t % KP, g = t / KP, ng = 256 / KP;
    T s = T(0);
    for (int j = g; j < nc; j += ng) s += inv[(size_t)i * nc + j] * b[(size_t)j * KP + q];
    part[t] = s;
    __syncthreads();
    if (t < KP) {
        T a = T(0);
        for (int k = 0; k < ng; ++k) a += part[k * KP + t];
        x[(size_t)i * KP + t] = a;
    }
}

__global__ void mg_diag_inv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                            const double* __restrict__ val, double* __restrict__ dinv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int t = rp[i]; t < rp[i + 1]; ++t)
        if (ci[t] == i) d = val[t];
    dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
}

inline dim3 grid_for(size_t work) { return dim3((unsigned)((work + 255) / 256)); }

// Gauss-Jordan inverse with partial pivoting (n <= a few hundred).
bool dense_inverse(int n, std::vector<double>& a, std::vector<double>& inv)
{
    inv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(a[(size_t)r * n + c]) > std::fabs(a[(size_t)p * n + c])) p = r;
        if (a[(size_t)p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) {
                std::swap(a[(size_t)p * n + k], a[(size_t)c * n + k]);
                std::swap(inv[(size_t)p * n + k], inv[(size_t)c * n + k]);
            }
        const double d = 1.0 / a[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) {
            a[(size_t)c * n + k] *= d;
            inv[(size_t)c * n + k] *= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
                inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
            }
        }
    }
    return true;
}

int level_of(const SphereMesh& m)
{
    for (int k = 0; k <= 12; ++k)
        if ((long)m.n_points == 10L * (1L << (2 * k)) + 2) return k;
    return -1;
}

// Prolongation parents of every fine inner row and the restriction rows.
int build_transfer(const SphereMesh& f, const SphereMesh& c, MgLevel& fl, int nc)
{
    std::map<std::array<double, 3>, int> where;
    for (int p = 0; p < c.n_points; ++p)
        where[{c.pr[3 * p], c.pr[3 * p + 1], c.pr[3 * p + 2]}] = p;
    const int nf = (int)f.inner.size();
    std::vector<int> p0(nf, -1), p1(nf, -1);
    auto coarse_of = [&](int p) {
        auto it = where.find({f.pr[3 * p], f.pr[3 * p + 1], f.pr[3 * p + 2]});
        return it == where.end() ? -1 : it->second;
    };
    auto cin = [&](int cp) { return c.is_boundary[cp] ? -1 : c.p2io[cp]; };
    for (int i = 0; i < nf; ++i) {
        const int p = f.inner[i];
        const int cp = coarse_of(p);
        if (cp >= 0) {
            p0[i] = cin(cp);
            p1[i] = -2;
            continue;
        }
        int par[2], np = 0;
        for (int e = f.adj_ptr[p]; e < f.adj_ptr[p + 1]; ++e)
            for (int a = 0; a < 3; ++a) {
                const int q = f.tri[3 * f.adj[e] + a];
                if (q == p) continue;
                const int cq = coarse_of(q);
                if (cq < 0) continue;
                bool seen = false;
                for (int t = 0; t < np; ++t) seen = seen || par[t] == cq;
                if (!seen) {
                    if (np == 2) {
                        set_error("mg: a fine vertex has more than two coarse neighbours");
                        return PHG_ERR_ARG;
                    }
                    par[np++] = cq;
                }
            }
        if (np != 2) {
            set_error("mg: meshes are not nested (midpoint without two coarse neighbours)");
            return PHG_ERR_ARG;
        }
        p0[i] = cin(par[0]);
        p1[i] = cin(par[1]);
    }
    // restriction = transpose of the prolongation
    std::vector<std::vector<std::pair<int, double>>> rows(nc);
    for (int i = 0; i < nf; ++i) {
        if (p1[i] == -2) {
            if (p0[i] >= 0) rows[p0[i]].push_back({i, 1.0});
        } else {
            if (p0[i] >= 0) rows[p0[i]].push_back({i, 0.5});
            if (p1[i] >= 0) rows[p1[i]].push_back({i, 0.5});
        }
    }
    std::vector<int> rrp(nc + 1, 0), ridx;
    std::vector<double> rw;
    for (int j = 0; j < nc; ++j) {
        std::sort(rows[j].begin(), rows[j].end());
        for (auto& e : rows[j]) {
            ridx.push_back(e.first);
            rw.push_back(e.second);
        }
        rrp[j + 1] = (int)ridx.size();
    }
    PHG_HIP(fl.p0.upload(p0.data(), nf));
    PHG_HIP(fl.p1.upload(p1.data(), nf));
    PHG_HIP(fl.rrp.upload(rrp.data(), nc + 1));
    PHG_HIP(fl.ridx.upload(ridx.data(), ridx.size()));
    PHG_HIP(fl.rw.upload(rw.data(), rw.size()));
    return PHG_OK;
}

}  // namespace

namespace phg {

BatchPrecond* mg_precond(phg_mg_s* M)
{
    M->pc.M = M;
    return &M->pc;
}

// Work vectors for kp columns (re-allocated when kp grows).
int mg_reserve(phg_mg_s* M, int kp)
{
    if (M->kp >= kp) return PHG_OK;
    for (auto& l : M->lv) {
        const size_t sz = (size_t)l->n * kp;
        PHG_HIP(l->x0.alloc(sz));
        PHG_HIP(l->x1.alloc(sz));
        PHG_HIP(l->b.alloc(sz));
        PHG_HIP(l->r.alloc(sz));
    }
    M->kp = kp;
    return PHG_OK;
}

// Level buffers of one precision.
template <typename T>
struct MgT;
template <>
struct MgT<double> {
    using MT = double;
    static const Ell<double>& E(MgLevel& l) { return l.A->ell64; }
    static double* dinv(MgLevel& l) { return l.dinv.p; }
    static double* rw(MgLevel& l) { return l.rw.p; }
    static double* x0(MgLevel& l) { return l.x0.p; }
    static double* x1(MgLevel& l) { return l.x1.p; }
    static double* b(MgLevel& l) { return l.b.p; }
    static double* r(MgLevel& l) { return l.r.p; }
    static double* cinv(phg_mg_s* M) { return M->cinv.p; }
};
template <>
struct MgT<float> {
    using MT = float;
    static const Ell<float>& E(MgLevel& l) { return l.A->ell32; }
    static float* dinv(MgLevel& l) { return l.dinvf.p; }
    static float* rw(MgLevel& l) { return l.rwf.p; }
    static float* x0(MgLevel& l) { return l.x0f.p; }
    static float* x1(MgLevel& l) { return l.x1f.p; }
    static float* b(MgLevel& l) { return l.bf.p; }
    static float* r(MgLevel& l) { return l.rf.p; }
    static float* cinv(phg_mg_s* M) { return M->cinvf.p; }
};

// z = V-cycle(v): v, z are n x KP (finest level); enqueued on s.
template <typename T>
void mg_apply_t(phg_mg_s* M, int KP, const T* v, T* z, hipStream_t s)
{
    using B = MgT<T>;
    using MT = typename B::MT;
    const T om = (T)MG_OMEGA;
    const int nlev = (int)M->lv.size();
    // descend: level 0 takes b = v (read directly), result to z
    const T* bk = v;
    for (int k = 0; k < nlev - 1; ++k) {
        MgLevel& l = *M->lv[k];
        const auto& E = B::E(l);
        const size_t w = (size_t)l.n * KP;
        T* xa = (k == 0) ? z : B::x0(l);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_jacobi<T, MT>), grid_for(w), dim3(256), 0, s, l.n,
                           KP, E.width, E.col.p, E.val.p, B::dinv(l), (const T*)nullptr, bk,
                           B::x1(l), om);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_jacobi<T, MT>), grid_for(w), dim3(256), 0, s, l.n,
                           KP, E.width, E.col.p, E.val.p, B::dinv(l), B::x1(l), bk, xa, om);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_residual<T, MT>), grid_for(w), dim3(256), 0, s, l.n,
                           KP, E.width, E.col.p, E.val.p, xa, bk, B::r(l));
        MgLevel& c = *M->lv[k + 1];
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_restrict<T>), grid_for((size_t)c.n * KP), dim3(256),
                           0, s, c.n, KP, l.rrp.p, l.ridx.p, B::rw(l), B::r(l), B::b(c));
        bk = B::b(c);
    }
    MgLevel& cl = *M->lv[nlev - 1];
    T* xcoarse = (nlev == 1) ? z : B::x0(cl);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_dense<T>), dim3((unsigned)M->nc), dim3(256), 0, s, M->nc,
                       KP, B::cinv(M), bk, xcoarse);
    // ascend
    for (int k = nlev - 2; k >= 0; --k) {
        MgLevel& l = *M->lv[k];
        MgLevel& c = *M->lv[k + 1];
        const auto& E = B::E(l);
        const size_t w = (size_t)l.n * KP;
        const T* bl = (k == 0) ? v : B::b(l);
        T* xa = (k == 0) ? z : B::x0(l);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_prolong_add<T>), grid_for(w), dim3(256), 0, s, l.n, KP,
                           l.p0.p, l.p1.p, B::x0(c), xa);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_jacobi<T, MT>), grid_for(w), dim3(256), 0, s, l.n,
                           KP, E.width, E.col.p, E.val.p, B::dinv(l), xa, bl, B::x1(l), om);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(mg_jacobi<T, MT>), grid_for(w), dim3(256), 0, s, l.n,
                           KP, E.width, E.col.p, E.val.p, B::dinv(l), B::x1(l), bl, xa, om);
    }
}

void mg_apply(phg_mg_s* M, int KP, const double* v, double* z, hipStream_t s)
{
    mg_apply_t<double>(M, KP, v, z, s);
}

void mg_apply_f32(phg_mg_s* M, int KP, const float* v, float* z, hipStream_t s)
{
    mg_apply_t<float>(M, KP, v, z, s);
}

// FP32 copies of every level's operator values, inverse diagonal and
// restriction weights, the coarse inverse, and FP32 work vectors for kp
// columns.
int mg_reserve_f32(phg_mg_s* M, int kp)
{
    if (!M->cinvf.p) {
        for (auto& l : M->lv) {
            int rc = ensure_ell32(l->A, nullptr);
            if (rc) return rc;
            PHG_HIP(l->dinvf.alloc(l->n));
            if (l->n)
                hipLaunchKernelGGL(mg_to_f32, grid_for(l->n), dim3(256), 0, nullptr,
                                   (size_t)l->n, l->dinv.p, l->dinvf.p);
            if (l->rw.n) {
                PHG_HIP(l->rwf.alloc(l->rw.n));
                hipLaunchKernelGGL(mg_to_f32, grid_for(l->rw.n), dim3(256), 0, nullptr, l->rw.n,
                                   l->rw.p, l->rwf.p);
            }
        }
        const size_t nci = (size_t)M->nc * M->nc;
        PHG_HIP(M->cinvf.alloc(nci));
        hipLaunchKernelGGL(mg_to_f32, grid_for(nci), dim3(256), 0, nullptr, nci, M->cinv.p,
                           M->cinvf.p);
        PHG_HIP(hipGetLastError());
        PHG_HIP(hipDeviceSynchronize());
    }
    if (M->kpf >= kp) return PHG_OK;
    for (auto& l : M->lv) {
        const size_t sz = (size_t)l->n * kp;
        PHG_HIP(l->x0f.alloc(sz));
        PHG_HIP(l->x1f.alloc(sz));
        PHG_HIP(l->bf.alloc(sz));
        PHG_HIP(l->rf.alloc(sz));
    }
    M->kpf = kp;
    return PHG_OK;
}

int mg_rows(const phg_mg_s* M) { return M->lv.empty() ? 0 : M->lv[0]->n; }

}  // namespace phg

extern "C" {

int phg_mg_create(phg_mesh_t fine, int kind, double c_mass, double c_lap, phg_stream_t stream,
                  phg_mg_t* out)
{
    if (!fine || !out || kind < 0 || kind > PHG_HELMHOLTZ) {
        set_error("phg_mg_create: bad argument (kind must be MASS, LAPLACE or HELMHOLTZ)");
        return PHG_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    const int L = level_of(fine->h);
    if (L < 0 || fine->h.outer.size() != 1) {
        set_error("phg_mg_create: the mesh is not a full icosahedral sphere (mke_mesh_sphere)");
        return PHG_ERR_ARG;
    }
    std::unique_ptr<phg_mg_s> M(new phg_mg_s());
    M->kind = kind;
    M->L = L;
    M->c_mass = c_mass;
    M->c_lap = c_lap;
    const int coarse = std::min(L, MG_COARSE);
    for (int k = L; k >= coarse; --k) {
        std::unique_ptr<MgLevel> l(new MgLevel());
        if (k == L) {
            l->mesh = fine;
        } else {
            l->mesh = new phg_mesh_s();
            l->own_mesh = true;
            std::string err;
            const std::string text = sphere_mesh_text(k);
            if (!load_mesh_text(text.data(), text.size(), l->mesh->h, err)) {
                set_error("phg_mg_create: " + err);
                return PHG_ERR_FORMAT;
            }
        }
        int rc = assemble_device(l->mesh, kind, c_mass, c_lap, s, &l->A);
        if (rc) return rc;
        rc = ensure_ell64(l->A, s);
        if (rc) return rc;
        if (l->A->ell64.width > ELL_MAXW) {
            set_error("phg_mg_create: rows longer than 8 entries");
            return PHG_ERR_ARG;
        }
        l->n = l->A->n_rows;
        PHG_HIP(l->dinv.alloc(l->n));
        hipLaunchKernelGGL(mg_diag_inv, grid_for(l->n), dim3(256), 0, s, l->n, l->A->rp.p,
                           l->A->ci.p, l->A->val.p, l->dinv.p);
        PHG_HIP(hipGetLastError());
        M->lv.push_back(std::move(l));
    }
    for (size_t k = 0; k + 1 < M->lv.size(); ++k) {
        int rc = build_transfer(M->lv[k]->mesh->h, M->lv[k + 1]->mesh->h, *M->lv[k],
                                M->lv[k + 1]->n);
        if (rc) return rc;
    }
    // inverse of the coarsest operator (host, dense)
    MgLevel& cl = *M->lv.back();
    const int nc = cl.n;
    std::vector<double> vals(cl.A->nnz);
    PHG_HIP(hipMemcpy(vals.data(), cl.A->val.p, vals.size() * sizeof(double),
                      hipMemcpyDeviceToHost));
    std::vector<double> a((size_t)nc * nc, 0.0), inv;
    for (int i = 0; i < nc; ++i)
        for (int t = cl.A->h_rp[i]; t < cl.A->h_rp[i + 1]; ++t)
            a[(size_t)i * nc + cl.A->h_ci[t]] += vals[t];
    if (!dense_inverse(nc, a, inv)) {
        set_error("phg_mg_create: coarsest operator is singular");
        return PHG_ERR_ARG;
    }
    PHG_HIP(M->cinv.upload(inv.data(), inv.size(), s));
    M->nc = nc;
    int rc = mg_reserve(M.get(), 1);
    if (rc) return rc;
    PHG_HIP(hipStreamSynchronize(s));
    *out = M.release();
    return PHG_OK;
}

void phg_mg_destroy(phg_mg_t M) { delete M; }

int phg_mg_matrix(phg_mg_t M, phg_matrix_t* out)
{
    if (!M || !out) {
        set_error("phg_mg_matrix: null argument");
        return PHG_ERR_ARG;
    }
    *out = M->lv[0]->A;
    return PHG_OK;
}

int phg_mg_levels(phg_mg_t M) { return M ? (int)M->lv.size() : 0; }

}  // extern "C"
