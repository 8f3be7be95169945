// Internal objects behind the C ABI of include/phelm_gpu.h.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/phelm_gpu.h"
#include "mesh.h"

namespace phg {

void set_error(const std::string& msg);

// f(lo, hi) over [0, n) split into contiguous ranges on up to 16 host
// threads (host-side integer work of the assembly plans; each range writes
// only its own rows, so the result does not depend on the split).
template <typename F>
void parallel_ranges(int n, F f)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int T = (int)std::min(16u, std::max(1u, hw));
    if (n < 8192 || T == 1) {
        f(0, n);
        return;
    }
    const int chunk = (n + T - 1) / T;
    std::vector<std::thread> th;
    th.reserve(T);
    for (int t = 0; t < T; ++t) {
        const int lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) continue;
        try {
            th.emplace_back(f, lo, hi);
        } catch (...) {
            f(lo, hi);  // no thread to be had: this range on the caller's
        }
    }
    for (auto& x : th) x.join();
}

#define PHG_HIP(call)                                                          \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) {                                                \
            ::phg::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return PHG_ERR_HIP;                                                \
        }                                                                      \
    } while (0)

// Device buffer owned by an object; freed in the destructor.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    hipError_t upload(const T* h, size_t count, hipStream_t s = nullptr)
    {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        e = hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(s);
    }
};

// ELL copy of a CSR matrix in slot-major order: entry k of row i sits at
// k * n_rows + i, so a wavefront of consecutive rows reads consecutive words.
// Short rows are padded with value 0 and a column of the same row.
template <typename V>
struct Ell {
    int width = 0;
    DevBuf<int> col;
    DevBuf<V> val;
};

struct GmresWork;  // gmres.hip
struct BlockWork;  // block.hip
struct BatchWork;  // gmres_batch.hip
void destroy_gmres_work(GmresWork* w);
void block_destroy(BlockWork* w);
void batch_destroy(BatchWork* w);
struct GmresWorkDeleter {
    void operator()(GmresWork* w) const { destroy_gmres_work(w); }
};

}  // namespace phg

struct phg_mesh_s {
    phg::SphereMesh h;
    // device copies for gather/scatter and assembly (uploaded on first use)
    phg::DevBuf<int> d_inner, d_outer, d_p2io, d_isb;
    phg::DevBuf<double> d_coords;
    phg::DevBuf<int> d_tri;
};

struct phg_matrix_s {
    int n_rows = 0, n_cols = 0;
    long nnz = 0;
    std::vector<int> h_rp, h_ci;   // host CSR structure (always kept)
    phg::DevBuf<int> rp, ci;
    phg::DevBuf<double> val;       // device CSR values (FP64)
    phg::Ell<double> ell64;
    phg::Ell<float> ell32;
    phg::Ell<uint16_t> ell16;      // bfloat16 bit patterns
    phg::Ell<double> ell64d;       // columns scaled by dinv (diagonal right preconditioning)
    phg::DevBuf<double> dinv;      // 1 / diagonal
    // per precision (PHG_F64, PHG_F32, PHG_BF16) and orthogonalisation (MGS, CGS2)
    std::unique_ptr<phg::GmresWork, phg::GmresWorkDeleter> work[6];
    phg::BlockWork* bwork = nullptr;  // block GMRES workspace
    phg::BatchWork* batch[2] = {nullptr, nullptr};  // batched GMRES: FP64, mixed
    ~phg_matrix_s();
};

namespace phg {
// assembly (assemble.hip): inner x inner operator, or every point x every
// point (generate_full_matrix) when full; the inner x outer boundary matrix
int assemble_device(phg_mesh_s* M, int kind, double c_mass, double c_lap, hipStream_t stream,
                    phg_matrix_s** out, bool full = false);
int assemble_boundary_device(phg_mesh_s* M, int kind, double c_mass, double c_lap,
                             hipStream_t stream, phg_matrix_s** out);
// multigrid preconditioner (mg.hip)
int mg_reserve(phg_mg_s* M, int kp);
void mg_apply(phg_mg_s* M, int KP, const double* v, double* z, hipStream_t s);
int mg_reserve_f32(phg_mg_s* M, int kp);
void mg_apply_f32(phg_mg_s* M, int KP, const float* v, float* z, hipStream_t s);
int mg_rows(const phg_mg_s* M);
// Right preconditioner of the batched GMRES: z = P^-1 v for n x KP FP64
// vectors, enqueued on s (captured into the cycle graph, so a fixed kernel
// sequence); the object's address identifies it in the graph key.
// GMRES restart length of the batched mass solves (the Jacobian's eight
// systems per member and the L() solves): the diagonally scaled mass matrix
// converges in about 17 steps either way, and a short cycle keeps the
// orthogonalisation reads small (measured on the 8-member level-8 BarVortex
// step: restart 32 -> 8 cuts the step from 1081 to 775 ms; 136 -> 140 steps).
constexpr int MASS_RESTART = 8;

struct BatchPrecond {
    virtual ~BatchPrecond() {}
    virtual int rows() const = 0;
    virtual int reserve(int kp) = 0;
    virtual void apply(int KP, const double* v, double* z, hipStream_t s) = 0;
    // FP32 form for the mixed-precision solves (optional)
    virtual bool f32() const { return false; }
    virtual int reserve_f32(int) { return PHG_ERR_ARG; }
    virtual void apply_f32(int, const float*, float*, hipStream_t) {}
};
BatchPrecond* mg_precond(phg_mg_s* M);
int gmres_batch_solve(phg_matrix_s* A, int K, const double* B, double* X,
                      const phg_gmres_opts* opts, phg_gmres_info* infos, hipStream_t s,
                      phg_mg_s* pc = nullptr);
int gmres_batch_solve_pc(phg_matrix_s* A, int K, const double* B, double* X,
                         const phg_gmres_opts* opts, phg_gmres_info* infos, hipStream_t s,
                         BatchPrecond* pc);
int gmres_batch_tune(phg_matrix_s* A, int K, const phg_gmres_opts* opts, hipStream_t s,
                     int* best, double* ms);
int batch_shape_get(int n, int K, int prec);
int batch_shape_set(int n, int K, int prec, int rpt);
int ensure_mesh_device(phg_mesh_s* m);
int ensure_ell64(phg_matrix_s* A, hipStream_t s);
// A D^-1 (D = diag A) as an ELL copy, and D^-1 (square matrices with a non-zero diagonal)
int ensure_ell64_diag(phg_matrix_s* A, hipStream_t s);
int ensure_ell32(phg_matrix_s* A, hipStream_t s);
int ensure_ell16(phg_matrix_s* A, hipStream_t s);
// matrix with host structure given, device values uploaded from host
int matrix_from_csr_host(int n_rows, int n_cols, const int* rp, const int* ci,
                         const double* vals, phg_matrix_s** out);
}  // namespace phg
