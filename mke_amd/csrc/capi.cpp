// C ABI entry points (include/phelm_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "phg_internal.h"

namespace phg {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

int spmv_f64(phg_matrix_s* A, const double* x, double* y, hipStream_t s);
int spmv_f32(phg_matrix_s* A, int prec, const float* x, float* y, hipStream_t s);
int spmm_f32(phg_matrix_s* A, int prec, int nrhs, const float* X, float* Y, hipStream_t s);
int dot_f64(int n, const double* x, const double* y, double* result, hipStream_t s);
int dot_noalloc_f64(int n, const double* x, const double* y, double* result, hipStream_t s);
int noalloc_scratch(double** out);
int axpby_f64(int n, double k1, const double* a, double k2, const double* b, double* r,
              hipStream_t s);
int u2p_f64(phg_mesh_s* m, const double* u, double* p, hipStream_t s);
int p2u_f64(phg_mesh_s* m, const double* p, const double* bnd, double* u, hipStream_t s);
int proj_bnd_f64(phg_mesh_s* m, const double* u, double* b, hipStream_t s);
int set_bnd_f64(phg_mesh_s* m, const double* bnd, double* u, hipStream_t s);
int gmres_solve(phg_matrix_s* A, const double* b, double* x, const phg_gmres_opts* opts,
                phg_gmres_info* info, hipStream_t s);
void csr_structure(const SphereMesh& m, std::vector<int>& rp, std::vector<int>& ci);
int block_gmres(phg_matrix_s* A, int nrhs, const double* B, double* X, const phg_gmres_opts* opts,
                phg_block_info* info, hipStream_t s);

int ensure_mesh_device(phg_mesh_s* m)
{
    if (m->d_coords.p || m->h.n_points == 0) return PHG_OK;
    PHG_HIP(m->d_inner.upload(m->h.inner.data(), m->h.inner.size()));
    PHG_HIP(m->d_outer.upload(m->h.outer.data(), m->h.outer.size()));
    PHG_HIP(m->d_p2io.upload(m->h.p2io.data(), m->h.p2io.size()));
    PHG_HIP(m->d_isb.upload(m->h.is_boundary.data(), m->h.is_boundary.size()));
    PHG_HIP(m->d_tri.upload(m->h.tri.data(), m->h.tri.size()));
    PHG_HIP(m->d_coords.upload(m->h.coords.data(), m->h.coords.size()));
    return PHG_OK;
}

int matrix_from_csr_host(int n_rows, int n_cols, const int* rp, const int* ci,
                         const double* vals, phg_matrix_s** out)
{
    std::unique_ptr<phg_matrix_s> A(new phg_matrix_s());
    A->n_rows = n_rows;
    A->n_cols = n_cols;
    A->nnz = rp[n_rows];
    A->h_rp.assign(rp, rp + n_rows + 1);
    A->h_ci.assign(ci, ci + A->nnz);
    PHG_HIP(A->rp.upload(rp, n_rows + 1));
    PHG_HIP(A->ci.upload(ci, A->nnz));
    PHG_HIP(A->val.upload(vals, A->nnz));
    *out = A.release();
    return PHG_OK;
}

}  // namespace phg

phg_matrix_s::~phg_matrix_s()
{
    for (auto& w : work) w.reset();
    phg::block_destroy(bwork);
    for (auto* b : batch) phg::batch_destroy(b);
}

using namespace phg;

extern "C" {

const char* phg_version(void) { return "phelm-mi355x 0.1 (gfx950)"; }

const char* phg_last_error(void) { return g_error.c_str(); }

int phg_init(int device)
{
    PHG_HIP(hipSetDevice(device));
    // the device's phg_dot_noalloc scratch address, looked up here once
    double* scratch = nullptr;
    return noalloc_scratch(&scratch);
}

void phg_free(void* p) { free(p); }

int phg_mesh_sphere_text(int level, char** text, size_t* len)
{
    if (level < 0 || level > 10 || !text || !len) {
        set_error("phg_mesh_sphere_text: level must be in 0..10");
        return PHG_ERR_ARG;
    }
    std::string s = sphere_mesh_text(level);
    char* buf = (char*)malloc(s.size() + 1);
    if (!buf) return PHG_ERR_NOMEM;
    memcpy(buf, s.data(), s.size());
    buf[s.size()] = 0;
    *text = buf;
    *len = s.size();
    return PHG_OK;
}

int phg_mesh_load(const char* text, size_t len, phg_mesh_t* out)
{
    if (!text || !out) {
        set_error("phg_mesh_load: null argument");
        return PHG_ERR_ARG;
    }
    std::unique_ptr<phg_mesh_s> m(new phg_mesh_s());
    std::string err;
    if (!load_mesh_text(text, len, m->h, err)) {
        set_error(err);
        return PHG_ERR_FORMAT;
    }
    *out = m.release();
    return PHG_OK;
}

int phg_mesh_sphere(int level, phg_mesh_t* out)
{
    if (level < 0 || level > 10 || !out) {
        set_error("phg_mesh_sphere: level must be in 0..10");
        return PHG_ERR_ARG;
    }
    std::string s = sphere_mesh_text(level);
    return phg_mesh_load(s.data(), s.size(), out);
}

int phg_mesh_save(phg_mesh_t m, const char* path)
{
    if (!m || !path) {
        set_error("phg_mesh_save: null argument");
        return PHG_ERR_ARG;
    }
    std::string err;
    if (!save_mesh_binary(m->h, path, err)) {
        set_error(err);
        return PHG_ERR_ARG;
    }
    return PHG_OK;
}

int phg_mesh_load_file(const char* path, phg_mesh_t* out)
{
    if (!path || !out) {
        set_error("phg_mesh_load_file: null argument");
        return PHG_ERR_ARG;
    }
    std::unique_ptr<phg_mesh_s> m(new phg_mesh_s());
    std::string err;
    if (!load_mesh_binary(path, m->h, err)) {
        set_error(err);
        return PHG_ERR_FORMAT;
    }
    *out = m.release();
    return PHG_OK;
}

int phg_mesh_sphere_cached(int level, const char* dir, phg_mesh_t* out)
{
    if (level < 0 || level > 10 || !dir || !out) {
        set_error("phg_mesh_sphere_cached: bad argument (level 0..10, dir, out)");
        return PHG_ERR_ARG;
    }
    const std::string path = std::string(dir) + "/sphere_L" + std::to_string(level) + ".phgmesh";
    if (phg_mesh_load_file(path.c_str(), out) == PHG_OK) return PHG_OK;
    int rc = phg_mesh_sphere(level, out);
    if (rc) return rc;
    std::string err;
    (void)save_mesh_binary((*out)->h, path.c_str(), err);  // a cache: failure to write is not an error
    return PHG_OK;
}

int phg_print_function(phg_mesh_t m, const double* h_u, int sphere_xyz, char** text, size_t* len)
{
    if (!m || !h_u || !text || !len) {
        set_error("phg_print_function: null argument");
        return PHG_ERR_ARG;
    }
    const std::string s = print_function_text(m->h, h_u, sphere_xyz != 0);
    char* buf = (char*)std::malloc(s.size() + 1);
    if (!buf) {
        set_error("phg_print_function: out of memory");
        return PHG_ERR_NOMEM;
    }
    std::memcpy(buf, s.data(), s.size());
    buf[s.size()] = 0;
    *text = buf;
    *len = s.size();
    return PHG_OK;
}

int phg_mesh_info(phg_mesh_t m, int* n_points, int* n_triangles, int* n_inner, int* n_outer,
                  int* n_zones)
{
    if (!m) return PHG_ERR_ARG;
    if (n_points) *n_points = m->h.n_points;
    if (n_triangles) *n_triangles = m->h.n_triangles;
    if (n_inner) *n_inner = (int)m->h.inner.size();
    if (n_outer) *n_outer = (int)m->h.outer.size();
    if (n_zones) *n_zones = m->h.n_zones;
    return PHG_OK;
}

int phg_mesh_arrays(phg_mesh_t m, double* pr, double* coords, int* tri, int* tri_zone,
                    int* adj_ptr, int* adj, int* is_boundary, int* inner, int* outer, int* p2io)
{
    if (!m) return PHG_ERR_ARG;
    const SphereMesh& h = m->h;
    auto cp = [](auto* dst, const auto& v) {
        if (dst && !v.empty()) memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(pr, h.pr);
    cp(coords, h.coords);
    cp(tri, h.tri);
    cp(tri_zone, h.tri_zone);
    cp(adj_ptr, h.adj_ptr);
    cp(adj, h.adj);
    cp(is_boundary, h.is_boundary);
    cp(inner, h.inner);
    cp(outer, h.outer);
    cp(p2io, h.p2io);
    return PHG_OK;
}

void phg_mesh_destroy(phg_mesh_t m) { delete m; }

int phg_mesh_csr_structure(phg_mesh_t m, int* row_ptr, int* col_idx, long* nnz)
{
    if (!m || !row_ptr) return PHG_ERR_ARG;
    std::vector<int> rp, ci;
    csr_structure(m->h, rp, ci);
    memcpy(row_ptr, rp.data(), rp.size() * sizeof(int));
    if (col_idx && !ci.empty()) memcpy(col_idx, ci.data(), ci.size() * sizeof(int));
    if (nnz) *nnz = (long)ci.size();
    return PHG_OK;
}

int phg_matrix_from_triplets(int n_rows, int n_cols, long n, const int* rows, const int* cols,
                             const double* vals, phg_matrix_t* out)
{
    if (n_rows < 0 || n_cols < 0 || n < 0 || !out || (n && (!rows || !cols || !vals))) {
        set_error("phg_matrix_from_triplets: bad argument");
        return PHG_ERR_ARG;
    }
    for (long e = 0; e < n; ++e)
        if (rows[e] < 0 || rows[e] >= n_rows || cols[e] < 0 || cols[e] >= n_cols) {
            set_error("phg_matrix_from_triplets: index out of range");
            return PHG_ERR_ARG;
        }
    // stable sort by (row, col): duplicates keep their call order and are
    // summed from 0.0 in that order, like repeated Solver::add
    std::vector<long> idx(n);
    std::iota(idx.begin(), idx.end(), 0L);
    std::stable_sort(idx.begin(), idx.end(), [&](long a, long b) {
        return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
    });
    std::vector<int> rp(n_rows + 1, 0), ci;
    std::vector<double> v;
    for (long t = 0; t < n; ++t) {
        const long e = idx[t];
        if (!ci.empty() && t > 0 && rows[idx[t - 1]] == rows[e] && ci.back() == cols[e]) {
            v.back() += vals[e];
        } else {
            ci.push_back(cols[e]);
            v.push_back(0.0 + vals[e]);
            rp[rows[e] + 1]++;
        }
    }
    for (int i = 0; i < n_rows; ++i) rp[i + 1] += rp[i];
    phg_matrix_s* A = nullptr;
    int rc = matrix_from_csr_host(n_rows, n_cols, rp.data(), ci.data(), v.data(), &A);
    if (rc) return rc;
    *out = A;
    return PHG_OK;
}

int phg_matrix_from_csr(int n_rows, int n_cols, const int* row_ptr, const int* col_idx,
                        const double* vals, phg_matrix_t* out)
{
    if (n_rows < 0 || n_cols < 0 || !row_ptr || !out) {
        set_error("phg_matrix_from_csr: bad argument");
        return PHG_ERR_ARG;
    }
    if (row_ptr[0] != 0) {
        set_error("phg_matrix_from_csr: row_ptr[0] must be 0");
        return PHG_ERR_ARG;
    }
    if (row_ptr[n_rows] > 0 && (!col_idx || !vals)) {
        set_error("phg_matrix_from_csr: entries given without col_idx or vals");
        return PHG_ERR_ARG;
    }
    for (int i = 0; i < n_rows; ++i) {
        if (row_ptr[i + 1] < row_ptr[i]) {
            set_error("phg_matrix_from_csr: row_ptr not monotone");
            return PHG_ERR_ARG;
        }
        for (int e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            if (col_idx[e] < 0 || col_idx[e] >= n_cols ||
                (e > row_ptr[i] && col_idx[e] <= col_idx[e - 1])) {
                set_error("phg_matrix_from_csr: columns out of range or not ascending");
                return PHG_ERR_ARG;
            }
    }
    phg_matrix_s* A = nullptr;
    int rc = matrix_from_csr_host(n_rows, n_cols, row_ptr, col_idx, vals, &A);
    if (rc) return rc;
    *out = A;
    return PHG_OK;
}

int phg_assemble(phg_mesh_t m, int kind, double c_mass, double c_lap, phg_stream_t stream,
                 phg_matrix_t* out)
{
    if (!m || !out || kind < 0 || kind > PHG_DIFF2_COS) {
        set_error("phg_assemble: bad argument");
        return PHG_ERR_ARG;
    }
    return assemble_device(m, kind, c_mass, c_lap, (hipStream_t)stream, out);
}

int phg_assemble_boundary(phg_mesh_t m, int kind, double c_mass, double c_lap, phg_matrix_t* out)
{
    if (!m || !out || kind < 0 || kind > PHG_DIFF2_COS) {
        set_error("phg_assemble_boundary: bad argument");
        return PHG_ERR_ARG;
    }
    return assemble_boundary_device(m, kind, c_mass, c_lap, nullptr, out);
}

int phg_assemble_full(phg_mesh_t m, int kind, double c_mass, double c_lap, phg_stream_t stream,
                      phg_matrix_t* out)
{
    if (!m || !out || kind < 0 || kind > PHG_DIFF2_COS) {
        set_error("phg_assemble_full: bad argument");
        return PHG_ERR_ARG;
    }
    return assemble_device(m, kind, c_mass, c_lap, (hipStream_t)stream, out, true);
}

int phg_matrix_info(phg_matrix_t A, int* n_rows, int* n_cols, long* nnz)
{
    if (!A) return PHG_ERR_ARG;
    if (n_rows) *n_rows = A->n_rows;
    if (n_cols) *n_cols = A->n_cols;
    if (nnz) *nnz = A->nnz;
    return PHG_OK;
}

int phg_matrix_csr(phg_matrix_t A, int* row_ptr, int* col_idx, double* vals)
{
    if (!A) return PHG_ERR_ARG;
    if (row_ptr) memcpy(row_ptr, A->h_rp.data(), A->h_rp.size() * sizeof(int));
    if (col_idx && A->nnz) memcpy(col_idx, A->h_ci.data(), A->nnz * sizeof(int));
    if (vals && A->nnz) PHG_HIP(hipMemcpy(vals, A->val.p, A->nnz * sizeof(double), hipMemcpyDeviceToHost));
    return PHG_OK;
}

void phg_matrix_destroy(phg_matrix_t A) { delete A; }

int phg_spmv(phg_matrix_t A, const double* d_x, double* d_y, phg_stream_t s)
{
    if (!A || (A->n_rows && (!d_x || !d_y))) return PHG_ERR_ARG;
    return spmv_f64(A, d_x, d_y, (hipStream_t)s);
}

int phg_spmv_f32(phg_matrix_t A, int prec, const float* d_x, float* d_y, phg_stream_t s)
{
    if (!A || (prec != PHG_F32 && prec != PHG_BF16)) return PHG_ERR_ARG;
    return spmv_f32(A, prec, d_x, d_y, (hipStream_t)s);
}

int phg_spmm_f32(phg_matrix_t A, int prec, int nrhs, const float* d_X, float* d_Y,
                 phg_stream_t s)
{
    if (!A || nrhs < 1 || (prec != PHG_F32 && prec != PHG_BF16)) return PHG_ERR_ARG;
    return spmm_f32(A, prec, nrhs, d_X, d_Y, (hipStream_t)s);
}

int phg_dot(int n, const double* d_x, const double* d_y, double* result, phg_stream_t s)
{
    if (n < 0 || !result) return PHG_ERR_ARG;
    return dot_f64(n, d_x, d_y, result, (hipStream_t)s);
}

int phg_dot_noalloc(int n, const double* d_x, const double* d_y, double* d_result,
                    phg_stream_t s)
{
    return dot_noalloc_f64(n, d_x, d_y, d_result, (hipStream_t)s);
}

int phg_nrm2(int n, const double* d_x, double* result, phg_stream_t s)
{
    if (n < 0 || !result) return PHG_ERR_ARG;
    double d = 0.0;
    int rc = dot_f64(n, d_x, d_x, &d, (hipStream_t)s);
    *result = std::sqrt(d);
    return rc;
}

int phg_axpby(int n, double k1, const double* d_a, double k2, const double* d_b, double* d_r,
              phg_stream_t s)
{
    if (n < 0) return PHG_ERR_ARG;
    return axpby_f64(n, k1, d_a, k2, d_b, d_r, (hipStream_t)s);
}

int phg_u2p(phg_mesh_t m, const double* d_u, double* d_p, phg_stream_t s)
{
    if (!m) return PHG_ERR_ARG;
    return u2p_f64(m, d_u, d_p, (hipStream_t)s);
}

int phg_p2u(phg_mesh_t m, const double* d_p, const double* d_bnd, double* d_u, phg_stream_t s)
{
    if (!m) return PHG_ERR_ARG;
    return p2u_f64(m, d_p, d_bnd, d_u, (hipStream_t)s);
}

int phg_proj_bnd(phg_mesh_t m, const double* d_u, double* d_b, phg_stream_t s)
{
    if (!m || !d_u || !d_b) {
        set_error("phg_proj_bnd: null argument");
        return PHG_ERR_ARG;
    }
    return proj_bnd_f64(m, d_u, d_b, (hipStream_t)s);
}

int phg_set_bnd(phg_mesh_t m, const double* d_bnd, double* d_u, phg_stream_t s)
{
    if (!m || !d_u) {
        set_error("phg_set_bnd: null argument");
        return PHG_ERR_ARG;
    }
    return set_bnd_f64(m, d_bnd, d_u, (hipStream_t)s);
}

void phg_gmres_default_opts(phg_gmres_opts* o)
{
    o->restart = 30;
    o->max_iters = 200000;
    o->rtol = 1e-10;
    o->ortho = PHG_MGS;
    o->precision = PHG_F64;
    o->inner_rtol = 1e-5;
    o->use_x0 = 0;
    o->graph = 1;
    o->diag_scale = 0;
    o->persistent = 1;
}

int phg_gmres(phg_matrix_t A, const double* d_b, double* d_x, const phg_gmres_opts* opts,
              phg_gmres_info* info, phg_stream_t s)
{
    if (!A || !d_b || !d_x) {
        set_error("phg_gmres: null argument");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o;
    if (opts) o = *opts;
    else phg_gmres_default_opts(&o);
    if (o.precision < PHG_F64 || o.precision > PHG_BF16 || o.max_iters < 0 || !(o.rtol >= 0)) {
        set_error("phg_gmres: bad options");
        return PHG_ERR_ARG;
    }
    return gmres_solve(A, d_b, d_x, &o, info, (hipStream_t)s);
}

int phg_block_gmres(phg_matrix_t A, int nrhs, const double* d_B, double* d_X,
                    const phg_gmres_opts* opts, phg_block_info* info, phg_stream_t s)
{
    if (!A || !d_B || !d_X) {
        set_error("phg_block_gmres: null argument");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o;
    if (opts) o = *opts;
    else phg_gmres_default_opts(&o);
    if (o.max_iters < 0 || !(o.rtol >= 0)) {
        set_error("phg_block_gmres: bad options");
        return PHG_ERR_ARG;
    }
    return block_gmres(A, nrhs, d_B, d_X, &o, info, (hipStream_t)s);
}

int phg_gmres_batch_mg(phg_mg_t M, int nrhs, const double* d_B, double* d_X,
                       const phg_gmres_opts* opts, phg_gmres_info* infos, phg_stream_t s)
{
    if (!M || !d_B || !d_X) {
        set_error("phg_gmres_batch_mg: null argument");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o;
    if (opts) o = *opts;
    else {
        phg_gmres_default_opts(&o);
        o.ortho = PHG_CGS2;
    }
    if (o.max_iters < 0 || !(o.rtol >= 0)) {
        set_error("phg_gmres_batch_mg: bad options");
        return PHG_ERR_ARG;
    }
    phg_matrix_t A = nullptr;
    int rc = phg_mg_matrix(M, &A);
    if (rc) return rc;
    return gmres_batch_solve(A, nrhs, d_B, d_X, &o, infos, (hipStream_t)s, M);
}

int phg_gmres_batch(phg_matrix_t A, int nrhs, const double* d_B, double* d_X,
                    const phg_gmres_opts* opts, phg_gmres_info* infos, phg_stream_t s)
{
    if (!A || !d_B || !d_X) {
        set_error("phg_gmres_batch: null argument");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o;
    if (opts) o = *opts;
    else {
        phg_gmres_default_opts(&o);
        o.ortho = PHG_CGS2;
    }
    if (o.max_iters < 0 || !(o.rtol >= 0)) {
        set_error("phg_gmres_batch: bad options");
        return PHG_ERR_ARG;
    }
    return gmres_batch_solve(A, nrhs, d_B, d_X, &o, infos, (hipStream_t)s);
}

int phg_gmres_batch_tune(phg_matrix_t A, int nrhs, const phg_gmres_opts* opts, phg_stream_t s,
                         int* rows_per_thread, double* ms)
{
    if (!A) {
        set_error("phg_gmres_batch_tune: null argument");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o;
    if (opts) o = *opts;
    else {
        phg_gmres_default_opts(&o);
        o.ortho = PHG_CGS2;
    }
    return gmres_batch_tune(A, nrhs, &o, (hipStream_t)s, rows_per_thread, ms);
}

int phg_gmres_batch_shape(int n_rows, int nrhs, int precision)
{
    if (n_rows < 1 || nrhs < 1 || nrhs > 64 || precision < PHG_F64 || precision > PHG_BF16) {
        set_error("phg_gmres_batch_shape: bad argument");
        return PHG_ERR_ARG;
    }
    return batch_shape_get(n_rows, nrhs, precision);
}

int phg_gmres_batch_set_shape(int n_rows, int nrhs, int precision, int rows_per_thread)
{
    return batch_shape_set(n_rows, nrhs, precision, rows_per_thread);
}

}  // extern "C"
