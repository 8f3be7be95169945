// GPU assembly of the sphere operators into CSR (generate_matrix,
// src/generators.cpp:57-139) and of the inner x outer boundary matrix
// (generate_boundary_matrix, src/generators.cpp:141-212).  Integer structure
// (row_ptr, col_idx, entry positions) is built on the host; all element
// values are computed on the device.
//
// The reference assembles row i in the chart of the first triangle around
// vertex inner[i] (generators.cpp:76) and adds the per-triangle values of row
// i in the order of that vertex's triangle list (mesh.cpp:312-320).  On the
// GPU this is done atomic-free in two passes:
//   1. one thread per (row, adjacent-triangle slot) evaluates that
//      triangle's three entries of the row in the row's chart
//      (elem_contrib): 30 cos evaluations and 24 GK15 sums shared by the
//      nine (phi_i, phi_j) products of the triangle;
//   2. one thread per row adds the slot values into the row's CSR entries in
//      slot order (row_reduce), i.e. exactly the reference's accumulation
//      order, so the values are bit-identical to a CPU restatement built
//      with the same element arithmetic (fem_math.h).
// A triangle colouring would also be atomic-free but would change the
// summation order of each entry.
#include <algorithm>
#include <cstring>

#include "fem_math.h"
#include "phg_internal.h"

namespace phg {
namespace {

struct AsmPlan {
    // host
    std::vector<int> rp, ci;          // CSR structure
    std::vector<int> slot_row, slot_tri, slot_local, row_zone, slot_ptr;
    std::vector<int> pos;             // [S*3] CSR index or -1 (boundary column)
};

}  // namespace

// Row i holds p2io of every inner vertex of every triangle around inner[i],
// columns ascending: the pattern Solver::add receives from generate_matrixv
// (generators.cpp:71-110).
void csr_structure(const SphereMesh& m, std::vector<int>& rp, std::vector<int>& ci)
{
    const int rs = (int)m.inner.size();
    // the sorted distinct inner columns of row i (into tmp)
    auto row_cols = [&m](int i, std::vector<int>& tmp) {
        const int p = m.inner[i];
        tmp.clear();
        for (int e = m.adj_ptr[p]; e < m.adj_ptr[p + 1]; ++e) {
            const int t = m.adj[e];
            for (int a = 0; a < 3; ++a) {
                const int q = m.tri[3 * t + a];
                if (!m.is_boundary[q]) tmp.push_back(m.p2io[q]);
            }
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    };
    // counts, prefix sum, then the columns, rows in parallel
    rp.assign(rs + 1, 0);
    parallel_ranges(rs, [&](int lo, int hi) {
        std::vector<int> tmp;
        for (int i = lo; i < hi; ++i) {
            row_cols(i, tmp);
            rp[i + 1] = (int)tmp.size();
        }
    });
    for (int i = 0; i < rs; ++i) rp[i + 1] += rp[i];
    ci.assign((size_t)rp[rs], 0);
    parallel_ranges(rs, [&](int lo, int hi) {
        std::vector<int> tmp;
        for (int i = lo; i < hi; ++i) {
            row_cols(i, tmp);
            std::copy(tmp.begin(), tmp.end(), ci.begin() + rp[i]);
        }
    });
}

namespace {

void build_plan(const SphereMesh& m, AsmPlan& P)
{
    const int rs = (int)m.inner.size();
    csr_structure(m, P.rp, P.ci);
    P.slot_ptr.assign(rs + 1, 0);
    P.row_zone.resize(rs);
    for (int i = 0; i < rs; ++i) {
        const int p = m.inner[i];
        P.slot_ptr[i + 1] = P.slot_ptr[i] + (m.adj_ptr[p + 1] - m.adj_ptr[p]);
        P.row_zone[i] = m.tri_zone[m.adj[m.adj_ptr[p]]];
    }
    const int S = P.slot_ptr[rs];
    P.slot_row.resize(S);
    P.slot_tri.resize(S);
    P.slot_local.resize(S);
    P.pos.resize((size_t)S * 3);
    parallel_ranges(rs, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            const int p = m.inner[i];
            int s = P.slot_ptr[i];
            for (int e = m.adj_ptr[p]; e < m.adj_ptr[p + 1]; ++e, ++s) {
                const int t = m.adj[e];
                P.slot_row[s] = i;
                P.slot_tri[s] = t;
                int loc = 0;
                while (m.tri[3 * t + loc] != p) ++loc;  // Triangle::elem1(p1, zone)
                P.slot_local[s] = loc;
                for (int a = 0; a < 3; ++a) {
                    const int q = m.tri[3 * t + a];
                    int where = -1;
                    if (!m.is_boundary[q]) {
                        const int j = m.p2io[q];
                        for (int c = P.rp[i]; c < P.rp[i + 1]; ++c)
                            if (P.ci[c] == j) where = c;
                    }
                    P.pos[(size_t)s * 3 + a] = where;
                }
            }
        }
    });
}

// generate_full_matrix (src/generators.h:200-234): every point is a row, in
// the zone of its first triangle; columns are point ids; nothing is skipped.
void build_plan_full(const SphereMesh& m, AsmPlan& P)
{
    const int V = m.n_points;
    P.rp.assign(V + 1, 0);
    P.ci.clear();
    P.slot_ptr.assign(V + 1, 0);
    P.row_zone.resize(V);
    std::vector<int> tmp;
    for (int p = 0; p < V; ++p) {
        tmp.clear();
        for (int e = m.adj_ptr[p]; e < m.adj_ptr[p + 1]; ++e)
            for (int a = 0; a < 3; ++a) tmp.push_back(m.tri[3 * m.adj[e] + a]);
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        P.ci.insert(P.ci.end(), tmp.begin(), tmp.end());
        P.rp[p + 1] = (int)P.ci.size();
        P.slot_ptr[p + 1] = P.slot_ptr[p] + (m.adj_ptr[p + 1] - m.adj_ptr[p]);
        P.row_zone[p] = m.tri_zone[m.adj[m.adj_ptr[p]]];
    }
    const int S = P.slot_ptr[V];
    P.slot_row.resize(S);
    P.slot_tri.resize(S);
    P.slot_local.resize(S);
    P.pos.resize((size_t)S * 3);
    for (int p = 0; p < V; ++p) {
        int s = P.slot_ptr[p];
        for (int e = m.adj_ptr[p]; e < m.adj_ptr[p + 1]; ++e, ++s) {
            const int t = m.adj[e];
            P.slot_row[s] = p;
            P.slot_tri[s] = t;
            int loc = 0;
            while (m.tri[3 * t + loc] != p) ++loc;
            P.slot_local[s] = loc;
            for (int a = 0; a < 3; ++a) {
                const int q = m.tri[3 * t + a];
                int where = -1;
                for (int c = P.rp[p]; c < P.rp[p + 1]; ++c)
                    if (P.ci[c] == q) where = c;
                P.pos[(size_t)s * 3 + a] = where;
            }
        }
    }
}

__global__ void __launch_bounds__(256)
elem_contrib(int S, int kind, double c_mass, double c_lap, int n_zones,
             const double* __restrict__ coords, const int* __restrict__ tri,
             const int* __restrict__ slot_row, const int* __restrict__ slot_tri,
             const int* __restrict__ slot_local, const int* __restrict__ row_zone,
             double* __restrict__ contrib)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int t = slot_tri[s];
    const int z = row_zone[slot_row[s]];
    double xs[3], ys[3];
    for (int a = 0; a < 3; ++a) {
        const int p = tri[3 * t + a];
        xs[a] = coords[((size_t)p * n_zones + z) * 2 + 0];
        ys[a] = coords[((size_t)p * n_zones + z) * 2 + 1];
    }
    ElemTables tab;
    if (elem_needs_tables(kind))
        elem_tables(trap_geom(xs[0], xs[1], xs[2], ys[0], ys[1], ys[2]), tab);
    P1Poly phi[3];
    p1_basis(xs, ys, phi);
    const P1Poly& phi_i = phi[slot_local[s]];
    for (int a = 0; a < 3; ++a)
        contrib[(size_t)s * 3 + a] = elem_value(kind, tab, xs, ys, phi_i, phi[a], c_mass, c_lap);
}

__global__ void __launch_bounds__(256)
row_reduce(int rs, const int* __restrict__ rp, const int* __restrict__ slot_ptr,
           const int* __restrict__ pos, const double* __restrict__ contrib,
           double* __restrict__ vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rs) return;
    const int r0 = rp[i], r1 = rp[i + 1];
    const int width = r1 - r0;
    if (width <= 16) {
        double acc[16];
        for (int c = 0; c < 16; ++c) acc[c] = 0.0;
        for (int s = slot_ptr[i]; s < slot_ptr[i + 1]; ++s)
            for (int a = 0; a < 3; ++a) {
                const int q = pos[(size_t)s * 3 + a];
                if (q >= 0) acc[q - r0] += contrib[(size_t)s * 3 + a];
            }
        for (int c = 0; c < width; ++c) vals[r0 + c] = acc[c];
    } else {
        for (int c = r0; c < r1; ++c) vals[c] = 0.0;
        for (int s = slot_ptr[i]; s < slot_ptr[i + 1]; ++s)
            for (int a = 0; a < 3; ++a) {
                const int q = pos[(size_t)s * 3 + a];
                if (q >= 0) vals[q] += contrib[(size_t)s * 3 + a];
            }
    }
}

}  // namespace

int assemble_device(phg_mesh_s* M, int kind, double c_mass, double c_lap,
                    hipStream_t stream, phg_matrix_s** out, bool full)
{
    const SphereMesh& m = M->h;
    AsmPlan P;
    if (full) build_plan_full(m, P);
    else build_plan(m, P);
    const int rs = full ? m.n_points : (int)m.inner.size();
    const int S = P.slot_ptr[rs];
    int rc = ensure_mesh_device(M);
    if (rc) return rc;

    DevBuf<int> slot_row, slot_tri, slot_local, row_zone, slot_ptr, pos;
    DevBuf<double> contrib;
    PHG_HIP(slot_row.upload(P.slot_row.data(), S, stream));
    PHG_HIP(slot_tri.upload(P.slot_tri.data(), S, stream));
    PHG_HIP(slot_local.upload(P.slot_local.data(), S, stream));
    PHG_HIP(row_zone.upload(P.row_zone.data(), rs, stream));
    PHG_HIP(slot_ptr.upload(P.slot_ptr.data(), rs + 1, stream));
    PHG_HIP(pos.upload(P.pos.data(), (size_t)S * 3, stream));
    PHG_HIP(contrib.alloc((size_t)S * 3));

    std::unique_ptr<phg_matrix_s> A(new phg_matrix_s());
    A->n_rows = rs;
    A->n_cols = rs;
    A->nnz = (long)P.ci.size();
    A->h_rp = P.rp;
    A->h_ci = P.ci;
    PHG_HIP(A->rp.upload(P.rp.data(), rs + 1, stream));
    PHG_HIP(A->ci.upload(P.ci.data(), P.ci.size(), stream));
    PHG_HIP(A->val.alloc(P.ci.size()));

    if (S > 0) {
        hipLaunchKernelGGL(elem_contrib, dim3((S + 255) / 256), dim3(256), 0, stream, S,
                           kind, c_mass, c_lap, m.n_zones, M->d_coords.p, M->d_tri.p,
                           slot_row.p, slot_tri.p, slot_local.p, row_zone.p, contrib.p);
        PHG_HIP(hipGetLastError());
        hipLaunchKernelGGL(row_reduce, dim3((rs + 255) / 256), dim3(256), 0, stream, rs,
                           A->rp.p, slot_ptr.p, pos.p, contrib.p, A->val.p);
        PHG_HIP(hipGetLastError());
    }
    PHG_HIP(hipStreamSynchronize(stream));
    *out = A.release();
    return PHG_OK;
}

// generate_boundary_matrix (generators.cpp:141-212): the inner x outer
// matrix.  Structure and entry positions on the host; values on the device,
// one thread per outer point walking its triangles serially in the
// reference's order (6 entries on the full sphere).
namespace {
__global__ void bnd_assemble(int os, int kind, double c_mass, double c_lap, int n_zones,
                             const double* __restrict__ coords, const int* __restrict__ tri,
                             const int* __restrict__ outer_zone,
                             const int* __restrict__ oslot_ptr, const int* __restrict__ oslot_tri,
                             const int* __restrict__ oslot_local, const int* __restrict__ opos,
                             double* __restrict__ vals)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= os) return;
    const int z = outer_zone[j];
    for (int s = oslot_ptr[j]; s < oslot_ptr[j + 1]; ++s) {
        const int t = oslot_tri[s];
        double xs[3], ys[3];
        for (int a = 0; a < 3; ++a) {
            const int p = tri[3 * t + a];
            xs[a] = coords[((size_t)p * n_zones + z) * 2 + 0];
            ys[a] = coords[((size_t)p * n_zones + z) * 2 + 1];
        }
        ElemTables tab;
        if (elem_needs_tables(kind))
            elem_tables(trap_geom(xs[0], xs[1], xs[2], ys[0], ys[1], ys[2]), tab);
        P1Poly phi[3];
        p1_basis(xs, ys, phi);
        const P1Poly& phi_j = phi[oslot_local[s]];
        for (int a = 0; a < 3; ++a) {
            const int q = opos[(size_t)s * 3 + a];
            if (q >= 0)
                vals[q] += elem_value(kind, tab, xs, ys, phi[a], phi_j, c_mass, c_lap);
        }
    }
}
}  // namespace

int assemble_boundary_device(phg_mesh_s* M, int kind, double c_mass, double c_lap,
                             hipStream_t stream, phg_matrix_s** out)
{
    const SphereMesh& m = M->h;
    const int rs = (int)m.inner.size();
    const int os = (int)m.outer.size();
    std::vector<std::vector<int>> cols(rs);
    std::vector<int> oslot_ptr(os + 1, 0), oslot_tri, oslot_local, outer_zone(os);
    for (int j = 0; j < os; ++j) {
        const int p2 = m.outer[j];
        outer_zone[j] = m.tri_zone[m.adj[m.adj_ptr[p2]]];
        for (int e = m.adj_ptr[p2]; e < m.adj_ptr[p2 + 1]; ++e) {
            const int t = m.adj[e];
            oslot_tri.push_back(t);
            int loc = 0;
            while (m.tri[3 * t + loc] != p2) ++loc;
            oslot_local.push_back(loc);
            for (int a = 0; a < 3; ++a) {
                const int p = m.tri[3 * t + a];
                if (m.is_boundary[p]) continue;
                auto& c = cols[m.p2io[p]];
                if (std::find(c.begin(), c.end(), j) == c.end()) c.push_back(j);
            }
        }
        oslot_ptr[j + 1] = (int)oslot_tri.size();
    }
    std::vector<int> rp(rs + 1, 0), ci;
    for (int i = 0; i < rs; ++i) {
        std::sort(cols[i].begin(), cols[i].end());
        ci.insert(ci.end(), cols[i].begin(), cols[i].end());
        rp[i + 1] = (int)ci.size();
    }
    const int S = oslot_ptr[os];
    std::vector<int> opos((size_t)S * 3, -1);
    for (int j = 0; j < os; ++j)
        for (int s = oslot_ptr[j]; s < oslot_ptr[j + 1]; ++s)
            for (int a = 0; a < 3; ++a) {
                const int p = m.tri[3 * oslot_tri[s] + a];
                if (m.is_boundary[p]) continue;
                const int i = m.p2io[p];
                for (int c = rp[i]; c < rp[i + 1]; ++c)
                    if (ci[c] == j) opos[(size_t)s * 3 + a] = c;
            }
    std::vector<double> zeros(ci.size(), 0.0);
    phg_matrix_s* A = nullptr;
    int rc = matrix_from_csr_host(rs, os, rp.data(), ci.data(), zeros.data(), &A);
    if (rc) return rc;
    std::unique_ptr<phg_matrix_s> hold(A);
    if (S > 0) {
        rc = ensure_mesh_device(M);
        if (rc) return rc;
        DevBuf<int> d_zone, d_ptr, d_tri, d_loc, d_pos;
        PHG_HIP(d_zone.upload(outer_zone.data(), os, stream));
        PHG_HIP(d_ptr.upload(oslot_ptr.data(), os + 1, stream));
        PHG_HIP(d_tri.upload(oslot_tri.data(), S, stream));
        PHG_HIP(d_loc.upload(oslot_local.data(), S, stream));
        PHG_HIP(d_pos.upload(opos.data(), (size_t)S * 3, stream));
        hipLaunchKernelGGL(bnd_assemble, dim3((os + 63) / 64), dim3(64), 0, stream, os, kind,
                           c_mass, c_lap, m.n_zones, M->d_coords.p, M->d_tri.p, d_zone.p,
                           d_ptr.p, d_tri.p, d_loc.p, d_pos.p, A->val.p);
        PHG_HIP(hipGetLastError());
        PHG_HIP(hipStreamSynchronize(stream));
    }
    *out = hold.release();
    return PHG_OK;
}

}  // namespace phg
