// Kernels on per-member vectors stored slot-fastest: a V x S (every mesh
// point) or n x S (inner points) row-major array holds S members' vectors,
// entry (point, slot) at point * S + slot.  Every kernel touches only the
// columns whose mask entry is set, so a slot's arithmetic never depends on
// the others.  Used by the BarVortex and Baroclin engines; internal linkage
// (each translation unit has its own copy).
#pragma once

#include <hip/hip_runtime.h>

namespace {

constexpr int BV_MAXS = 8;   // slots per engine
constexpr int BV_BS = 256;   // threads per block

inline dim3 g256(size_t work) { return dim3((unsigned)((work + BV_BS - 1) / BV_BS)); }

// Y[r][s] = beta * Y[r][s] + sum_c A[r][c] X[c][s] (CSR order, no fma); only
// columns with mask[s] != 0.
__global__ void bv_spmm(int rows, int S, const int* __restrict__ rp, const int* __restrict__ ci,
                        const double* __restrict__ val, const double* __restrict__ X,
                        double* __restrict__ Y, int accumulate, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * S) return;
    const int r = (int)(e / S), s = (int)(e % S);
    if (!mask[s]) return;
    double t = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) t += val[k] * X[(size_t)ci[k] * S + s];
    Y[e] = accumulate ? Y[e] + t : t;
}

// p[i][s] = u[inner[i]][s]
__global__ void bv_u2p(int n, int S, const int* __restrict__ inner, const double* __restrict__ u,
                       double* __restrict__ p, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * S) return;
    const int i = (int)(e / S), s = (int)(e % S);
    p[e] = mask[s] ? u[(size_t)inner[i] * S + s] : 0.0;
}

// u[k][s] = boundary ? 0 : p[p2io[k]][s]  (p2u with zero boundary values)
__global__ void bv_p2u(int V, int S, const int* __restrict__ isb, const int* __restrict__ p2io,
                       const double* __restrict__ p, double* __restrict__ u,
                       const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    const int k = (int)(e / S), s = (int)(e % S);
    if (!mask[s]) return;
    u[e] = isb[k] ? 0.0 : p[(size_t)p2io[k] * S + s];
}

// u[k][s] = boundary ? bnd[p2io[k]][s] : p[p2io[k]][s]  (p2u, src/proj.cpp:49-75;
// bnd is outer x S and p2io numbers the boundary points in outer order)
__global__ void bv_p2u_bnd(int V, int S, const int* __restrict__ isb, const int* __restrict__ p2io,
                           const double* __restrict__ p, const double* __restrict__ bnd,
                           double* __restrict__ u, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    const int k = (int)(e / S), s = (int)(e % S);
    if (!mask[s]) return;
    u[e] = (isb[k] ? bnd : p)[(size_t)p2io[k] * S + s];
}

// column s of a rows x S array = v (or 0 when v is null)
// u = set_bnd(u, bnd) for the masked columns (src/proj.cpp:143-160): the
// boundary points take bnd (outer x S, indexed by outer number), the inner
// points keep their values.
__global__ void bv_set_bnd_cols(int V, int S, const int* __restrict__ isb,
                                const int* __restrict__ p2io, const double* __restrict__ bnd,
                                double* __restrict__ u, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    const int k = (int)(e / S), s = (int)(e % S);
    if (mask[s] && isb[k]) u[e] = bnd[(size_t)p2io[k] * S + s];
}

__global__ void bv_col_set(int rows, int S, int s, const double* __restrict__ v,
                           double* __restrict__ M)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < rows) M[(size_t)k * S + s] = v ? v[k] : 0.0;
}

// r = k1 a + k2 b (vec_sum1), masked columns; zero elsewhere when zero_rest.
__global__ void bv_axpby(size_t rows, int S, double k1, const double* __restrict__ a, double k2,
                         const double* __restrict__ b, double* __restrict__ r,
                         const int* __restrict__ mask, int zero_rest)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * S) return;
    const int s = (int)(e % S);
    if (mask[s]) r[e] = k1 * a[e] + k2 * b[e];
    else if (zero_rest) r[e] = 0.0;
}

// r = a + k * lh (vec_sum2 with a mesh function shared by all slots)
__global__ void bv_add_shared(int V, int S, const double* __restrict__ a, double k,
                              const double* __restrict__ lh, double* __restrict__ r,
                              const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    const int p = (int)(e / S), s = (int)(e % S);
    if (mask[s]) r[e] = a[e] + k * lh[p];
}

// F[i][s] = FC[inner[i]][s] - jac[i][s]
__global__ void bv_rhs(int n, int S, const int* __restrict__ inner, const double* __restrict__ FC,
                       const double* __restrict__ jac, double* __restrict__ F,
                       const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * S) return;
    const int i = (int)(e / S), s = (int)(e % S);
    F[e] = mask[s] ? FC[(size_t)inner[i] * S + s] - jac[e] : 0.0;
}

__global__ void bv_zero_cols(size_t rows, int S, double* __restrict__ x,
                             const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * S) return;
    if (!mask[e % S]) x[e] = 0.0;
}

__global__ void bv_copy_cols(size_t rows, int S, const double* __restrict__ src,
                             double* __restrict__ dst, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * S) return;
    if (mask[e % S]) dst[e] = src[e];
}

// column s of a V x S array <-> a contiguous vector
__global__ void bv_col_in(int V, int S, int s, const double* __restrict__ v, double* __restrict__ M)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < V) M[(size_t)k * S + s] = v[k];
}
__global__ void bv_col_out(int V, int S, int s, const double* __restrict__ M, double* __restrict__ v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < V) v[k] = M[(size_t)k * S + s];
}

// Block partials of d^T N d per column, d = a - b (b may be null): N is the
// full V x V mass matrix (CSR).  Thread per (row, column), 256/S rows per
// block; part[s * G + block].
__global__ void __launch_bounds__(BV_BS)
bv_normsq(int V, int S, const int* __restrict__ rp, const int* __restrict__ ci,
          const double* __restrict__ val, const double* __restrict__ a,
          const double* __restrict__ b, double* __restrict__ part, int G,
          const int* __restrict__ mask)
{
    __shared__ double sh[BV_BS / 64][BV_MAXS];
    const int s = threadIdx.x % S;
    const int r = blockIdx.x * (BV_BS / S) + (int)threadIdx.x / S;
    double acc = 0.0;
    if (r < V && mask[s]) {
        double t = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) {
            const size_t o = (size_t)ci[k] * S + s;
            t += val[k] * (b ? a[o] - b[o] : a[o]);
        }
        const size_t o = (size_t)r * S + s;
        acc = (b ? a[o] - b[o] : a[o]) * t;
    }
    for (int off = S; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane < S) sh[wid][lane] = acc;
    __syncthreads();
    if ((int)threadIdx.x < S) {
        double t = 0.0;
        for (int w = 0; w < BV_BS / 64; ++w) t += sh[w][threadIdx.x];
        part[(size_t)threadIdx.x * G + blockIdx.x] = t;
    }
}

// out[s] = sqrt(sum of the G partials of column s), fixed order, one wavefront per column.
__global__ void bv_sum_cols(int S, const double* __restrict__ part, int G, double* __restrict__ out)
{
    const int s = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (s >= S) return;
    double t = 0.0;
    for (int b = lane; b < G; b += 64) t += part[(size_t)s * G + b];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (lane == 0) out[s] = sqrt(t);
}

// Ub[o][s] = u[outer[o]][s]
__global__ void bv_outer(int os, int S, const int* __restrict__ outer, const double* __restrict__ u,
                         double* __restrict__ Ub, const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)os * S) return;
    const int o = (int)(e / S), s = (int)(e % S);
    Ub[e] = mask[s] ? u[(size_t)outer[o] * S + s] : 0.0;
}

// x[e] = 0 for every e < count, every column (a fill kernel rather than a
// memset, so that it may sit in captured work)
__global__ void bv_fill0(size_t count, double* __restrict__ x)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) x[e] = 0.0;
}

__global__ void bv_set_cols(size_t rows, int S, double v, double* __restrict__ x,
                            const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * S) return;
    if (mask[e % S]) x[e] = v;
}

// r[p][s] = lh[p] for the masked columns, 0 elsewhere
__global__ void bv_bcast(int V, int S, const double* __restrict__ lh, double* __restrict__ r,
                         const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    r[e] = mask[e % S] ? lh[e / S] : 0.0;
}

// Right side of the linearised and adjoint steps, in the reference's order:
//   jt = ((ka A + kb B) + kc C) * sj
//   r  = (((base or 0) + c1 P1) + c2 P2) + jt,  0 on boundary points if zb
// (calc_L: A, B, C = J(u, lh), J(z, L u), J(u, L z), ka = k2, kb = kc = k1;
// calc_LT: A, B, C = J^T(v1, L z), L J^T(z, v1), J^T(v1, lh), ka = kb = k1,
// kc = k2.)  Masked columns only.
__global__ void bv_lin_sum(int V, int S, const double* __restrict__ base,
                           const double* __restrict__ P1, double c1,
                           const double* __restrict__ P2, double c2,
                           const double* __restrict__ A, double ka,
                           const double* __restrict__ B, double kb,
                           const double* __restrict__ C, double kc, double sj,
                           const int* __restrict__ isb, int zb, double* __restrict__ r,
                           const int* __restrict__ mask)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)V * S) return;
    if (!mask[e % S]) return;
    const double jt = ((ka * A[e] + kb * B[e]) + kc * C[e]) * sj;
    const double p1 = P1[e] * c1, p2 = P2[e] * c2;
    const double v = (((base ? base[e] : 0.0) + p1) + p2) + jt;
    r[e] = zb && isb[e / S] ? 0.0 : v;
}

}  // namespace
