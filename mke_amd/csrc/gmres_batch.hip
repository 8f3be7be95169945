// Batched restarted GMRES(m): K <= 64 independent systems A x_q = b_q that
// share one matrix, on gfx950.
//
// Each column keeps its own Krylov space, Hessenberg matrix, Givens
// rotations and stopping test, exactly as K separate solves would (no block
// Krylov coupling), while every matrix read of an Arnoldi step serves all K
// columns (one SpMM instead of K SpMVs).  It is the form the callers with
// several right-hand sides per operator need: the eight mass solves of one
// SphereJacobian::calc2 (src/sjacobian.cpp:136-362: 16 idt_.solve calls with
// 8 distinct right-hand sides) and the implicit-step solves of the ensemble
// members that share one mesh level.
//
// Layout: vectors are n x KP row-major (entry (i, q) at i * KP + q), KP = K
// rounded up to a power of two; a thread owns one (row, column) pair, so a
// wavefront covers 64 / KP consecutive rows and the gathers of one ELL entry
// read KP consecutive words.  The caller's B and X are n x K, row-major.
//
// Orthogonalisation: classical Gram-Schmidt with one re-orthogonalisation
// pass (CGS2), or a single pass (CGS1) for well-conditioned operators such
// as the mass matrix.  Precision: FP64 throughout, or the mixed form (FP32
// basis, FP32/BF16 matrix values, FP64 dots, Hessenberg, Givens, update and
// a true FP64 residual at every restart, i.e. iterative refinement).
//
// Reductions are deterministic and do not depend on the column's position
// or on the other columns: a column's result is bit-identical whatever else
// is in the batch, provided KP is the same.  A restart cycle is captured in
// one hipGraph and replayed; the host reads a K-entry summary per cycle.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "ell.h"
#include "phg_internal.h"

namespace phg {

constexpr int BB_BS = 256;
constexpr int BB_MAXM = 32;
constexpr int BB_MAXK = 64;
constexpr int BB_RPT = 16;  // largest rows per thread of a launch shape
constexpr int BB_LB = 4;   // basis vectors per load batch

// A restart cycle runs its Arnoldi steps in chunks ending at steps 4, 8, 16,
// 24 and 32; after each chunk but the last the host reads how many columns
// are still running and, when none is, ends the cycle there.  The steps
// skipped would have found every column stopped: same results, without the
// launches (and the preconditioner applications) of a full cycle, which
// matters for solves that stop after a few steps (multigrid-preconditioned,
// warm-started).
constexpr int BB_NCHUNK = 5;
__host__ __device__ constexpr int bb_chunk_end(int c) { return c < 2 ? 4 * (c + 1) : 8 * c; }

struct BCol {
    double H[(BB_MAXM + 1) * BB_MAXM];  // unrotated Hessenberg, H[i * BB_MAXM + j]
    double R[BB_MAXM * BB_MAXM];        // rotated, upper triangular
    double cs[BB_MAXM], sn[BB_MAXM], g[BB_MAXM], carry[BB_MAXM + 1], y[BB_MAXM];
    double hc[BB_MAXM + 1];             // coefficients of the current CGS pass
    double beta, bnorm, rel, hn;
    int k, active, converged, iters, iters0, cycles, pad0, pad1;
};

struct BSum {
    double rel;
    int active, converged, iters, cycles;
};

struct BatchWork {
    int n = 0, m = 0, kp = 0, G = 0, prec = 0, rpt = 0;
    DevBuf<char> V, W;    // (m+1) x n x kp and 2 x n x kp of the Krylov type
    DevBuf<char> Z;       // m x n x kp preconditioned vectors (with a preconditioner)
    DevBuf<double> part;  // (m+2) x kp x G block partials
    DevBuf<BCol> cols;    // kp
    DevBuf<BSum> sum;     // kp
    BSum* host = nullptr;
    DevBuf<int> nact;      // columns still running in the current cycle
    int* hnact = nullptr;  // its host copy (pinned)
    hipStream_t cap = nullptr;
    // one graph per chunk of Arnoldi steps of a cycle, then the cycle's end
    hipGraphExec_t exec[BB_NCHUNK + 1] = {};
    bool captured = false;
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    int key_m = 0, key_ortho = -1, key_k = 0, key_maxit = 0;
    const void* key_pc = nullptr;
    double key_rtol = 0.0, key_inner = 0.0;
    void drop_graphs()
    {
        for (auto& e : exec)
            if (e) (void)hipGraphExecDestroy(e), e = nullptr;
        captured = false;
    }
    ~BatchWork()
    {
        drop_graphs();
        if (cap) (void)hipStreamDestroy(cap);
        if (host) (void)hipHostFree(host);
        if (hnact) (void)hipHostFree(hnact);
    }
};

void batch_destroy(BatchWork* w) { delete w; }

namespace {

__device__ __forceinline__ double wave_sum64(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the lanes of a wavefront that hold the same column (lane % KP).
__device__ __forceinline__ double col_wave_sum(double v, int KP)
{
    for (int off = KP; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Fixed-order sum of np partials by one wavefront; every lane returns it.
__device__ __forceinline__ double wave_sum_partials(const double* __restrict__ p, int np)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    int b = lane;
    for (; b + 64 * 7 < np; b += 64 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[b + 64 * u];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; b < np; b += 64) s += p[b];
    return wave_sum64(s);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename T>
__device__ __forceinline__ double dd(T v) { return (double)v; }

// Row of A times column q of X (n x KP), source scaled by s: FP64 without
// fused multiply-add, or FP32 with fma over FP32/BF16 values.
template <typename VT, typename MT>
struct BRow;
template <>
struct BRow<double, double> {
    __device__ __forceinline__ static double run(int n, int width, int row,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ X, int KP, int q,
                                                 double s)
    {
        if (width > ELL_MAXW) {  // long rows (block systems): the same sum, not unrolled
            double r = 0.0;
            for (int k = 0; k < width; ++k) {
                const size_t o = (size_t)k * n + row;
                r += val[o] * (X[(size_t)col[o] * KP + q] * s);
            }
            return r;
        }
        int c[ELL_MAXW];
        double v[ELL_MAXW], xv[ELL_MAXW];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = (size_t)k * n + row;
            c[k] = k < width ? col[o] : 0;
            v[k] = k < width ? val[o] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < width ? X[(size_t)c[k] * KP + q] : 0.0;
        double r = 0.0;
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k)
            if (k < width) r += v[k] * (xv[k] * s);
        return r;
    }
};
template <typename MT>
struct BRow<float, MT> {
    __device__ __forceinline__ static float run(int n, int width, int row,
                                                const int* __restrict__ col,
                                                const MT* __restrict__ val,
                                                const float* __restrict__ X, int KP, int q,
                                                float s)
    {
        if (width > ELL_MAXW) {
            float r = 0.0f;
            for (int k = 0; k < width; ++k) {
                const size_t o = (size_t)k * n + row;
                r = fmaf(ell_val_f32(val[o]), X[(size_t)col[o] * KP + q] * s, r);
            }
            return r;
        }
        int c[ELL_MAXW];
        MT v[ELL_MAXW];
        float xv[ELL_MAXW];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = (size_t)k * n + row;
            c[k] = k < width ? col[o] : 0;
            v[k] = k < width ? val[o] : MT(0);
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < width ? X[(size_t)c[k] * KP + q] : 0.0f;
        float r = 0.0f;
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k)
            if (k < width) r = fmaf(ell_val_f32(v[k]), xv[k] * s, r);
        return r;
    }
};

// One term of BRow's sum: r + val * (x * s) in FP64 (no fused multiply-add),
// fmaf(val, x * s, r) in FP32.
template <typename VT, typename MT>
struct BTerm;
template <>
struct BTerm<double, double> {
    __device__ __forceinline__ static double add(double r, double v, double x, double s)
    {
        return r + v * (x * s);
    }
};
template <typename MT>
struct BTerm<float, MT> {
    __device__ __forceinline__ static float add(float r, MT v, float x, float s)
    {
        return fmaf(ell_val_f32(v), x * s, r);
    }
};

// BRow of this thread's rows base + t * RB (t < RPT) of column q, and xo[t]
// = X(row, q); w[t] = xo[t] = 0 for rows at or past n and inactive columns.
// For rows of at most ELL_MAXW entries, RG rows at a time: every row's own
// entry and column/value slots (ell_slot: past the width the last slot
// again, in bounds) are loaded first, then every row's gathers, then the
// sums in BRow's order; two memory round trips per group instead of three
// per row.
template <typename VT, typename MT, int RPT, int RG>
__device__ __forceinline__ void brows(int n, int width, int KP, int q, int base, int RB, bool act,
                                      const int* __restrict__ col, const MT* __restrict__ val,
                                      const VT* __restrict__ X, VT s, VT (&xo)[RPT],
                                      VT (&w)[RPT])
{
    if (width >= 1 && width <= ELL_MAXW) {
#pragma unroll
        for (int g = 0; g < RPT; g += RG) {
            int c[RG][ELL_MAXW];
            MT v[RG][ELL_MAXW];
            VT xv[RG][ELL_MAXW];
#pragma unroll
            for (int u = 0; u < RG; ++u) {
                const int row = min(base + (g + u) * RB, n - 1);  // in bounds; unused past n
                xo[g + u] = X[(size_t)row * KP + q];
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k) {
                    const size_t o = ell_slot(k, width, n, row);
                    c[u][k] = col[o];
                    v[u][k] = val[o];
                }
            }
#pragma unroll
            for (int u = 0; u < RG; ++u)
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k) xv[u][k] = X[(size_t)c[u][k] * KP + q];
#pragma unroll
            for (int u = 0; u < RG; ++u) {
                VT r = VT(0);
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k) {
                    const VT t = BTerm<VT, MT>::add(r, v[u][k], xv[u][k], s);
                    r = k < width ? t : r;
                }
                const bool in = act && base + (g + u) * RB < n;
                w[g + u] = in ? r : VT(0);
                xo[g + u] = in ? xo[g + u] : VT(0);
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < RPT; ++t) {
            const int row = base + t * RB;
            w[t] = VT(0);
            xo[t] = VT(0);
            if (act && row < n) {
                xo[t] = X[(size_t)row * KP + q];
                w[t] = BRow<VT, MT>::run(n, width, row, col, val, X, KP, q, s);
            }
        }
    }
}

// Block partials of <V_i, w> for i < nd (i == own_i taken from `own`), or
// of ||w||^2 when nd == 0, for this thread's rows; written to
// part[((out0 + i) * KP + q) * G + block].  All threads of the block call it.
template <typename VT, int RPT>
__device__ __forceinline__ void bdots(int nd, int own_i, const VT* own, const VT* __restrict__ V,
                                      size_t bs, int n, int KP, int q, int base, int RB,
                                      bool act, const VT* w, double* __restrict__ part, int G,
                                      int out0)
{
    __shared__ double sh[BB_BS / 64][BB_LB][BB_MAXK];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nout = nd == 0 ? 1 : nd;
    for (int i0 = 0; i0 < nout; i0 += BB_LB) {
        double acc[BB_LB];
        if (nd == 0) {
            double a = 0.0;
#pragma unroll
            for (int t = 0; t < RPT; ++t) {
                const int row = base + t * RB;
                if (act && row < n) a += dd(w[t]) * dd(w[t]);
            }
            acc[0] = a;
#pragma unroll
            for (int u = 1; u < BB_LB; ++u) acc[u] = 0.0;
        } else {
            VT vb[BB_LB][RPT];
#pragma unroll
            for (int u = 0; u < BB_LB; ++u) {
                const int i = i0 + u;
#pragma unroll
                for (int t = 0; t < RPT; ++t) {
                    const int row = base + t * RB;
                    vb[u][t] = (act && i < nd && i != own_i && row < n)
                                   ? V[(size_t)i * bs + (size_t)row * KP + q] : VT(0);
                }
            }
#pragma unroll
            for (int u = 0; u < BB_LB; ++u) {
                const int i = i0 + u;
                double a = 0.0;
#pragma unroll
                for (int t = 0; t < RPT; ++t) {
                    const int row = base + t * RB;
                    if (act && i < nd && row < n) a += dd(w[t]) * dd(i == own_i ? own[t] : vb[u][t]);
                }
                acc[u] = a;
            }
        }
#pragma unroll
        for (int u = 0; u < BB_LB; ++u) {
            const double s = col_wave_sum(acc[u], KP);
            if (lane < KP) sh[wid][u][lane] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < BB_LB * KP; e += BB_BS) {
            const int u = e / KP, qq = e % KP;
            if (i0 + u < nout) {
                double s = 0.0;
#pragma unroll
                for (int ww = 0; ww < BB_BS / 64; ++ww) s += sh[ww][u][qq];
                part[((size_t)(out0 + i0 + u) * KP + qq) * G + lblock()] = s;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- kernels --
// Arnoldi step j: v_j = W_src / h (stored), W_dst = A v_j, partial dots of
// W_dst with v_0..v_j (first CGS pass).
template <typename VT, typename MT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_step_spmm(int n, int KP, int K, int width, const int* __restrict__ col,
             const MT* __restrict__ val, const VT* __restrict__ Wsrc, VT* __restrict__ V,
             size_t bs, VT* __restrict__ Wdst, int j, double* __restrict__ part, int G,
             const BCol* __restrict__ cols)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K && cols[q].active;
    double h = 1.0;
    if (act) h = j == 0 ? cols[q].beta : cols[q].hn;
    const VT inv = (VT)(1.0 / h);
    VT* Vj = V + (size_t)j * bs;
    VT w[RPT], vj[RPT], xo[RPT];
    brows<VT, MT, RPT, (RPT < 4 ? RPT : 4)>(n, width, KP, q, base, RB, act, col, val, Wsrc, inv,
                                            xo, w);
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        vj[t] = VT(0);
        if (act && row < n) {
            const size_t o = (size_t)row * KP + q;
            vj[t] = xo[t] * inv;
            Vj[o] = vj[t];
            Wdst[o] = w[t];
        }
    }
    bdots<VT, RPT>(j + 1, j, vj, V, bs, n, KP, q, base, RB, act, w, part, G, 0);
}

// Right-preconditioned step: v_j = W_src / h (stored) for the preconditioner.
template <typename VT>
__global__ void __launch_bounds__(BB_BS)
bk_scale(int n, int KP, int K, const VT* __restrict__ Wsrc, VT* __restrict__ Vj, int j,
         const BCol* __restrict__ cols)
{
    const size_t e = (size_t)lblock() * BB_BS + threadIdx.x;
    if (e >= (size_t)n * KP) return;
    const int q = (int)(e % KP);
    if (q >= K || !cols[q].active) return;
    const double h = j == 0 ? cols[q].beta : cols[q].hn;
    Vj[e] = Wsrc[e] * (VT)(1.0 / h);
}

// W_dst = A z_j (z_j = M^-1 v_j, the preconditioned vector) and the partial
// dots of W_dst with v_0..v_j.
template <typename VT, typename MT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_step_spmm_pc(int n, int KP, int K, int width, const int* __restrict__ col,
                const MT* __restrict__ val, const VT* __restrict__ Zj, const VT* __restrict__ V,
                size_t bs, VT* __restrict__ Wdst, int j, double* __restrict__ part, int G,
                const BCol* __restrict__ cols)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K && cols[q].active;
    VT w[RPT], xo[RPT];
    brows<VT, MT, RPT, (RPT < 4 ? RPT : 4)>(n, width, KP, q, base, RB, act, col, val, Zj, VT(1),
                                            xo, w);
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        if (act && row < n) Wdst[(size_t)row * KP + q] = w[t];
    }
    bdots<VT, RPT>(j + 1, -1, nullptr, V, bs, n, KP, q, base, RB, act, w, part, G, 0);
}

// Sums the partials of the nd dots of a pass, one wavefront per (i, column):
// pass 0 sets H(i, j), pass 1 adds the re-orthogonalisation correction.
__global__ void __launch_bounds__(BB_BS)
bk_reduce(int KP, int K, int nd, int j, int pass, const double* __restrict__ part, int G,
          BCol* __restrict__ cols)
{
    const int wv = lblock() * (BB_BS / 64) + ((int)threadIdx.x >> 6);
    if (wv >= nd * KP) return;
    const int i = wv / KP, q = wv % KP;
    if (q >= K || !cols[q].active) return;
    const double s = wave_sum_partials(part + ((size_t)i * KP + q) * G, G);
    if ((threadIdx.x & 63) == 0) {
        cols[q].hc[i] = s;
        if (pass == 0) cols[q].H[i * BB_MAXM + j] = s;
        else cols[q].H[i * BB_MAXM + j] += s;
    }
}

// w -= sum_{i<=j} hc_i v_i, then the partial dots with v_0..v_j (next pass)
// or the partial of ||w||^2 (final pass).
template <typename VT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_update(int n, int KP, int K, VT* __restrict__ W, const VT* __restrict__ V, size_t bs, int j,
          int final_pass, double* __restrict__ part, int G, const BCol* __restrict__ cols)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K && cols[q].active;
    VT w[RPT];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        w[t] = (act && row < n) ? W[(size_t)row * KP + q] : VT(0);
    }
    for (int i0 = 0; i0 <= j; i0 += BB_LB) {
        VT vb[BB_LB][RPT];
        VT hv[BB_LB];
#pragma unroll
        for (int u = 0; u < BB_LB; ++u) {
            const int i = i0 + u;
            hv[u] = (act && i <= j) ? (VT)cols[q].hc[i] : VT(0);
#pragma unroll
            for (int t = 0; t < RPT; ++t) {
                const int row = base + t * RB;
                vb[u][t] = (act && i <= j && row < n) ? V[(size_t)i * bs + (size_t)row * KP + q]
                                                      : VT(0);
            }
        }
#pragma unroll
        for (int u = 0; u < BB_LB; ++u)
#pragma unroll
            for (int t = 0; t < RPT; ++t) w[t] = w[t] - hv[u] * vb[u][t];
    }
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        if (act && row < n) W[(size_t)row * KP + q] = w[t];
    }
    if (final_pass) bdots<VT, RPT>(0, -1, nullptr, V, bs, n, KP, q, base, RB, act, w, part, G, j + 1);
    else bdots<VT, RPT>(j + 1, -1, nullptr, V, bs, n, KP, q, base, RB, act, w, part, G, 0);
}

// First CGS2 pass fused with the dots of the second: w -= sum_{i<=j} hc_i v_i
// row by row, with the row's v_0..v_j kept in registers for the partial
// dots of the updated w with v_0..v_j.  The same operations in the same
// order as bk_update (non-final pass) followed by its bdots, so the same
// bits, with one read of the basis instead of two.
template <typename VT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_update_dots(int n, int KP, int K, VT* __restrict__ W, const VT* __restrict__ V, size_t bs,
               int j, double* __restrict__ part, int G, const BCol* __restrict__ cols)
{
    __shared__ VT shv[BB_MAXM][BB_MAXK];
    __shared__ double sh[BB_BS / 64][BB_LB][BB_MAXK];
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K && cols[q].active;
    for (int e = threadIdx.x; e < BB_MAXM * KP; e += BB_BS) {
        const int i = e / KP, qq = e % KP;
        shv[i][qq] = (qq < K && i <= j && cols[qq].active) ? (VT)cols[qq].hc[i] : VT(0);
    }
    __syncthreads();
    double acc[BB_MAXM];
#pragma unroll
    for (int i = 0; i < BB_MAXM; ++i) acc[i] = 0.0;
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        if (!act || row >= n) continue;
        const size_t o = (size_t)row * KP + q;
        VT vv[BB_MAXM];
#pragma unroll
        for (int i = 0; i < BB_MAXM; ++i) vv[i] = i <= j ? V[(size_t)i * bs + o] : VT(0);
        VT w = W[o];
#pragma unroll
        for (int i = 0; i < BB_MAXM; ++i)
            if (i <= j) w = w - shv[i][q] * vv[i];
        W[o] = w;
#pragma unroll
        for (int i = 0; i < BB_MAXM; ++i)
            if (i <= j) acc[i] += dd(w) * dd(vv[i]);
    }
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i0 = 0; i0 < BB_MAXM; i0 += BB_LB) {
        if (i0 > j) break;
#pragma unroll
        for (int u = 0; u < BB_LB; ++u) {
            const double s = col_wave_sum(acc[i0 + u], KP);
            if (lane < KP) sh[wid][u][lane] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < BB_LB * KP; e += BB_BS) {
            const int u = e / KP, qq = e % KP;
            if (i0 + u <= j) {
                double s = 0.0;
#pragma unroll
                for (int ww = 0; ww < BB_BS / 64; ++ww) s += sh[ww][u][qq];
                part[((size_t)(i0 + u) * KP + qq) * G + lblock()] = s;
            }
        }
        __syncthreads();
    }
}

// Givens rotation of step j for each column (one wavefront per column) from
// H(0..j, j) and the norm partials; decides whether the column goes on.
__global__ void __launch_bounds__(BB_BS)
bk_givens(int KP, int K, int j, int m, const double* __restrict__ part, int G,
          BCol* __restrict__ cols, double rtol, double inner_rtol, int max_iters)
{
    const int q = lblock() * (BB_BS / 64) + ((int)threadIdx.x >> 6);
    if (q >= K) return;
    BCol* c = cols + q;
    if (!c->active) return;
    const int lane = threadIdx.x & 63;
    const double s = wave_sum_partials(part + ((size_t)(j + 1) * KP + q) * G, G);
    const double hreg = lane <= j ? c->H[lane * BB_MAXM + j] : 0.0;
    const double creg = lane < j ? c->cs[lane] : 0.0;
    const double sreg = lane < j ? c->sn[lane] : 0.0;
    const double hn = sqrt(s);
    double cur = readlane_d(hreg, 0);
    for (int i = 0; i < j; ++i) {
        const double cc = readlane_d(creg, i), sn = readlane_d(sreg, i);
        const double nxt = readlane_d(hreg, i + 1);
        const double t = cc * cur + sn * nxt;
        cur = -sn * cur + cc * nxt;
        if (lane == 0) c->R[i * BB_MAXM + j] = t;
    }
    const double r = hypot(cur, hn);
    const double cj = cur / r, sj = hn / r;
    const double carry = c->carry[j];
    const double gj = cj * carry, cn = -sj * carry;
    const int iters = c->iters0 + j + 1;
    const double stop = fmax(rtol * c->bnorm, inner_rtol * c->beta);
    const bool cont = !(fabs(cn) <= stop || !(hn > 0.0) || iters >= max_iters || j + 1 == m);
    if (lane == 0) {
        c->R[j * BB_MAXM + j] = r;
        c->H[(j + 1) * BB_MAXM + j] = hn;
        c->hn = hn;
        c->cs[j] = cj;
        c->sn[j] = sj;
        c->g[j] = gj;
        c->carry[j + 1] = cn;
        c->k = j + 1;
        c->iters = iters;
        if (!cont) c->active = 0;
    }
}

// y = R^{-1} g for each column that ran this cycle (one wavefront each:
// R staged in LDS by all lanes, back substitution by lane 0).
__global__ void __launch_bounds__(BB_BS)
bk_solve_y(int K, BCol* __restrict__ cols)
{
    __shared__ double Rs[BB_BS / 64][BB_MAXM * BB_MAXM];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = lblock() * (BB_BS / 64) + wid;
    if (q >= K) return;
    BCol* c = cols + q;
    const int k = c->k;
    if (c->converged || k == 0) return;
    double* R = Rs[wid];
    for (int e = lane; e < k * k; e += 64) R[e] = c->R[(e / k) * BB_MAXM + (e % k)];
    const double gl = lane < k ? c->g[lane] : 0.0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double y[BB_MAXM];
    for (int i = k - 1; i >= 0; --i) {
        double s = readlane_d(gl, i);
        for (int l = i + 1; l < k; ++l) s -= R[i * k + l] * y[l];
        y[i] = s / R[i * k + i];
    }
    if (lane == 0)
        for (int i = 0; i < k; ++i) c->y[i] = y[i];
}

// X += D (V y) per column (FP64 accumulation; D = dscale, or 1 when null: the
// diagonal right preconditioner); X is n x K (the caller's layout).
template <typename VT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_xupdate(int n, int KP, int K, const VT* __restrict__ V, size_t bs, double* __restrict__ X,
           const BCol* __restrict__ cols, const double* __restrict__ dscale)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    if (q >= K) return;
    const BCol* c = cols + q;
    const int k = c->k;
    if (c->converged || k == 0) return;
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        if (row >= n) break;
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += dd(V[(size_t)i * bs + (size_t)row * KP + q]) * c->y[i];
        X[(size_t)row * K + q] += dscale ? dscale[row] * s : s;
    }
}

// r = b - A x (FP64), W0 = r in the Krylov type, partial ||r||^2 per column.
template <typename VT, int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_residual(int n, int KP, int K, int width, const int* __restrict__ col,
            const double* __restrict__ val, const double* __restrict__ B,
            const double* __restrict__ X, VT* __restrict__ W0, double* __restrict__ part, int G,
            const BCol* __restrict__ cols)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K && !cols[q].converged;
    VT w[RPT];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        w[t] = VT(0);
        if (act && row < n) {
            double s = 0.0;
            if (width > ELL_MAXW) {  // long rows: the same sum, not unrolled
                for (int k = 0; k < width; ++k) {
                    const size_t o = (size_t)k * n + row;
                    s += val[o] * X[(size_t)col[o] * K + q];
                }
            } else {
                int c[ELL_MAXW];
                double v[ELL_MAXW], xv[ELL_MAXW];
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k) {
                    const size_t o = (size_t)k * n + row;
                    c[k] = k < width ? col[o] : 0;
                    v[k] = k < width ? val[o] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k)
                    xv[k] = k < width ? X[(size_t)c[k] * K + q] : 0.0;
#pragma unroll
                for (int k = 0; k < ELL_MAXW; ++k)
                    if (k < width) s += v[k] * xv[k];
            }
            const double r = B[(size_t)row * K + q] - s;
            w[t] = (VT)r;
            W0[(size_t)row * KP + q] = w[t];
        }
    }
    bdots<VT, RPT>(0, -1, nullptr, nullptr, 0, n, KP, q, base, RB, act, w, part, G, 0);
}

// ||b_q||^2 partials: the same reduction with B read as the vector.
template <int RPT>
__global__ void __launch_bounds__(BB_BS)
bk_sumsq(int n, int KP, int K, const double* __restrict__ B, double* __restrict__ part, int G)
{
    const int q = threadIdx.x & (KP - 1);
    const int RB = BB_BS / KP;
    const int base = lblock() * RB * RPT + (int)threadIdx.x / KP;
    const bool act = q < K;
    double w[RPT];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int row = base + t * RB;
        w[t] = (act && row < n) ? B[(size_t)row * K + q] : 0.0;
    }
    bdots<double, RPT>(0, -1, nullptr, nullptr, 0, n, KP, q, base, RB, act, w, part, G, 0);
}

__global__ void bk_set_bnorm(int KP, int K, const double* __restrict__ part, int G,
                             BCol* __restrict__ cols)
{
    const int q = lblock() * (BB_BS / 64) + ((int)threadIdx.x >> 6);
    if (q >= K) return;
    const double s = wave_sum_partials(part + (size_t)q * G, G);
    if ((threadIdx.x & 63) == 0) {
        BCol* c = cols + q;
        const double bn = sqrt(s);
        c->bnorm = bn > 0.0 ? bn : 1.0;
        c->iters = 0;
        c->cycles = 0;
        c->converged = 0;
        c->active = 0;
        c->k = 0;
    }
}

// Start of a restart cycle from the residual partials; writes the summary.
__global__ void bk_cycle_start(int KP, int K, const double* __restrict__ part, int G,
                               BCol* __restrict__ cols, BSum* __restrict__ sum, double rtol,
                               int max_iters)
{
    const int q = lblock() * (BB_BS / 64) + ((int)threadIdx.x >> 6);
    if (q >= K) return;
    BCol* c = cols + q;
    const bool done = c->converged;
    const double s = done ? 0.0 : wave_sum_partials(part + (size_t)q * G, G);
    if ((threadIdx.x & 63) != 0) return;
    if (!done) {
        const double beta = sqrt(s);
        // a whole restart cycle that lowered the true residual by less than
        // 1 % means it sits at its rounding floor above rtol: the column
        // stops there (not converged)
        const bool stalled = c->cycles > 0 && c->iters > c->iters0 && beta > 0.99 * c->beta;
        c->beta = beta;
        c->rel = beta / c->bnorm;
        c->k = 0;
        c->iters0 = c->iters;
        if (c->rel <= rtol || beta == 0.0) {
            c->converged = 1;
            c->active = 0;
        } else if (c->iters >= max_iters || stalled || !isfinite(beta)) {
            c->active = 0;
        } else {
            c->active = 1;
            c->carry[0] = beta;
            c->cycles += 1;
        }
    }
    BSum o;
    o.rel = c->rel;
    o.active = c->active;
    o.converged = c->converged;
    o.iters = c->iters;
    o.cycles = c->cycles;
    sum[q] = o;
}

__global__ void bk_fill0(size_t n, double* __restrict__ x)
{
    const size_t i = (size_t)lblock() * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// ---------------------------------------------------------------- driver --
// The preconditioner works on FP64 vectors (the form it is used with).
template <typename VT>
void precond_apply(BatchPrecond* pc, int KP, const VT* v, VT* z, hipStream_t s);
template <>
void precond_apply<double>(BatchPrecond* pc, int KP, const double* v, double* z, hipStream_t s)
{
    pc->apply(KP, v, z, s);
}
template <>
void precond_apply<float>(BatchPrecond* pc, int KP, const float* v, float* z, hipStream_t s)
{
    pc->apply_f32(KP, v, z, s);
}

// The fused first-pass kernel (default on; PHG_BATCH_FUSED=0 selects the
// two-read form, which gives the same bits).
bool fused_update_dots()
{
    static const bool on = [] {
        const char* e = getenv("PHG_BATCH_FUSED");
        return !(e && e[0] == '0');
    }();
    return on;
}

int next_pow2(int k)
{
    int p = 1;
    while (p < k) p <<= 1;
    return p;
}

int blocks_for(int n, int kp, int rpt)
{
    const int rows_per_block = (BB_BS / kp) * rpt;
    return std::max(1, (n + rows_per_block - 1) / rows_per_block);
}

int prepare_batch(phg_matrix_s* A, int slot, int prec, int m, int kp, int rpt, bool pc,
                  BatchWork** out)
{
    BatchWork*& w = A->batch[slot];
    const int n = A->n_rows;
    if (!w || w->m < m || w->kp != kp || w->prec != prec) {
        batch_destroy(w);
        w = nullptr;
        std::unique_ptr<BatchWork> b(new BatchWork());
        b->n = n;
        b->m = m;
        b->kp = kp;
        b->prec = prec;
        const size_t es = prec == PHG_F64 ? 8 : 4;
        const size_t bs = (size_t)n * kp;
        PHG_HIP(b->V.alloc((size_t)(m + 1) * bs * es));
        PHG_HIP(b->W.alloc(2 * bs * es));
        // sized for the smallest rows-per-thread (most blocks)
        PHG_HIP(b->part.alloc((size_t)(m + 2) * kp * blocks_for(n, kp, 1)));
        PHG_HIP(b->cols.alloc(kp));
        PHG_HIP(b->sum.alloc(kp));
        PHG_HIP(hipHostMalloc((void**)&b->host, sizeof(BSum) * kp, hipHostMallocDefault));
        PHG_HIP(b->nact.alloc(1));
        PHG_HIP(hipHostMalloc((void**)&b->hnact, sizeof(int), hipHostMallocDefault));
        w = b.release();
    }
    if (w->rpt != rpt) {  // another launch shape: new block count, new graph
        w->rpt = rpt;
        w->G = blocks_for(n, kp, rpt);
        w->drop_graphs();
    }
    if (pc && !w->Z.p) PHG_HIP(w->Z.alloc((size_t)w->m * w->n * w->kp * (prec == PHG_F64 ? 8 : 4)));
    *out = w;
    return PHG_OK;
}

// Number of columns still running in the cycle, for the host's early exit.
__global__ void bk_count_active(int K, const BCol* __restrict__ cols, int* __restrict__ out)
{
    const int lane = threadIdx.x;
    int a = 0;
    for (int q = lane; q < K; q += 64) a += cols[q].active ? 1 : 0;
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (lane == 0) out[0] = a;
}

// Arnoldi steps j0 .. j1-1 of a cycle of m steps (chunk c: steps up to
// bb_chunk_end(c)); a chunk that is not the cycle's last ends by counting
// the columns still running.
template <typename VT, typename MT, int RPT>
void enqueue_batch_steps(phg_matrix_s* A, const Ell<MT>& E, BatchWork* w, int K, int j0, int j1,
                         int m, int passes, double rtol, double inner_rtol, int max_iters,
                         BatchPrecond* pc, hipStream_t s)
{
    (void)A;
    const int n = w->n, KP = w->kp, G = w->G;
    const size_t bs = (size_t)n * KP;
    VT* V = (VT*)w->V.p;
    VT* W0 = (VT*)w->W.p;
    VT* W1 = W0 + bs;
    double* part = w->part.p;
    BCol* cols = w->cols.p;
    const dim3 grid(G), block(BB_BS);
    const int cb = (K + 3) / 4;  // blocks of one wavefront per column
    for (int j = j0; j < j1; ++j) {
        VT* src = (j & 1) ? W1 : W0;
        VT* dst = (j & 1) ? W0 : W1;
        if (pc) {
            // right preconditioning (flexible form): z_j = M^-1 v_j kept for the update
            VT* Vj = V + (size_t)j * bs;
            VT* Zj = (VT*)w->Z.p + (size_t)j * bs;
            hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_scale<VT>), dim3((unsigned)((bs + BB_BS - 1) / BB_BS)),
                               block, 0, s, n, KP, K, src, Vj, j, cols);
            precond_apply<VT>(pc, KP, Vj, Zj, s);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_step_spmm_pc<VT, MT, RPT>), grid, block, 0, s, n, KP,
                               K, E.width, E.col.p, E.val.p, Zj, V, bs, dst, j, part, G, cols);
        } else {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_step_spmm<VT, MT, RPT>), grid, block, 0, s, n, KP, K,
                               E.width, E.col.p, E.val.p, src, V, bs, dst, j, part, G, cols);
        }
        for (int p = 0; p < passes; ++p) {
            const int nd = j + 1;
            hipLaunchKernelGGL(bk_reduce, dim3((nd * KP + 3) / 4), block, 0, s, KP, K, nd, j, p,
                               part, G, cols);
            if (p < passes - 1 && fused_update_dots())
                hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_update_dots<VT, RPT>), grid, block, 0, s, n,
                                   KP, K, dst, V, bs, j, part, G, cols);
            else
                hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_update<VT, RPT>), grid, block, 0, s, n, KP,
                                   K, dst, V, bs, j, p == passes - 1 ? 1 : 0, part, G, cols);
        }
        hipLaunchKernelGGL(bk_givens, dim3(cb), block, 0, s, KP, K, j, m, part, G, cols, rtol,
                           inner_rtol, max_iters);
    }
    if (j1 < m) hipLaunchKernelGGL(bk_count_active, dim3(1), dim3(64), 0, s, K, cols, w->nact.p);
}

// The end of a cycle: y, the x update, the true residual and the next
// cycle's start (the host summary).
template <typename VT, typename MT, int RPT>
void enqueue_batch_finish(phg_matrix_s* A, const Ell<MT>& E, BatchWork* w, int K,
                          const double* B, double* X, double rtol, int max_iters,
                          BatchPrecond* pc, hipStream_t s)
{
    const int n = w->n, KP = w->kp, G = w->G;
    const size_t bs = (size_t)n * KP;
    VT* V = (VT*)w->V.p;
    VT* W0 = (VT*)w->W.p;
    double* part = w->part.p;
    BCol* cols = w->cols.p;
    const dim3 grid(G), block(BB_BS);
    const int cb = (K + 3) / 4;
    hipLaunchKernelGGL(bk_solve_y, dim3(cb), block, 0, s, K, cols);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_xupdate<VT, RPT>), grid, block, 0, s, n, KP, K,
                       pc ? (const VT*)w->Z.p : (const VT*)V, bs, X, cols,
                       (!pc && (const void*)E.val.p == (const void*)A->ell64d.val.p)
                           ? A->dinv.p : (const double*)nullptr);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_residual<VT, RPT>), grid, block, 0, s, n, KP, K,
                       A->ell64.width, A->ell64.col.p, A->ell64.val.p, B, X, W0, part, G, cols);
    hipLaunchKernelGGL(bk_cycle_start, dim3(cb), block, 0, s, KP, K, part, G, cols, w->sum.p,
                       rtol, max_iters);
}

template <typename VT, typename MT, int RPT>
int batch_run_t(phg_matrix_s* A, const Ell<MT>& E, BatchWork* w, int K, const double* B, double* X,
              const phg_gmres_opts& o, phg_gmres_info* infos, BatchPrecond* pc, hipStream_t s)
{
    const int n = w->n, KP = w->kp, G = w->G, m = o.restart;
    const size_t bs = (size_t)n * KP;
    VT* W0 = (VT*)w->W.p;
    double* part = w->part.p;
    BCol* cols = w->cols.p;
    const dim3 grid(G), block(BB_BS);
    const int cb = (K + 3) / 4;
    const int passes = o.ortho == PHG_CGS1 ? 1 : 2;
    const double inner = o.precision == PHG_F64 ? 0.0 : o.inner_rtol;
    (void)bs;

    if (!o.use_x0) {
        const size_t tot = (size_t)n * K;
        hipLaunchKernelGGL(bk_fill0, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, tot, X);
    }
    hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_sumsq<RPT>), grid, block, 0, s, n, KP, K, B, part, G);
    hipLaunchKernelGGL(bk_set_bnorm, dim3(cb), block, 0, s, KP, K, part, G, cols);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(bk_residual<VT, RPT>), grid, block, 0, s, n, KP, K,
                       A->ell64.width, A->ell64.col.p, A->ell64.val.p, B, X, W0, part, G, cols);
    hipLaunchKernelGGL(bk_cycle_start, dim3(cb), block, 0, s, KP, K, part, G, cols, w->sum.p,
                       o.rtol, o.max_iters);
    PHG_HIP(hipGetLastError());

    int nch = 0;  // chunks of the cycle
    while (nch < BB_NCHUNK && (nch == 0 || bb_chunk_end(nch - 1) < m)) ++nch;
    auto chunk_lo = [&](int c) { return c == 0 ? 0 : std::min(m, bb_chunk_end(c - 1)); };
    auto chunk_hi = [&](int c) { return c == nch - 1 ? m : std::min(m, bb_chunk_end(c)); };
    if (o.graph) {
        const void* key[4] = {B, X, E.val.p, A->ell64.val.p};
        const bool same = w->captured && w->key_m == m && w->key_ortho == o.ortho &&
                          w->key_k == K && w->key_rtol == o.rtol && w->key_inner == inner &&
                          w->key_maxit == o.max_iters && w->key_pc == (const void*)pc &&
                          !memcmp(key, w->key, sizeof key);
        if (!same) {
            w->drop_graphs();
            // recorded on the caller's stream, or for the null stream (which
            // cannot be captured) on a stream of the workspace
            hipStream_t cs = s;
            if (!cs) {
                if (!w->cap) PHG_HIP(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
                cs = w->cap;
            }
            for (int c = 0; c <= nch; ++c) {
                hipGraph_t g = nullptr;
                PHG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
                if (c < nch)
                    enqueue_batch_steps<VT, MT, RPT>(A, E, w, K, chunk_lo(c), chunk_hi(c), m,
                                                     passes, o.rtol, inner, o.max_iters, pc,
                                                     cs);
                else
                    enqueue_batch_finish<VT, MT, RPT>(A, E, w, K, B, X, o.rtol, o.max_iters, pc,
                                                      cs);
                PHG_HIP(hipStreamEndCapture(cs, &g));
                hipError_t e = hipGraphInstantiate(&w->exec[c], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                PHG_HIP(e);
            }
            w->captured = true;
            memcpy(w->key, key, sizeof key);
            w->key_m = m;
            w->key_ortho = o.ortho;
            w->key_k = K;
            w->key_rtol = o.rtol;
            w->key_inner = inner;
            w->key_maxit = o.max_iters;
            w->key_pc = pc;
        }
    }
    for (;;) {
        PHG_HIP(hipMemcpyAsync(w->host, w->sum.p, sizeof(BSum) * K, hipMemcpyDeviceToHost, s));
        PHG_HIP(hipStreamSynchronize(s));
        bool any = false;
        for (int q = 0; q < K; ++q) any = any || w->host[q].active;
        if (!any) break;
        for (int c = 0; c < nch; ++c) {
            if (o.graph) {
                PHG_HIP(hipGraphLaunch(w->exec[c], s));
            } else {
                enqueue_batch_steps<VT, MT, RPT>(A, E, w, K, chunk_lo(c), chunk_hi(c), m, passes,
                                                 o.rtol, inner, o.max_iters, pc, s);
                PHG_HIP(hipGetLastError());
            }
            if (c < nch - 1) {
                PHG_HIP(hipMemcpyAsync(w->hnact, w->nact.p, sizeof(int), hipMemcpyDeviceToHost, s));
                PHG_HIP(hipStreamSynchronize(s));
                if (*w->hnact == 0) break;
            }
        }
        if (o.graph) {
            PHG_HIP(hipGraphLaunch(w->exec[nch], s));
        } else {
            enqueue_batch_finish<VT, MT, RPT>(A, E, w, K, B, X, o.rtol, o.max_iters, pc, s);
            PHG_HIP(hipGetLastError());
        }
    }
    bool all = true;
    for (int q = 0; q < K; ++q) {
        const BSum& h = w->host[q];
        all = all && h.converged;
        if (infos) {
            memset(&infos[q], 0, sizeof infos[q]);
            infos[q].iterations = h.iters;
            infos[q].restarts = h.cycles;
            infos[q].refinements = o.precision == PHG_F64 ? 0 : h.cycles;
            infos[q].rel_residual = h.rel;
            infos[q].converged = h.converged;
        }
    }
    return all ? PHG_OK : PHG_ERR_NOCONV;
}

template <typename VT, typename MT>
int batch_run(phg_matrix_s* A, const Ell<MT>& E, BatchWork* w, int K, const double* B, double* X,
              const phg_gmres_opts& o, phg_gmres_info* infos, BatchPrecond* pc, hipStream_t s)
{
    switch (w->rpt) {
    case 1: return batch_run_t<VT, MT, 1>(A, E, w, K, B, X, o, infos, pc, s);
    case 2: return batch_run_t<VT, MT, 2>(A, E, w, K, B, X, o, infos, pc, s);
    case 4: return batch_run_t<VT, MT, 4>(A, E, w, K, B, X, o, infos, pc, s);
    case 8: return batch_run_t<VT, MT, 8>(A, E, w, K, B, X, o, infos, pc, s);
    default: return batch_run_t<VT, MT, BB_RPT>(A, E, w, K, B, X, o, infos, pc, s);
    }
}

// ------------------------------------------------------- launch shapes --
// Rows per thread of the row kernels for an (n, KP, precision) shape.  A
// shape's value fixes the block partials of every dot product, so it is kept
// per process: every matrix of that shape uses the same value and a column's
// result does not depend on which object solved it.  Untuned shapes take
// the largest value that still gives >= 1024 workgroups (4 per CU), or 1.
struct ShapeKey {
    int n, kp, prec;
    bool operator<(const ShapeKey& o) const
    {
        return n != o.n ? n < o.n : kp != o.kp ? kp < o.kp : prec < o.prec;
    }
};
std::mutex g_shape_mu;
std::map<ShapeKey, int> g_shape_rpt;

int default_rpt(int n, int kp)
{
    for (int r = 8; r > 1; r >>= 1)
        if (blocks_for(n, kp, r) >= 1024) return r;
    return 1;
}

int shape_rpt(int n, int kp, int prec)
{
    std::lock_guard<std::mutex> g(g_shape_mu);
    auto it = g_shape_rpt.find({n, kp, prec});
    return it == g_shape_rpt.end() ? default_rpt(n, kp) : it->second;
}

void set_shape_rpt(int n, int kp, int prec, int rpt)
{
    std::lock_guard<std::mutex> g(g_shape_mu);
    g_shape_rpt[{n, kp, prec}] = rpt;
}

// Deterministic right-hand sides for timing: values in [-1, 1) from a hash.
__global__ void bk_fill_hash(size_t n, double* __restrict__ x)
{
    const size_t i = (size_t)lblock() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long h = (i + 1) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    x[i] = (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

}  // namespace

int gmres_batch_solve(phg_matrix_s* A, int K, const double* B, double* X,
                      const phg_gmres_opts* opts, phg_gmres_info* infos, hipStream_t s,
                      phg_mg_s* mg)
{
    return gmres_batch_solve_pc(A, K, B, X, opts, infos, s, mg ? mg_precond(mg) : nullptr);
}

int gmres_batch_solve_pc(phg_matrix_s* A, int K, const double* B, double* X,
                         const phg_gmres_opts* opts, phg_gmres_info* infos, hipStream_t s,
                         BatchPrecond* pc)
{
    phg_gmres_opts o = *opts;
    if (K < 1 || K > BB_MAXK) {
        set_error("gmres_batch: nrhs must be in 1..64");
        return PHG_ERR_ARG;
    }
    if (o.restart < 1 || o.restart > BB_MAXM) {
        set_error("gmres_batch: restart must be in 1..32");
        return PHG_ERR_ARG;
    }
    if (o.ortho != PHG_CGS2 && o.ortho != PHG_CGS1) {
        set_error("gmres_batch: ortho must be PHG_CGS2 or PHG_CGS1");
        return PHG_ERR_ARG;
    }
    if (o.precision < PHG_F64 || o.precision > PHG_BF16) {
        set_error("gmres_batch: bad precision");
        return PHG_ERR_ARG;
    }
    if (A->n_rows != A->n_cols) {
        set_error("gmres_batch: matrix is not square");
        return PHG_ERR_ARG;
    }
    if (A->n_rows == 0) {
        if (infos)
            for (int q = 0; q < K; ++q) memset(&infos[q], 0, sizeof infos[q]), infos[q].converged = 1;
        return PHG_OK;
    }
    int rc = ensure_ell64(A, s);
    if (rc) return rc;
    const int kp = next_pow2(K);
    if (pc) {
        if ((o.precision != PHG_F64 && !pc->f32()) || pc->rows() != A->n_rows) {
            set_error("gmres_batch: the preconditioner does not fit the matrix or the precision");
            return PHG_ERR_ARG;
        }
        rc = o.precision == PHG_F64 ? pc->reserve(kp) : pc->reserve_f32(kp);
        if (rc) return rc;
    }
    BatchWork* w = nullptr;
    const int slot = o.precision == PHG_F64 ? 0 : 1;
    // the autotuner times the plain form, so a preconditioned solve keeps the
    // untuned default shape: its speed and bits do not depend on which
    // tunings ran earlier in the process
    const int rpt = pc ? default_rpt(A->n_rows, kp) : shape_rpt(A->n_rows, kp, o.precision);
    rc = prepare_batch(A, slot, o.precision == PHG_F64 ? PHG_F64 : PHG_F32, o.restart, kp, rpt,
                       pc != nullptr, &w);
    if (rc) return rc;
    if (o.precision == PHG_F64) {
        if (o.diag_scale && !pc) {  // right preconditioning by D^-1: solve (A D^-1) y = b, x = D^-1 y
            rc = ensure_ell64_diag(A, s);
            if (rc) return rc;
            return batch_run<double, double>(A, A->ell64d, w, K, B, X, o, infos, nullptr, s);
        }
        return batch_run<double, double>(A, A->ell64, w, K, B, X, o, infos, pc, s);
    }
    if (o.precision == PHG_F32) {
        rc = ensure_ell32(A, s);
        if (rc) return rc;
        return batch_run<float, float>(A, A->ell32, w, K, B, X, o, infos, pc, s);
    }
    rc = ensure_ell16(A, s);
    if (rc) return rc;
    return batch_run<float, uint16_t>(A, A->ell16, w, K, B, X, o, infos, pc, s);
}

int batch_shape_get(int n, int K, int prec)
{
    return shape_rpt(n, next_pow2(K), prec);
}

int batch_shape_set(int n, int K, int prec, int rpt)
{
    if (n < 1 || K < 1 || K > BB_MAXK || prec < PHG_F64 || prec > PHG_BF16 ||
        (rpt != 0 && rpt != 1 && rpt != 2 && rpt != 4 && rpt != 8 && rpt != 16)) {
        set_error("gmres_batch_set_shape: bad argument (rows per thread 0, 1, 2, 4, 8 or 16)");
        return PHG_ERR_ARG;
    }
    if (rpt == 0) {
        std::lock_guard<std::mutex> g(g_shape_mu);
        g_shape_rpt.erase({n, next_pow2(K), prec});
    } else {
        set_shape_rpt(n, next_pow2(K), prec, rpt);
    }
    return PHG_OK;
}

// Autotuner: times one restart cycle of `K` hash-valued right-hand sides for
// each rows-per-thread value and keeps the fastest for the shape.
int gmres_batch_tune(phg_matrix_s* A, int K, const phg_gmres_opts* opts, hipStream_t s,
                     int* best_out, double* ms_out)
{
    static const int cand[5] = {1, 2, 4, 8, 16};
    const int n = A->n_rows;
    if (K < 1 || K > BB_MAXK || n < 1) {
        set_error("gmres_batch_tune: nrhs must be in 1..64 and the matrix non-empty");
        return PHG_ERR_ARG;
    }
    phg_gmres_opts o = *opts;
    o.max_iters = o.restart;   // exactly one cycle of `restart` steps per column
    o.rtol = 0.0;
    o.inner_rtol = 0.0;
    o.use_x0 = 0;
    o.graph = 1;
    DevBuf<double> B, X;
    PHG_HIP(B.alloc((size_t)n * K));
    PHG_HIP(X.alloc((size_t)n * K));
    const size_t tot = (size_t)n * K;
    hipLaunchKernelGGL(bk_fill_hash, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, tot,
                       B.p);
    PHG_HIP(hipGetLastError());
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PHG_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        PHG_HIP(he);
    }
    const int kp = next_pow2(K);
    int best = BB_RPT, rc = PHG_OK;
    float best_ms = 0.0f;
    for (int c = 0; c < 5 && rc == PHG_OK; ++c) {
        set_shape_rpt(n, kp, o.precision, cand[c]);
        float t = 0.0f;
        for (int rep = 0; rep < 4 && rc == PHG_OK; ++rep) {  // rep 0 captures the graph
            if (hipEventRecord(e0, s) != hipSuccess) rc = PHG_ERR_HIP;
            int r = gmres_batch_solve(A, K, B.p, X.p, &o, nullptr, s);
            if (r != PHG_OK && r != PHG_ERR_NOCONV) rc = r;
            if (rc == PHG_OK && (hipEventRecord(e1, s) != hipSuccess ||
                                 hipEventSynchronize(e1) != hipSuccess))
                rc = PHG_ERR_HIP;
            float ms = 0.0f;
            if (rc == PHG_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PHG_ERR_HIP;
            if (rc == PHG_OK && rep > 0 && (rep == 1 || ms < t)) t = ms;
        }
        if (rc == PHG_OK) {
            if (ms_out) ms_out[c] = t;
            if (c == 0 || t < best_ms) {
                best_ms = t;
                best = cand[c];
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == PHG_ERR_HIP) set_error("gmres_batch_tune: HIP event error");
    set_shape_rpt(n, kp, o.precision, rc == PHG_OK ? best : default_rpt(n, kp));
    if (best_out) *best_out = best;
    return rc;
}

}  // namespace phg
