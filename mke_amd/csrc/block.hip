// Batched form: 64 right-hand sides against one shared matrix, solved by
// block GMRES(m) with a fused SpMM and MFMA block Gram-Schmidt.
//
// Block Arnoldi with block size p = 64 (Saad, Iterative Methods for Sparse
// Linear Systems, 2nd ed., sec. 6.12): the basis V = [V_0 .. V_m] is stored
// block-major in FP32, each block n x 64 row-major (an SpMM gather of a row
// of V_j and every GEMM tile read contiguous 256-byte rows).  Per step j:
//   W = A V_j                                  (SpMM, wave per row, lane per RHS)
//   two passes of block classical Gram-Schmidt (BCGS2):
//     H = [V_0..V_j]^T W   tall-skinny split-K GEMM on v_mfma_f32_32x32x2f32,
//                          partials summed in block order in FP64
//     W -= [V_0..V_j] H    row-tiled GEMM, V staged through LDS
//   CholQR2 of W: Gram W^T W accumulated in FP64 from exact FP32 products,
//     Cholesky and triangular inverse in FP64, V_{j+1} = W R^{-1} (MFMA)
//   the new block column of the block Hessenberg is reduced by Householder
//     reflectors (FP64, one wavefront per column) and the transformed
//     right-hand side gives every column's residual estimate.
// Each restart starts from the true FP64 residual R = B - A X (FP64 matrix),
// so the FP32 Krylov work sits under FP64 iterative refinement; X += V Y is
// accumulated in FP64.
#include <cmath>
#include <cstring>

#include "ell.h"
#include "phg_internal.h"

namespace phg {

constexpr int BP = 64;          // block size (RHS)
constexpr int B_MAXM = 32;      // max restart length
constexpr int B_LDSP = BP + 1;  // padded LDS row (bank-conflict free column reads)

typedef float f32x16 __attribute__((ext_vector_type(16)));

// Value of v in lane `lane` (wave-uniform index): two v_readlane_b32.
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

struct BlockWork {
    int n = 0, m = 0;
    size_t bs = 0;                // block stride of V = n * 64
    DevBuf<float> V, W, T;        // (m+1) blocks of n x 64, work, scratch
    DevBuf<double> R;             // n x 64 FP64 residual
    DevBuf<float> part;           // split-K partials, up to 2048 x 4096 (nb x chunks)
    DevBuf<double> gpart;         // Gram partials, up to 512 x 4096
    DevBuf<double> gsum;          // 64 x 64 Gram matrix
    DevBuf<double> Hbar;          // (m+1)p x mp, column-major (ld (m+1)p)
    DevBuf<double> U;             // reflectors: m x 64 x (p+1)
    DevBuf<double> tau;           // m x 64
    DevBuf<double> Tm;            // m x 64 x 64 compact-WY factors (row-major, upper)
    DevBuf<double> G;             // (m+1)p x p transformed RHS, column-major
    DevBuf<double> Hc;            // (m+1)p x p coefficients of one step (FP64)
    DevBuf<float> Hs;             // (m+1)p x p coefficients (FP32, row-major)
    DevBuf<double> Rm;            // 64 x 64 Cholesky factor (upper), row-major
    DevBuf<double> Racc;          // 64 x 64 accumulated R of CholQR2
    DevBuf<float> Rinv;           // 64 x 64 inverse (FP32, row-major)
    DevBuf<double> Y;             // mp x p solution of the projected problem
    DevBuf<float> Y32;
    DevBuf<double> colsq;         // per-column sums: partials (512 x 64) + result
    DevBuf<double> stat;          // [0..63] res estimates, [64..127] beta, [128..191] bnorm
    DevBuf<int> flags;            // [0] breakdown
    double* hstat = nullptr;      // pinned copy of stat
    ~BlockWork()
    {
        if (hstat) (void)hipHostFree(hstat);
    }
};

namespace {

// ------------------------------------------------------------------ SpMM --
// W[r][q] = sum_k a_rk V[c_k][off + q]: one wavefront per row, lane = RHS.
template <typename MT>
__global__ void __launch_bounds__(256)
k_spmm(int n, int width, const int* __restrict__ col, const MT* __restrict__ val,
       const float* __restrict__ X, size_t ldx, float* __restrict__ Y, size_t ldy)
{
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    int c[ELL_MAXW];
    float v[ELL_MAXW], xv[ELL_MAXW];
    const int w = width < ELL_MAXW ? width : ELL_MAXW;
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) {
        const size_t o = (size_t)k * n + row;
        c[k] = k < w ? col[o] : 0;
        v[k] = k < w ? ell_val_f32(val[o]) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < w ? X[(size_t)c[k] * ldx + lane] : 0.0f;
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < ELL_MAXW; ++k)
        if (k < w) s = fmaf(v[k], xv[k], s);
    for (int k = ELL_MAXW; k < width; ++k) {
        const size_t o = (size_t)k * n + row;
        s = fmaf(ell_val_f32(val[o]), X[(size_t)col[o] * ldx + lane], s);
    }
    Y[(size_t)row * ldy + lane] = s;
}

// R = B - A X in FP64 (FP64 values), per-block partial column sums of R^2.
__global__ void __launch_bounds__(256)
k_residual64(int n, int width, const int* __restrict__ col, const double* __restrict__ val,
             const double* __restrict__ B, const double* __restrict__ X, double* __restrict__ R,
             float* __restrict__ R32, double* __restrict__ colpart)
{
    __shared__ double sh[4][BP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc = 0.0;
    for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
        int c[ELL_MAXW];
        double v[ELL_MAXW], xv[ELL_MAXW];
        const int w = width < ELL_MAXW ? width : ELL_MAXW;
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = (size_t)k * n + row;
            c[k] = k < w ? col[o] : 0;
            v[k] = k < w ? val[o] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = k < w ? X[(size_t)c[k] * BP + lane] : 0.0;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k)
            if (k < w) s += v[k] * xv[k];
        for (int k = ELL_MAXW; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            s += val[o] * X[(size_t)col[o] * BP + lane];
        }
        const double r = B[(size_t)row * BP + lane] - s;
        R[(size_t)row * BP + lane] = r;
        R32[(size_t)row * BP + lane] = (float)r;
        acc += r * r;
    }
    sh[wv][lane] = acc;
    __syncthreads();
    if (threadIdx.x < BP)
        colpart[(size_t)blockIdx.x * BP + threadIdx.x] =
            sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

__global__ void k_colsum_b(int n, const double* __restrict__ B, double* __restrict__ colpart)
{
    __shared__ double sh[4][BP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double acc = 0.0;
    for (int row = blockIdx.x * 4 + wv; row < n; row += gridDim.x * 4) {
        const double b = B[(size_t)row * BP + lane];
        acc += b * b;
    }
    sh[wv][lane] = acc;
    __syncthreads();
    if (threadIdx.x < BP)
        colpart[(size_t)blockIdx.x * BP + threadIdx.x] =
            sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// stat[off + q] = sqrt(sum_blocks colpart[block][q]), in block order.
__global__ void k_colnorm(int nblk, const double* __restrict__ colpart, double* __restrict__ stat,
                          int off)
{
    const int q = threadIdx.x;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += colpart[(size_t)b * BP + q];
    stat[off + q] = sqrt(s);
}

// ------------------------------------------------- split-K  H = V^T W ------
// Blocks ip, ip+1 of V against W over row chunk c; the 64 x 64 partials go
// to part[(i * nc + c) * 4096].  V is block-major (block i at i * bs, rows
// of 64 floats).  Each wavefront takes every fourth 2-row step of the chunk
// and accumulates both blocks' full 64 x 64 outputs (8 tiles of
// v_mfma_f32_32x32x2f32: A[a][k] = V_i[r+k][a], B[k][b] = W[r+k][b]; lane l
// supplies row r + l/32, column l%32 of a 32-wide half), so every loaded
// element feeds two MFMAs.  The four wavefronts are added in order in LDS.
__global__ void __launch_bounds__(256)
k_atb(int n, const float* __restrict__ V, size_t bs, const float* __restrict__ W, int nb,
      int nc, int chunk, float* __restrict__ part)
{
    __shared__ float red[BP * BP];
    const int c = blockIdx.x, ip = 2 * blockIdx.y;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kr = lane >> 5, cl = lane & 31;
    const int r0 = c * chunk, r1 = min(n, r0 + chunk);
    const bool two = ip + 1 < nb;
    const float* V0 = V + (size_t)ip * bs + cl;
    const float* V1 = V + (size_t)(two ? ip + 1 : ip) * bs + cl;
    const float* Wb = W + cl;
    f32x16 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = f32x16{0};
    for (int r = r0 + 2 * wv; r < r1; r += 32) {
        float a[4][4][2], b[4][2];  // [unroll][blk*2+half] / [unroll][half]
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int row = r + 8 * u + kr;
            const bool in = row < r1;
            const size_t o = (size_t)row * BP;
            b[u][0] = in ? Wb[o] : 0.0f;
            b[u][1] = in ? Wb[o + 32] : 0.0f;
            a[u][0][0] = in ? V0[o] : 0.0f;
            a[u][0][1] = in ? V0[o + 32] : 0.0f;
            a[u][1][0] = (in && two) ? V1[o] : 0.0f;
            a[u][1][1] = (in && two) ? V1[o + 32] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int bk = 0; bk < 2; ++bk)
#pragma unroll
                for (int ah = 0; ah < 2; ++ah)
#pragma unroll
                    for (int bh = 0; bh < 2; ++bh) {
                        const int t = bk * 4 + ah * 2 + bh;
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u][bk][ah], b[u][bh], acc[t],
                                                                      0, 0, 0);
                    }
    }
    for (int bk = 0; bk < (two ? 2 : 1); ++bk) {
        for (int w = 0; w < 4; ++w) {
            if (wv == w) {
#pragma unroll
                for (int ah = 0; ah < 2; ++ah)
#pragma unroll
                    for (int bh = 0; bh < 2; ++bh) {
                        const f32x16 v16 = acc[bk * 4 + ah * 2 + bh];
#pragma unroll
                        for (int v = 0; v < 16; ++v) {
                            const int ii = 32 * ah + 8 * (v >> 2) + 4 * kr + (v & 3);
                            const int jj = 32 * bh + cl;
                            red[ii * BP + jj] = (w == 0) ? v16[v] : red[ii * BP + jj] + v16[v];
                        }
                    }
            }
            __syncthreads();
        }
        float* out = part + ((size_t)(ip + bk) * nc + c) * (BP * BP);
        for (int e = threadIdx.x; e < BP * BP; e += 256) out[e] = red[e];
        __syncthreads();
    }
}

// Sum the split-K partials in chunk order (FP64): H(64i+a, b).  Pass 0
// stores, pass 1 adds; Hs gets the FP32 copy of this pass's coefficients.
__global__ void __launch_bounds__(256)
k_atb_reduce(int nb, int nc, const float* __restrict__ part, double* __restrict__ Hc,
             float* __restrict__ Hs, int pass)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * BP * BP) return;
    const int i = e / (BP * BP), ab = e % (BP * BP);
    const float* p = part + (size_t)i * nc * (BP * BP) + ab;
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= nc; c += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[(size_t)(c + u) * (BP * BP)];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += (double)v[u];
    }
    for (; c < nc; ++c) s += (double)p[(size_t)c * (BP * BP)];
    const int a = ab / BP, b = ab % BP;
    Hs[(size_t)(i * BP + a) * BP + b] = (float)s;
    if (pass == 0) Hc[(size_t)(i * BP + a) * BP + b] = s;
    else Hc[(size_t)(i * BP + a) * BP + b] += s;
}

// -------------------------------------------- row-tiled  out = f(V H) ------
// acc(r, b) = sum_{k < K} V[r][k] * H[k][b] for a 128-row tile; V (block
// k/64 of a block-major basis, stride bs) is staged through LDS 64 values of
// k at a time, the stage's H fragments (rows k, 64 wide, shared by every
// tile and resident in L2) are loaded into registers.  Wavefront w owns rows
// 32w..32w+31 and both 32-column halves: per k step one A and two B
// fragments feed two v_mfma_f32_32x32x2f32.  Epilogues: 0: Out -= acc, 1: Out = acc,
// 2: X64 += acc.
template <int MODE>
__global__ void __launch_bounds__(256)
k_vh(int n, const float* __restrict__ V, size_t bs, int K, const float* __restrict__ H,
     float* __restrict__ Out, double* __restrict__ X64)
{
    __shared__ float tile[128 * B_LDSP];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int kr = lane >> 5, cl = lane & 31;
    const int row0 = blockIdx.x * 128;
    f32x16 acc0 = {0}, acc1 = {0};
    // registers of the next stage: 32 V values (rows wv, wv+4, ..) and the
    // 64 H fragments; loaded while the current stage runs its MFMAs
    float pv[32], hb0[32], hb1[32];
    auto load_stage = [&](int k0) {
        const float* Vk = V + (size_t)(k0 / 64) * bs;
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int row = row0 + wv + 4 * u;
            pv[u] = row < n ? Vk[(size_t)row * BP + lane] : 0.0f;
        }
        const float* Hk = H + (size_t)(k0 + kr) * BP + cl;
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            hb0[u] = Hk[(size_t)(2 * u) * BP];
            hb1[u] = Hk[(size_t)(2 * u) * BP + 32];
        }
    };
    load_stage(0);
    for (int k0 = 0; k0 < K; k0 += 64) {
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 32; ++u) tile[(wv + 4 * u) * B_LDSP + lane] = pv[u];
        float cb0[32], cb1[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            cb0[u] = hb0[u];
            cb1[u] = hb1[u];
        }
        __syncthreads();
        if (k0 + 64 < K) load_stage(k0 + 64);
        const float* trow = tile + (32 * wv + cl) * B_LDSP + kr;
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const float a = trow[2 * u];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, cb0[u], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, cb1[u], acc1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = row0 + 32 * wv + 8 * (v >> 2) + 4 * kr + (v & 3);
        if (row < n) {
            const size_t o = (size_t)row * BP + cl;
            if (MODE == 0) {
                Out[o] = Out[o] - acc0[v];
                Out[o + 32] = Out[o + 32] - acc1[v];
            } else if (MODE == 1) {
                Out[o] = acc0[v];
                Out[o + 32] = acc1[v];
            } else {
                X64[o] += (double)acc0[v];
                X64[o + 32] += (double)acc1[v];
            }
        }
    }
}

// ----------------------------------------------- Gram  G = W^T W (FP64) ----
// Products of FP32 values are exact in FP64; sums in FP64.  Each block takes
// a row chunk staged through LDS 32 rows at a time; thread t accumulates the
// 4 x 4 block G[4(t/16) .. +3][4(t%16) .. +3].
__global__ void __launch_bounds__(256)
k_gram(int n, const float* __restrict__ W, int chunk, double* __restrict__ gpart)
{
    __shared__ float rows[32][BP];
    const int ai = 4 * (threadIdx.x >> 4), bj = 4 * (threadIdx.x & 15);
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    const int r0 = blockIdx.x * chunk, r1 = min(n, r0 + chunk);
    for (int rb = r0; rb < r1; rb += 32) {
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * BP; e += 256) {
            const int rr = e / BP, q = e % BP;
            rows[rr][q] = (rb + rr < r1) ? W[(size_t)(rb + rr) * BP + q] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int rr = 0; rr < 32; ++rr) {
            double xa[4], xb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xa[u] = (double)rows[rr][ai + u];
                xb[u] = (double)rows[rr][bj + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] += xa[u] * xb[v];
        }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            gpart[(size_t)blockIdx.x * BP * BP + (ai + u) * BP + bj + v] = acc[u][v];
}

// G = sum of the Gram partials in chunk order, one thread per entry; the
// loads of 8 partials are issued together.
__global__ void __launch_bounds__(256)
k_gram_reduce(int ngp, const double* __restrict__ gpart, double* __restrict__ gsum)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= BP * BP) return;
    double s = 0.0;
    int c = 0;
    for (; c + 8 <= ngp; c += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = gpart[(size_t)(c + u) * BP * BP + e];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < ngp; ++c) s += gpart[(size_t)c * BP * BP + e];
    gsum[e] = s;
}

// Cholesky of the Gram matrix G (one block of 256 threads, FP64):
// G = R^T R, R upper; Rinv = R^{-1} (FP32 copy for the scaling GEMM);
// Racc = R * Racc (mode 1) or Racc = R (mode 0).  A non-positive pivot sets
// flags[0] and is replaced by a small positive value.
__global__ void __launch_bounds__(256)
k_chol(const double* __restrict__ gsum, double* __restrict__ Racc, float* __restrict__ Rinv32,
       int mode, int* __restrict__ flags)
{
    __shared__ double A[BP][BP + 1];
    __shared__ double Ri[BP][BP + 1];
    __shared__ double Rold[BP][BP + 1];
    __shared__ double scale;
    const int t = threadIdx.x;
    for (int e = t; e < BP * BP; e += 256) A[e / BP][e % BP] = gsum[e];
    __syncthreads();
    if (t == 0) {
        double sc = 0.0;
        for (int q = 0; q < BP; ++q) sc = fmax(sc, A[q][q]);
        scale = sc;
    }
    __syncthreads();
    // right-looking Cholesky on the upper triangle
    for (int k = 0; k < BP; ++k) {
        if (t == 0) {
            double d = A[k][k];
            if (!(d > 1e-30 * scale)) {
                flags[0] = 1;
                d = fmax(1e-30 * scale, 1e-300);
            }
            A[k][k] = sqrt(d);
        }
        __syncthreads();
        if (t > k && t < BP) A[k][t] = A[k][t] / A[k][k];
        __syncthreads();
        const int rem = BP - 1 - k;
        for (int e = t; e < rem * rem; e += 256) {
            const int r = k + 1 + e / rem, c = k + 1 + e % rem;
            if (c >= r) A[r][c] -= A[k][r] * A[k][c];
        }
        __syncthreads();
    }
    // inverse of the upper factor, column t by back substitution
    if (t < BP) {
        for (int i = 0; i < BP; ++i) Ri[i][t] = 0.0;
        for (int i = t; i >= 0; --i) {
            double sacc = (i == t) ? 1.0 : 0.0;
            for (int c = i + 1; c <= t; ++c) sacc -= A[i][c] * Ri[c][t];
            Ri[i][t] = sacc / A[i][i];
        }
    }
    if (mode == 1)
        for (int e = t; e < BP * BP; e += 256) Rold[e / BP][e % BP] = Racc[e];
    __syncthreads();
    for (int e = t; e < BP * BP; e += 256) {
        const int i = e / BP, c = e % BP;
        Rinv32[e] = (float)Ri[i][c];
        if (mode == 0) {
            Racc[e] = (i <= c) ? A[i][c] : 0.0;
        } else {
            double sacc = 0.0;
            for (int q = i; q <= c; ++q) sacc += A[i][q] * Rold[q][c];
            Racc[e] = sacc;
        }
    }
}

// ------------------------------------- projected least squares (FP64) -----
// Hbar is column-major with leading dimension ld = (m+1)p.  Column q of
// block column j holds H(0..(j+2)p-1, jp+q).

// Write block column j of Hbar from the BCGS coefficients Hc (rows
// 0..(j+1)p-1, row-major p wide) and the CholQR factor Racc (rows
// (j+1)p..(j+2)p-1).
__global__ void __launch_bounds__(256)
k_put_column(int j, int ld, const double* __restrict__ Hc, const double* __restrict__ Racc,
             double* __restrict__ Hbar)
{
    const int rows = (j + 2) * BP;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * BP; e += gridDim.x * blockDim.x) {
        const int r = e / BP, q = e % BP;
        double v;
        if (r < (j + 1) * BP) v = Hc[(size_t)r * BP + q];
        else v = Racc[(r - (j + 1) * BP) * BP + q];
        Hbar[(size_t)(j * BP + q) * ld + r] = v;
    }
}

// Reflector c of block column k acts on rows kp+c .. (k+1)p+c (length p+1):
// u_c = U[k][c][0..64], u_c[0] = 1.  The p reflectors of a block are applied
// in compact WY form (Schreiber & Van Loan 1989): H_0 H_1 .. H_{p-1} =
// I - U T U^T with T upper triangular, so Q^T x = x - U (T^T (U^T x)), three
// lane-parallel products instead of p dependent reductions.
// A wavefront holds rows kp .. kp+127 of one column (x0 = row l, x1 = row
// 64 + l); u is the block's reflectors in LDS (64 x 65), tcol[d] = T[d][l].
__device__ __forceinline__ void apply_wy(double& x0, double& x1, const double* u,
                                         const double* tcol)
{
    const int l = threadIdx.x & 63;
    // z_l = sum_k u_l[k] x[l + k]
    double z = 0.0;
    for (int k = 0; k <= BP; ++k) {
        const int src = l + k;  // row index 0..127
        const double a = __shfl(x0, src & 63, 64);
        const double b = __shfl(x1, src & 63, 64);
        z += u[l * (BP + 1) + k] * (src < BP ? a : b);
    }
    // z'_l = sum_{d <= l} T[d][l] z_d
    double zp = 0.0;
#pragma unroll
    for (int d = 0; d < BP; ++d) {
        const double zd = readlane_d(z, d);
        if (d <= l) zp += tcol[d] * zd;
    }
    // x[r] -= sum_c u_c[r - c] z'_c
    for (int c = 0; c < BP; ++c) {
        const double zc = readlane_d(zp, c);
        const double* uc = u + c * (BP + 1);
        if (l >= c) x0 -= uc[l - c] * zc;
        if (l <= c) x1 -= uc[BP + l - c] * zc;
    }
}

// Apply all earlier reflector blocks k < j to block column j of Hbar: 16
// blocks of 4 wavefronts (LDS traffic spread over 16 CUs), one column each;
// block k's reflectors are staged in LDS, its T column per lane in registers.
__global__ void __launch_bounds__(256)
k_ls_apply(int j, int ld, double* __restrict__ Hbar, const double* __restrict__ U,
           const double* __restrict__ Tm)
{
    __shared__ double us[BP * (BP + 1)];
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    double* col = Hbar + (size_t)(j * BP + q) * ld;
    for (int k = 0; k < j; ++k) {
        __syncthreads();
        for (int e = threadIdx.x; e < BP * (BP + 1); e += 256) us[e] = U[(size_t)k * BP * (BP + 1) + e];
        double tcol[BP];
        const double* Tk = Tm + (size_t)k * BP * BP;
#pragma unroll
        for (int d = 0; d < BP; ++d) tcol[d] = Tk[d * BP + l];
        __syncthreads();
        double x0 = col[k * BP + l], x1 = col[k * BP + BP + l];
        apply_wy(x0, x1, us, tcol);
        col[k * BP + l] = x0;
        col[k * BP + BP + l] = x1;
    }
}

// T of block j's reflectors: T[c][c] = tau_c, T[0:c, c] = -tau_c T[0:c, 0:c]
// (U^T u_c)[0:c], with (U^T u_c)[d] = sum over the rows both reflectors cover.
// One block of 1024 threads.
__global__ void __launch_bounds__(1024)
k_ls_wy(int j, const double* __restrict__ U, const double* __restrict__ tau, double* __restrict__ Tm)
{
    __shared__ double us[BP * (BP + 1)];
    __shared__ double S[BP][BP + 1];
    __shared__ double T[BP][BP + 1];
    const int t = threadIdx.x;
    for (int e = t; e < BP * (BP + 1); e += 1024) us[e] = U[(size_t)j * BP * (BP + 1) + e];
    __syncthreads();
    // S[d][c] = u_d . u_c over rows c .. d + 64 (d < c); u_d[r - d], u_c[r - c]
    for (int e = t; e < BP * BP; e += 1024) {
        const int d = e / BP, c = e % BP;
        double sacc = 0.0;
        if (d < c)
            for (int r = c; r <= d + BP; ++r) sacc += us[d * (BP + 1) + (r - d)] * us[c * (BP + 1) + (r - c)];
        S[d][c] = sacc;
        T[d][c] = 0.0;
    }
    __syncthreads();
    for (int c = 0; c < BP; ++c) {
        const double tc = tau[(size_t)j * BP + c];
        if (t < c) {
            double sacc = 0.0;
            for (int e = t; e < c; ++e) sacc += T[t][e] * S[e][c];
            T[t][c] = -tc * sacc;
        }
        if (t == c) T[c][c] = tc;
        __syncthreads();
    }
    for (int e = t; e < BP * BP; e += 1024) Tm[(size_t)j * BP * BP + e] = T[e / BP][e % BP];
}

// Householder reflectors that triangularise rows jp..(j+2)p-1 of block
// column j (one block of 1024 threads: wave w owns columns 4w..4w+3).
__global__ void __launch_bounds__(1024)
k_ls_factor(int j, int ld, double* __restrict__ Hbar, double* __restrict__ U,
            double* __restrict__ tau)
{
    __shared__ double uc[BP + 1];
    __shared__ double tc;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double x0[4], x1[4];
    for (int s = 0; s < 4; ++s) {
        const double* col = Hbar + (size_t)(j * BP + 4 * w + s) * ld + j * BP;
        x0[s] = col[l];
        x1[s] = col[BP + l];
    }
    double* Uj = U + (size_t)j * BP * (BP + 1);
    for (int c = 0; c < BP; ++c) {
        const int ow = c >> 2, os = c & 3;
        if (w == ow) {
            // segment rows c .. c+64 of column c: x0[l] for l >= c, x1[l] for l <= c
            const double s0 = (l >= c) ? x0[os] : 0.0;
            const double s1 = (l <= c) ? x1[os] : 0.0;
            double nrm2 = s0 * s0 + s1 * s1;
            for (int off = 32; off > 0; off >>= 1) nrm2 += __shfl_xor(nrm2, off, 64);
            const double alpha = __shfl(x0[os], c, 64);
            const double nrm = sqrt(nrm2);
            const double beta = alpha >= 0.0 ? -nrm : nrm;
            double t = 0.0, v0 = 1.0;
            if (nrm2 > 0.0 && nrm > 0.0) {
                v0 = alpha - beta;
                t = (beta - alpha) / beta;
            }
            // u = [1, x(c+1..c+64) / v0]
            if (l >= c) uc[l - c] = (l == c) ? 1.0 : ((v0 != 0.0) ? s0 / v0 : 0.0);
            if (l <= c) uc[l + BP - c] = (v0 != 0.0) ? s1 / v0 : 0.0;
            if (l == 0) tc = t;
        }
        __syncthreads();
        const double tt = tc;
        // apply to columns >= c (column c becomes [beta, 0, ...])
        for (int s = 0; s < 4; ++s) {
            const int q = 4 * w + s;
            if (q < c) continue;
            const double u0 = (l >= c) ? uc[l - c] : 0.0;
            const double u1 = (l <= c) ? uc[l + BP - c] : 0.0;
            double d = u0 * x0[s] + u1 * x1[s];
            for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
            const double f = tt * d;
            x0[s] -= f * u0;
            x1[s] -= f * u1;
        }
        if (threadIdx.x <= BP) Uj[(size_t)c * (BP + 1) + threadIdx.x] = uc[threadIdx.x];
        if (threadIdx.x == 0) tau[(size_t)j * BP + c] = tt;
        __syncthreads();
    }
    for (int s = 0; s < 4; ++s) {
        double* col = Hbar + (size_t)(j * BP + 4 * w + s) * ld + j * BP;
        col[l] = x0[s];
        col[BP + l] = x1[s];
    }
}

// Apply reflector block j to the transformed right-hand side G (rows
// jp..(j+2)p-1 of each column) and record the residual estimate of column q:
// || G((j+1)p .. (j+2)p-1, q) ||.  16 blocks of 4 wavefronts.
__global__ void __launch_bounds__(256)
k_ls_rhs(int j, int ld, double* __restrict__ G, const double* __restrict__ U,
         const double* __restrict__ Tm, double* __restrict__ stat)
{
    __shared__ double us[BP * (BP + 1)];
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    for (int e = threadIdx.x; e < BP * (BP + 1); e += 256) us[e] = U[(size_t)j * BP * (BP + 1) + e];
    double tcol[BP];
    const double* Tj = Tm + (size_t)j * BP * BP;
#pragma unroll
    for (int d = 0; d < BP; ++d) tcol[d] = Tj[d * BP + l];
    __syncthreads();
    double* col = G + (size_t)q * ld + j * BP;
    double x0 = col[l], x1 = col[BP + l];
    apply_wy(x0, x1, us, tcol);
    col[l] = x0;
    col[BP + l] = x1;
    double s = x1 * x1;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (l == 0) stat[q] = sqrt(s);
}

// G = [S; 0] with S = Racc (upper, row-major) for a new cycle.
__global__ void k_ls_init_rhs(int rows, int ld, const double* __restrict__ Racc, double* __restrict__ G)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * BP; e += gridDim.x * blockDim.x) {
        const int r = e % rows, q = e / rows;
        G[(size_t)q * ld + r] = (r < BP) ? Racc[r * BP + q] : 0.0;
    }
}

// Y = R^{-1} G(0..Jp-1), R the upper triangular leading part of Hbar, by
// block rows of 64 from the bottom (Z = G initially, kept in Y's storage,
// row-major Jp x 64):
//   k_ls_copy_rhs:  Z = G
//   k_ls_diag(b):   Y_b = R_bb^{-1} Z_b      (one block, column-parallel)
//   k_ls_update(b): Z_c -= R_cb Y_b, c < b   (32-row tiles, R_cb staged in LDS)
__global__ void k_ls_copy_rhs(int K, int ld, const double* __restrict__ G, double* __restrict__ Y)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * BP) return;
    const int r = e / BP, q = e % BP;
    Y[e] = G[(size_t)q * ld + r];
}

__global__ void __launch_bounds__(1024)
k_ls_diag(int b, int ld, const double* __restrict__ Hbar, double* __restrict__ Y,
          float* __restrict__ Y32)
{
    __shared__ double Rb[BP][BP + 1];
    __shared__ double Zb[BP][BP + 1];
    const int t = threadIdx.x;
    for (int e = t; e < BP * BP; e += 1024) {
        const int r = e % BP, c = e / BP;  // column-major source: consecutive r
        Rb[r][c] = Hbar[(size_t)(b * BP + c) * ld + b * BP + r];
        Zb[e / BP][e % BP] = Y[(size_t)(b * BP) * BP + e];
    }
    __syncthreads();
    for (int i = BP - 1; i >= 0; --i) {
        if (t < BP) Zb[i][t] = Zb[i][t] / Rb[i][i];
        __syncthreads();
        for (int e = t; e < i * BP; e += 1024) {
            const int r = e / BP, q = e % BP;
            Zb[r][q] -= Rb[r][i] * Zb[i][q];
        }
        __syncthreads();
    }
    for (int e = t; e < BP * BP; e += 1024) {
        const double v = Zb[e / BP][e % BP];
        Y[(size_t)(b * BP) * BP + e] = v;
        Y32[(size_t)(b * BP) * BP + e] = (float)v;
    }
}

__global__ void __launch_bounds__(256)
k_ls_update(int b, int ld, const double* __restrict__ Hbar, double* __restrict__ Y)
{
    __shared__ double Rt[32][BP + 1];
    __shared__ double Yb[BP][BP + 1];
    const int r0 = blockIdx.x * 32, t = threadIdx.x;
    for (int e = t; e < 32 * BP; e += 256) {
        const int rr = e % 32, c = e / 32;  // column-major source: consecutive rows
        Rt[rr][c] = Hbar[(size_t)(b * BP + c) * ld + r0 + rr];
    }
    for (int e = t; e < BP * BP; e += 256) Yb[e / BP][e % BP] = Y[(size_t)(b * BP) * BP + e];
    __syncthreads();
    for (int e = t; e < 32 * BP; e += 256) {
        const int rr = e / BP, q = e % BP;
        double sacc = 0.0;
        for (int c = 0; c < BP; ++c) sacc += Rt[rr][c] * Yb[c][q];
        Y[(size_t)(r0 + rr) * BP + q] -= sacc;
    }
}

__global__ void k_fill0_d(size_t cnt, double* __restrict__ x)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < cnt) x[e] = 0.0;
}

}  // namespace

namespace {
int gram_chunks(int n, int* chunk)
{
    int ng = std::min(512, std::max(1, (n + 255) / 256));
    *chunk = (n + ng - 1) / ng;
    *chunk = (*chunk + 31) / 32 * 32;
    ng = (n + *chunk - 1) / *chunk;
    return ng;
}

// W (n x 64, FP32) -> orthonormal Q in V block `blk`, Racc = R with W = Q R
// (CholQR2: two Gram/Cholesky passes, the first into T).
int cholqr2(BlockWork* w, const float* Wsrc, int blk, hipStream_t s)
{
    const int n = w->n;
    int chunk = 0;
    const int ng = gram_chunks(n, &chunk);
    const dim3 tiles((n + 127) / 128);
    for (int pass = 0; pass < 2; ++pass) {
        const float* src = pass == 0 ? Wsrc : w->T.p;
        float* dst = pass == 0 ? w->T.p : w->V.p + (size_t)blk * w->bs;
        hipLaunchKernelGGL(k_gram, dim3(ng), dim3(256), 0, s, n, src, chunk, w->gpart.p);
        hipLaunchKernelGGL(k_gram_reduce, dim3(BP * BP / 256), dim3(256), 0, s, ng, w->gpart.p,
                           w->gsum.p);
        hipLaunchKernelGGL(k_chol, dim3(1), dim3(256), 0, s, w->gsum.p, w->Racc.p, w->Rinv.p, pass,
                           w->flags.p);
        // dst = src * Rinv   (K = 64: src is its own single block)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_vh<1>), tiles, dim3(256), 0, s, n, src, (size_t)0, BP,
                           w->Rinv.p, dst, nullptr);
    }
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}
}  // namespace

int block_prepare(phg_matrix_s* A, int m, BlockWork*& w)
{
    const int n = A->n_rows;
    if (w && w->m >= m && w->n == n) return PHG_OK;
    delete w;
    w = new BlockWork();
    w->n = n;
    w->m = m;
    w->bs = (size_t)n * BP;
    const size_t ldh = (size_t)(m + 1) * BP;
    PHG_HIP(w->V.alloc((size_t)(m + 1) * w->bs));
    PHG_HIP(w->W.alloc((size_t)n * BP));
    PHG_HIP(w->T.alloc((size_t)n * BP));
    PHG_HIP(w->R.alloc((size_t)n * BP));
    PHG_HIP(w->part.alloc((size_t)2048 * BP * BP));
    PHG_HIP(w->gpart.alloc((size_t)512 * BP * BP));
    PHG_HIP(w->gsum.alloc(BP * BP));
    PHG_HIP(w->Hbar.alloc(ldh * (size_t)m * BP));
    PHG_HIP(w->U.alloc((size_t)m * BP * (BP + 1)));
    PHG_HIP(w->tau.alloc((size_t)m * BP));
    PHG_HIP(w->Tm.alloc((size_t)m * BP * BP));
    PHG_HIP(w->G.alloc(ldh * BP));
    PHG_HIP(w->Hc.alloc(ldh * BP));
    PHG_HIP(w->Hs.alloc(ldh * BP));
    PHG_HIP(w->Racc.alloc(BP * BP));
    PHG_HIP(w->Rinv.alloc(BP * BP));
    PHG_HIP(w->Y.alloc((size_t)m * BP * BP));
    PHG_HIP(w->Y32.alloc((size_t)m * BP * BP));
    PHG_HIP(w->colsq.alloc(1024 * BP));
    PHG_HIP(w->stat.alloc(4 * BP));
    PHG_HIP(w->flags.alloc(4));
    PHG_HIP(hipHostMalloc((void**)&w->hstat, 4 * BP * sizeof(double), hipHostMallocDefault));
    return PHG_OK;
}

void block_destroy(BlockWork* w) { delete w; }

template <typename MT>
int block_gmres_run(phg_matrix_s* A, const Ell<MT>& E, BlockWork* w, const double* B, double* X,
                    const phg_gmres_opts& o, phg_block_info* info, hipStream_t s)
{
    const int n = w->n, m = o.restart;
    const size_t bs = w->bs;
    const int ld = (m + 1) * BP;  // Hbar / G leading dimension for this solve
    const int rblk = std::min(1024, (n + 3) / 4);
    const dim3 tiles((n + 127) / 128);
    if (!o.use_x0) {
        const size_t cnt = (size_t)n * BP;
        hipLaunchKernelGGL(k_fill0_d, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, X);
    }
    PHG_HIP(hipMemsetAsync(w->flags.p, 0, 4 * sizeof(int), s));
    hipLaunchKernelGGL(k_colsum_b, dim3(rblk), dim3(256), 0, s, n, B, w->colsq.p);
    hipLaunchKernelGGL(k_colnorm, dim3(1), dim3(BP), 0, s, rblk, w->colsq.p, w->stat.p, 2 * BP);
    int iters = 0, cycles = 0, converged = 0;
    double maxrel = 0.0;
    for (;;) {
        // true FP64 residual and its column norms
        hipLaunchKernelGGL(k_residual64, dim3(rblk), dim3(256), 0, s, n, A->ell64.width,
                           A->ell64.col.p, A->ell64.val.p, B, X, w->R.p, w->W.p, w->colsq.p);
        hipLaunchKernelGGL(k_colnorm, dim3(1), dim3(BP), 0, s, rblk, w->colsq.p, w->stat.p, BP);
        PHG_HIP(hipMemcpyAsync(w->hstat, w->stat.p, 4 * BP * sizeof(double), hipMemcpyDeviceToHost, s));
        PHG_HIP(hipStreamSynchronize(s));
        maxrel = 0.0;
        for (int q = 0; q < BP; ++q) {
            const double bn = w->hstat[2 * BP + q] > 0 ? w->hstat[2 * BP + q] : 1.0;
            maxrel = std::max(maxrel, w->hstat[BP + q] / bn);
        }
        if (maxrel <= o.rtol) {
            converged = 1;
            break;
        }
        if (iters >= o.max_iters) break;
        ++cycles;
        // V_0 = qr(R), S -> G
        int rc = cholqr2(w, w->W.p, 0, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_ls_init_rhs, dim3(64), dim3(256), 0, s, ld, ld, w->Racc.p, w->G.p);
        int J = 0;
        for (int j = 0; j < m; ++j) {
            // W = A V_j
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spmm<MT>), dim3((n + 3) / 4), dim3(256), 0, s, n,
                               E.width, E.col.p, E.val.p, w->V.p + (size_t)j * bs, (size_t)BP,
                               w->W.p, (size_t)BP);
            const int nb = j + 1;
            const int npair = (nb + 1) / 2;
            const int nc = std::max(1, std::min(1024 / npair, (n + 255) / 256));
            const int chunk = (n + nc - 1) / nc;
            const int ncu = (n + chunk - 1) / chunk;
            const int passes = (o.ortho == PHG_CGS1) ? 1 : 2;
            for (int pass = 0; pass < passes; ++pass) {
                hipLaunchKernelGGL(k_atb, dim3(ncu, npair), dim3(256), 0, s, n, w->V.p, bs, w->W.p,
                                   nb, ncu, chunk, w->part.p);
                const int tot = nb * BP * BP;
                hipLaunchKernelGGL(k_atb_reduce, dim3((tot + 255) / 256), dim3(256), 0, s, nb, ncu,
                                   w->part.p, w->Hc.p, w->Hs.p, pass);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_vh<0>), tiles, dim3(256), 0, s, n, w->V.p, bs,
                                   nb * BP, w->Hs.p, w->W.p, nullptr);
            }
            int rc2 = cholqr2(w, w->W.p, j + 1, s);
            if (rc2) return rc2;
            hipLaunchKernelGGL(k_put_column, dim3(64), dim3(256), 0, s, j, ld, w->Hc.p, w->Racc.p,
                               w->Hbar.p);
            if (j > 0)
                hipLaunchKernelGGL(k_ls_apply, dim3(16), dim3(256), 0, s, j, ld, w->Hbar.p, w->U.p,
                                   w->Tm.p);
            hipLaunchKernelGGL(k_ls_factor, dim3(1), dim3(1024), 0, s, j, ld, w->Hbar.p, w->U.p,
                               w->tau.p);
            hipLaunchKernelGGL(k_ls_wy, dim3(1), dim3(1024), 0, s, j, w->U.p, w->tau.p, w->Tm.p);
            hipLaunchKernelGGL(k_ls_rhs, dim3(16), dim3(256), 0, s, j, ld, w->G.p, w->U.p, w->Tm.p,
                               w->stat.p);
            PHG_HIP(hipGetLastError());
            PHG_HIP(hipMemcpyAsync(w->hstat, w->stat.p, 4 * BP * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
            PHG_HIP(hipStreamSynchronize(s));
            ++iters;
            J = j + 1;
            bool stop = true;
            for (int q = 0; q < BP; ++q) {
                const double bn = w->hstat[2 * BP + q] > 0 ? w->hstat[2 * BP + q] : 1.0;
                const double lim = std::max(o.rtol * bn, o.inner_rtol * w->hstat[BP + q]);
                if (w->hstat[q] > lim) stop = false;
            }
            if (stop || iters >= o.max_iters) break;
        }
        // Y = R^{-1} G, X += V Y
        const int K = J * BP;
        hipLaunchKernelGGL(k_ls_copy_rhs, dim3((K * BP + 255) / 256), dim3(256), 0, s, K, ld,
                           w->G.p, w->Y.p);
        for (int b = J - 1; b >= 0; --b) {
            hipLaunchKernelGGL(k_ls_diag, dim3(1), dim3(1024), 0, s, b, ld, w->Hbar.p, w->Y.p,
                               w->Y32.p);
            if (b > 0)
                hipLaunchKernelGGL(k_ls_update, dim3(b * BP / 32), dim3(256), 0, s, b, ld, w->Hbar.p,
                                   w->Y.p);
        }
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_vh<2>), tiles, dim3(256), 0, s, n, w->V.p, bs, J * BP,
                           w->Y32.p, nullptr, X);
        PHG_HIP(hipGetLastError());
    }
    if (info) {
        info->iterations = iters;
        info->cycles = cycles;
        info->max_rel_residual = maxrel;
        info->converged = converged;
    }
    return converged ? PHG_OK : PHG_ERR_NOCONV;
}

int block_gmres(phg_matrix_s* A, int nrhs, const double* B, double* X, const phg_gmres_opts* opts,
                phg_block_info* info, hipStream_t s)
{
    phg_gmres_opts o = *opts;
    if (nrhs != BP) {
        set_error("block gmres: nrhs must be 64");
        return PHG_ERR_ARG;
    }
    if (o.restart < 1 || o.restart > B_MAXM) {
        set_error("block gmres: restart must be in 1..32");
        return PHG_ERR_ARG;
    }
    if (A->n_rows != A->n_cols || A->n_rows < 64) {
        set_error("block gmres: square matrix with at least 64 rows required");
        return PHG_ERR_ARG;
    }
    int rc = ensure_ell64(A, s);
    if (rc) return rc;
    rc = block_prepare(A, o.restart, A->bwork);
    if (rc) return rc;
    if (o.precision == PHG_BF16) {
        rc = ensure_ell16(A, s);
        if (rc) return rc;
        return block_gmres_run<uint16_t>(A, A->ell16, A->bwork, B, X, o, info, s);
    }
    rc = ensure_ell32(A, s);
    if (rc) return rc;
    return block_gmres_run<float>(A, A->ell32, A->bwork, B, X, o, info, s);
}

}  // namespace phg
