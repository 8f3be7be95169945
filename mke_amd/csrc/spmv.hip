// Sparse products, BLAS-1 and gather/scatter on gfx950.
//
// SpMV: thread per row over a slot-major ELL copy of the CSR.  Rows of the
// sphere operators hold 6-7 entries (nnz = 70*4^k - 11 for n = 10*4^k + 1
// rows), so ELL of width 7 carries almost no padding and every value/column
// load of a wavefront is one contiguous 256/512-byte segment.  The x gathers
// stay within the Cuthill-McKee bandwidth of the row block and hit L2.
// FP64 products are accumulated in CSR column order from 0.0 without fused
// multiply-add (this file is compiled with -ffp-contract=off), so y = A x is
// bit-identical to a plain CPU CSR loop.
#include <atomic>

#include "ell.h"
#include "phg_internal.h"

namespace phg {

__device__ __forceinline__ float bf16_to_f32(uint16_t b)
{
    return __uint_as_float(((uint32_t)b) << 16);
}

__host__ __device__ inline uint16_t f64_to_bf16(double v)
{
    // round to nearest even on the float value (NaN kept quiet)
    float f = (float)v;
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

template <typename V>
__device__ __forceinline__ float val_f32(V v);
template <>
__device__ __forceinline__ float val_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float val_f32<uint16_t>(uint16_t v) { return bf16_to_f32(v); }

namespace {

template <typename V>
__global__ void csr_to_ell(int n, int width, const int* __restrict__ rp,
                           const int* __restrict__ ci, const double* __restrict__ val,
                           int* __restrict__ ecol, V* __restrict__ eval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r0 = rp[i], r1 = rp[i + 1];
    const int pad_col = r1 > r0 ? ci[r0] : 0;
    for (int k = 0; k < width; ++k) {
        const int e = r0 + k;
        const size_t o = (size_t)k * n + i;
        if (e < r1) {
            ecol[o] = ci[e];
            if constexpr (sizeof(V) == 8) eval[o] = val[e];
            else if constexpr (sizeof(V) == 4) eval[o] = (float)val[e];
            else eval[o] = f64_to_bf16(val[e]);
        } else {
            ecol[o] = pad_col;
            eval[o] = V(0);
        }
    }
}

__global__ void __launch_bounds__(1024)
ell_spmv_f64(int n, int width, const int* __restrict__ col, const double* __restrict__ val,
             const double* __restrict__ x, double* __restrict__ y)
{
    const int i = lblock() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = ell_row_f64(n, width, i, col, val, x, 1.0, false);
}

// The same product for matrices larger than the L2s can hold (level 8 and
// up, rows of at most ELL_MAXW entries): the column/value stream is read
// with non-temporal loads, so it does not evict the gathered x, and XCD x
// takes the contiguous x-th eighth of the blocks, so its L2 fetches about
// an eighth of x plus the stencil's halo.  Every column slot is loaded
// before the value slots, so each gather waits for its column alone; WD is
// the row width when EXACT (the sphere operators have 7 entries a row, 6 at
// level 0), else ELL_MAXW with slots past the width clamped (ell_slot).
// One wavefront per workgroup: at level 8 (10,241 workgroups) the finer
// grain balances the CUs' share of the stream better than 128- or
// 256-thread workgroups (11.9 vs 12.1 and 12.5 us per launch over 8 matrix
// copies).  Same summation order as ell_row_f64 (bitwise the same y).
constexpr int SPMV_STREAM_BS = 64;
template <int WD, bool EXACT>
__global__ void __launch_bounds__(SPMV_STREAM_BS)
ell_spmv_f64_stream(int n, int width, const int* __restrict__ col,
                    const double* __restrict__ val, const double* __restrict__ x,
                    double* __restrict__ y)
{
    const int i = xcd_block((int)blockIdx.x, (int)gridDim.x) * SPMV_STREAM_BS + threadIdx.x;
    if (i >= n) return;
    int c[WD];
    double v[WD], xv[WD];
#pragma unroll
    for (int k = 0; k < WD; ++k)
        c[k] = __builtin_nontemporal_load(col + (EXACT ? (size_t)k * n + i : ell_slot(k, width, n, i)));
#pragma unroll
    for (int k = 0; k < WD; ++k)
        v[k] = __builtin_nontemporal_load(val + (EXACT ? (size_t)k * n + i : ell_slot(k, width, n, i)));
#pragma unroll
    for (int k = 0; k < WD; ++k) xv[k] = x[c[k]];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < WD; ++k) {
        const double t = s + v[k] * xv[k];
        s = (EXACT || k < width) ? t : s;
    }
    y[i] = s;
}

template <typename V>
__global__ void __launch_bounds__(1024)
ell_spmv_f32(int n, int width, const int* __restrict__ col, const V* __restrict__ val,
             const float* __restrict__ x, float* __restrict__ y)
{
    const int i = lblock() * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = ell_row_f32<V>(n, width, i, col, val, x, 1.0f);
}

// Y = A X, X/Y row-major n x nrhs.  One wavefront per row: lane r owns
// right-hand sides r, r+64, ...; the row's columns/values are wave-uniform.
template <typename V>
__global__ void __launch_bounds__(256)
ell_spmm_f32(int n, int width, int nrhs, const int* __restrict__ col,
             const V* __restrict__ val, const float* __restrict__ X, float* __restrict__ Y)
{
    const int lane = threadIdx.x & 63;
    const int row = lblock() * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (row >= n) return;
    for (int r = lane; r < nrhs; r += 64) {
        float s = 0.0f;
        for (int k = 0; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            s = fmaf(val_f32(val[o]), X[(size_t)col[o] * nrhs + r], s);
        }
        Y[(size_t)row * nrhs + r] = s;
    }
}

__global__ void __launch_bounds__(256)
dot_partial(int n, const double* __restrict__ x, const double* __restrict__ y,
            double* __restrict__ part)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        s += x[i] * y[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void axpby_kernel(int n, double k1, const double* __restrict__ a, double k2,
                             const double* __restrict__ b, double* __restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = a[i] * k1 + b[i] * k2;
}

__global__ void u2p_kernel(int rs, const int* __restrict__ inner, const double* __restrict__ u,
                           double* __restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rs) p[i] = u[inner[i]];
}

__global__ void p2u_kernel(int sz, const int* __restrict__ isb, const int* __restrict__ p2io,
                           const double* __restrict__ p, const double* __restrict__ bnd,
                           double* __restrict__ u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sz) return;
    if (isb[i]) u[i] = bnd ? bnd[p2io[i]] : 0.0;
    else u[i] = p[p2io[i]];
}

// F[k] = F1[outer[k]]  (proj_bnd)
__global__ void proj_bnd_kernel(int os, const int* __restrict__ outer,
                                const double* __restrict__ F1, double* __restrict__ F)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < os) F[i] = F1[outer[i]];
}

// u[i] = bnd ? bnd[p2io[i]] : 0 at boundary points, inner points untouched (set_bnd)
__global__ void set_bnd_kernel(int sz, const int* __restrict__ isb, const int* __restrict__ p2io,
                               const double* __restrict__ bnd, double* __restrict__ u)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < sz && isb[i]) u[i] = bnd ? bnd[p2io[i]] : 0.0;
}

int max_row(const phg_matrix_s* A)
{
    int w = 0;
    for (int i = 0; i < A->n_rows; ++i) w = std::max(w, A->h_rp[i + 1] - A->h_rp[i]);
    return std::max(w, 1);
}

template <typename V>
int build_ell(phg_matrix_s* A, Ell<V>& E, hipStream_t s)
{
    if (E.width) return PHG_OK;
    const int n = A->n_rows;
    const int w = max_row(A);
    PHG_HIP(E.col.alloc((size_t)w * n));
    PHG_HIP(E.val.alloc((size_t)w * n));
    if (n > 0) {
        hipLaunchKernelGGL(csr_to_ell<V>, dim3((n + 255) / 256), dim3(256), 0, s, n, w,
                           A->rp.p, A->ci.p, A->val.p, E.col.p, E.val.p);
        PHG_HIP(hipGetLastError());
    }
    PHG_HIP(hipStreamSynchronize(s));
    E.width = w;
    return PHG_OK;
}

}  // namespace

int ensure_ell64(phg_matrix_s* A, hipStream_t s) { return build_ell(A, A->ell64, s); }

namespace {
__global__ void diag_inverse(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                             const double* __restrict__ val, double* __restrict__ d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) a = val[k];
    d[i] = a != 0.0 ? 1.0 / a : 1.0;
}

// E2 = E with every value times d[column]
__global__ void ell_scale_cols(size_t m, const int* __restrict__ col,
                               const double* __restrict__ val, const double* __restrict__ d,
                               int* __restrict__ col2, double* __restrict__ val2)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    col2[e] = col[e];
    val2[e] = val[e] * d[col[e]];
}
}  // namespace

int ensure_ell64_diag(phg_matrix_s* A, hipStream_t s)
{
    if (A->ell64d.width) return PHG_OK;
    if (A->n_rows != A->n_cols) {
        set_error("diagonal scaling needs a square matrix");
        return PHG_ERR_ARG;
    }
    int rc = ensure_ell64(A, s);
    if (rc) return rc;
    const int n = A->n_rows;
    const size_t m = (size_t)A->ell64.width * n;
    PHG_HIP(A->dinv.alloc(n));
    PHG_HIP(A->ell64d.col.alloc(m));
    PHG_HIP(A->ell64d.val.alloc(m));
    if (n) {
        hipLaunchKernelGGL(diag_inverse, dim3((n + 255) / 256), dim3(256), 0, s, n, A->rp.p,
                           A->ci.p, A->val.p, A->dinv.p);
        hipLaunchKernelGGL(ell_scale_cols, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m,
                           A->ell64.col.p, A->ell64.val.p, A->dinv.p, A->ell64d.col.p,
                           A->ell64d.val.p);
        PHG_HIP(hipGetLastError());
    }
    PHG_HIP(hipStreamSynchronize(s));
    A->ell64d.width = A->ell64.width;
    return PHG_OK;
}
int ensure_ell32(phg_matrix_s* A, hipStream_t s) { return build_ell(A, A->ell32, s); }
int ensure_ell16(phg_matrix_s* A, hipStream_t s) { return build_ell(A, A->ell16, s); }

// Threads per block of the one-row-per-thread SpMV: the power of two nearest
// n / 1280 within 64..1024, so a launch is about 1280 workgroups (5 per CU)
// at every mesh level (measured at levels 7 and 8: 128 and 512 give the
// shortest launches, 12-14 % below 256).
// Rows from which spmv_f64 uses the streaming kernel (between levels 7 and
// 8: measured at level 8, 256 threads, non-temporal matrix loads and
// contiguous XCD eighths give 10.6 us repeated and 12.2 us when the matrix
// streams from HBM, against 12.4 / 14.7 us for the plain kernel; at level 7,
// whose 14 MB stay cached, they make the launch slower).
constexpr int SPMV_STREAM_ROWS = 1 << 18;

int spmv_block(int n)
{
    int bs = 64;
    while (bs < 1024 && (double)n / (bs * 2) >= 1280.0 / 1.41421356) bs *= 2;
    return bs;
}

int spmv_f64(phg_matrix_s* A, const double* x, double* y, hipStream_t s)
{
    int rc = ensure_ell64(A, s);
    if (rc) return rc;
    const int n = A->n_rows;
    if (n == 0) return PHG_OK;
    if (n >= SPMV_STREAM_ROWS && A->ell64.width >= 1 && A->ell64.width <= ELL_MAXW) {
        const dim3 grid((n + SPMV_STREAM_BS - 1) / SPMV_STREAM_BS), block(SPMV_STREAM_BS);
        const int w = A->ell64.width;
        if (w == 7)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(ell_spmv_f64_stream<7, true>), grid, block, 0, s,
                               n, w, A->ell64.col.p, A->ell64.val.p, x, y);
        else if (w == 6)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(ell_spmv_f64_stream<6, true>), grid, block, 0, s,
                               n, w, A->ell64.col.p, A->ell64.val.p, x, y);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(ell_spmv_f64_stream<ELL_MAXW, false>), grid, block,
                               0, s, n, w, A->ell64.col.p, A->ell64.val.p, x, y);
        PHG_HIP(hipGetLastError());
        return PHG_OK;
    }
    const int bs = spmv_block(n);
    hipLaunchKernelGGL(ell_spmv_f64, dim3((n + bs - 1) / bs), dim3(bs), 0, s, n,
                       A->ell64.width, A->ell64.col.p, A->ell64.val.p, x, y);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int spmv_f32(phg_matrix_s* A, int prec, const float* x, float* y, hipStream_t s)
{
    const int n = A->n_rows;
    if (prec == PHG_BF16) {
        int rc = ensure_ell16(A, s);
        if (rc) return rc;
        const int bs = spmv_block(n);
        if (n)
            hipLaunchKernelGGL(ell_spmv_f32<uint16_t>, dim3((n + bs - 1) / bs), dim3(bs), 0, s,
                               n, A->ell16.width, A->ell16.col.p, A->ell16.val.p, x, y);
    } else {
        int rc = ensure_ell32(A, s);
        if (rc) return rc;
        const int bs = spmv_block(n);
        if (n)
            hipLaunchKernelGGL(ell_spmv_f32<float>, dim3((n + bs - 1) / bs), dim3(bs), 0, s, n,
                               A->ell32.width, A->ell32.col.p, A->ell32.val.p, x, y);
    }
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int spmm_f32(phg_matrix_s* A, int prec, int nrhs, const float* X, float* Y, hipStream_t s)
{
    const int n = A->n_rows;
    const dim3 grid((n + 3) / 4), block(256);
    if (prec == PHG_BF16) {
        int rc = ensure_ell16(A, s);
        if (rc) return rc;
        if (n)
            hipLaunchKernelGGL(ell_spmm_f32<uint16_t>, grid, block, 0, s, n, A->ell16.width,
                               nrhs, A->ell16.col.p, A->ell16.val.p, X, Y);
    } else {
        int rc = ensure_ell32(A, s);
        if (rc) return rc;
        if (n)
            hipLaunchKernelGGL(ell_spmm_f32<float>, grid, block, 0, s, n, A->ell32.width, nrhs,
                               A->ell32.col.p, A->ell32.val.p, X, Y);
    }
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int dot_f64(int n, const double* x, const double* y, double* result, hipStream_t s)
{
    const int G = std::max(1, std::min(256, (n + 255) / 256));
    // the 256 partials: a device buffer and a pinned host mirror per calling
    // thread and device, allocated on the thread's first call on the device
    // and kept (a call never allocates again; threads never share them)
    constexpr int MAXDEV = 64;
    thread_local double* part[MAXDEV] = {};
    thread_local double* host[MAXDEV] = {};
    int dev = 0;
    PHG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAXDEV) {
        set_error("phg_dot: device index beyond the 64 this library tracks");
        return PHG_ERR_ARG;
    }
    if (!part[dev]) {
        PHG_HIP(hipMalloc((void**)&part[dev], 256 * sizeof(double)));
        PHG_HIP(hipHostMalloc((void**)&host[dev], 256 * sizeof(double), hipHostMallocDefault));
    }
    hipLaunchKernelGGL(dot_partial, dim3(G), dim3(256), 0, s, n, x, y, part[dev]);
    PHG_HIP(hipGetLastError());
    PHG_HIP(hipMemcpyAsync(host[dev], part[dev], G * sizeof(double), hipMemcpyDeviceToHost, s));
    PHG_HIP(hipStreamSynchronize(s));
    double r = 0.0;
    for (int g = 0; g < G; ++g) r += host[dev][g];
    *result = r;
    return PHG_OK;
}

// ---- phg_dot_noalloc: the dot product without allocating -----------------
// Its block partials live in `scratch`, a file-scope device array: the code
// object holds one copy per GPU from load until the process ends.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (kept per device in scratch_ptr); the array's name is not a
// device address on the host.  Every call on a GPU shares the array, so
// calls that overlap in time on one GPU race.  Each call rewrites every
// partial it reads (one per workgroup launched), so nothing needs clearing.
constexpr int NOALLOC_DOT_ELEMS = 1024;  // entries per workgroup: 256 threads x 4
constexpr int NOALLOC_DOT_MAXG = PHG_DOT_NOALLOC_MAX / NOALLOC_DOT_ELEMS;  // 4096 partials
static __device__ __align__(256) double scratch[NOALLOC_DOT_MAXG];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::atomic<double*> scratch_ptr[NOALLOC_MAX_DEVICES];

__device__ __forceinline__ double dot_wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// part[g] = sum of x[i] y[i] over i = g*1024 + t + 256 r (t < 256, r < 4):
// per thread in r order, xor-butterfly wave sums, then the four waves in order
__global__ void __launch_bounds__(256)
dot_partial_noalloc(int n, const double* __restrict__ x, const double* __restrict__ y,
                    double* __restrict__ part)
{
    __shared__ double sh[4];
    const int base = blockIdx.x * NOALLOC_DOT_ELEMS + threadIdx.x;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < NOALLOC_DOT_ELEMS / 256; ++r) {
        const int i = base + 256 * r;
        if (i < n) s += x[i] * y[i];
    }
    s = dot_wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// result[0] = the G partials: thread t adds t, t + 256, ... in order, then
// wave sums and the four waves in order (deterministic)
__global__ void __launch_bounds__(256)
dot_finish_noalloc(int G, const double* __restrict__ part, double* __restrict__ result)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int g = threadIdx.x; g < G; g += 256) s += part[g];
    s = dot_wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) result[0] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// The scratch address on the current device, looked up once per device.
int noalloc_scratch(double** out)
{
    int dev = 0;
    PHG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("phg_dot_noalloc: device index beyond the 64 this library tracks");
        return PHG_ERR_ARG;
    }
    double* p = scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        void* q = nullptr;
        PHG_HIP(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        p = (double*)q;
        scratch_ptr[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return PHG_OK;
}

int dot_noalloc_f64(int n, const double* x, const double* y, double* result, hipStream_t s)
{
    if (n < 0 || n > PHG_DOT_NOALLOC_MAX) {
        set_error("phg_dot_noalloc: n must be in 0..4194304 (the 4096 block partials its "
                  "32 KiB device scratch holds)");
        return PHG_ERR_ARG;
    }
    if (!result || (n > 0 && (!x || !y))) {
        set_error("phg_dot_noalloc: null argument");
        return PHG_ERR_ARG;
    }
    double* part = nullptr;
    int rc = noalloc_scratch(&part);
    if (rc) return rc;
    const int G = std::max(1, (n + NOALLOC_DOT_ELEMS - 1) / NOALLOC_DOT_ELEMS);
    hipLaunchKernelGGL(dot_partial_noalloc, dim3(G), dim3(256), 0, s, n, x, y, part);
    hipLaunchKernelGGL(dot_finish_noalloc, dim3(1), dim3(256), 0, s, G, part, result);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int axpby_f64(int n, double k1, const double* a, double k2, const double* b, double* r,
              hipStream_t s)
{
    if (n <= 0) return PHG_OK;
    hipLaunchKernelGGL(axpby_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, k1, a, k2, b, r);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int u2p_f64(phg_mesh_s* m, const double* u, double* p, hipStream_t s)
{
    int rc = ensure_mesh_device(m);
    if (rc) return rc;
    const int rs = (int)m->h.inner.size();
    if (rs)
        hipLaunchKernelGGL(u2p_kernel, dim3((rs + 255) / 256), dim3(256), 0, s, rs,
                           m->d_inner.p, u, p);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int p2u_f64(phg_mesh_s* m, const double* p, const double* bnd, double* u, hipStream_t s)
{
    int rc = ensure_mesh_device(m);
    if (rc) return rc;
    const int sz = m->h.n_points;
    if (sz)
        hipLaunchKernelGGL(p2u_kernel, dim3((sz + 255) / 256), dim3(256), 0, s, sz, m->d_isb.p,
                           m->d_p2io.p, p, bnd, u);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int proj_bnd_f64(phg_mesh_s* m, const double* u, double* b, hipStream_t s)
{
    int rc = ensure_mesh_device(m);
    if (rc) return rc;
    const int os = (int)m->h.outer.size();
    if (os)
        hipLaunchKernelGGL(proj_bnd_kernel, dim3((os + 255) / 256), dim3(256), 0, s, os,
                           m->d_outer.p, u, b);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

int set_bnd_f64(phg_mesh_s* m, const double* bnd, double* u, hipStream_t s)
{
    int rc = ensure_mesh_device(m);
    if (rc) return rc;
    const int sz = m->h.n_points;
    if (sz)
        hipLaunchKernelGGL(set_bnd_kernel, dim3((sz + 255) / 256), dim3(256), 0, s, sz,
                           m->d_isb.p, m->d_p2io.p, bnd, u);
    PHG_HIP(hipGetLastError());
    return PHG_OK;
}

}  // namespace phg
