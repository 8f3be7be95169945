// Row dot products over the slot-major ELL layout (entry k of row i at
// k * n + i), shared by the SpMV, residual and Arnoldi kernels.
//
// For rows of at most ELL_MAXW entries (the sphere operators have 6-7) the
// column/value loads of all slots are issued first, then all x gathers, then
// the products are added in slot order; a row's memory latency is paid twice
// instead of 2 * width times.  Sums start from 0 and add slots in order, so
// FP64 results equal a plain CSR loop over the same row.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace phg {

constexpr int ELL_MAXW = 8;

// Workgroup b of a grid of nb is dispatched to XCD b % 8 (round robin over
// the 8 XCDs, each with its own L2).  The row kernels take their block
// number from lblock(), which can renumber with xcd_block(b, nb): XCD x then
// owns the contiguous x-th eighth of the rows, the x entries its rows gather
// lie in that eighth plus the stencil's halo, and each L2 fetches about 1/8
// of x instead of most of it.  A bijection on [0, nb) for every nb.
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// Same idea with XCD x taking every 8th run of S consecutive blocks (runs of
// S * 256 rows); the grid's tail beyond the last whole group of 8 runs keeps
// its order.  A bijection on [0, nb) for every nb and S >= 1.
__device__ __forceinline__ int xcd_block_runs(int b, int nb, int S)
{
    const int full = nb / (8 * S) * (8 * S);
    if (b >= full) return b;
    const int x = b & 7, k = b >> 3;
    return ((k / S) * 8 + x) * S + k % S;
}

// Off by default: measured on the L7 solve (17 MB per SpMV), the contiguous
// eighths cut the SpMV's fetched bytes from 1.20x to 0.97x of the
// algorithmic bytes but not its time (3.48 vs 3.53 us), and the extra index
// arithmetic at the start of the latency-bound MGS/CGS kernels cost 4-8 %.
#ifndef PHG_XCD_RUN
#define PHG_XCD_RUN 0  // -1: one contiguous eighth per XCD; S > 0: runs of S blocks
#endif

__device__ __forceinline__ int lblock()
{
#if PHG_XCD_RUN < 0
    return xcd_block((int)blockIdx.x, (int)gridDim.x);
#elif PHG_XCD_RUN == 0
    return (int)blockIdx.x;
#else
    return xcd_block_runs((int)blockIdx.x, (int)gridDim.x, PHG_XCD_RUN);
#endif
}

template <typename MT>
__device__ __forceinline__ float ell_val_f32(MT v);
template <>
__device__ __forceinline__ float ell_val_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float ell_val_f32<uint16_t>(uint16_t v)
{
    return __uint_as_float(((uint32_t)v) << 16);
}

// Offset of slot k of row `row` in a slot-major ELL of width 1..ELL_MAXW
// (callers check width >= 1: an empty ELL has no slot to read), clamped to
// the row's last slot.  The unrolled row loops below issue every
// slot's loads without a branch (a slot beyond the width reads the last
// slot again, in bounds) and add only the products of slots k < width, so
// the sum is the same as a loop over the width; with no branch between the
// loads the compiler can wait for each gather's column alone instead of
// for every load issued before it.
__device__ __forceinline__ size_t ell_slot(int k, int width, int n, int row)
{
    return (size_t)(k < width ? k : width - 1) * n + row;
}

// s = sum_k val[k] * (x[col[k]] * scale), FP64 (no fused multiply-add; the
// including file is compiled with -ffp-contract=off).  scale = 1 multiplies
// exactly.
__device__ __forceinline__ double ell_row_f64(int n, int width, int row,
                                              const int* __restrict__ col,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, double scale,
                                              bool scaled)
{
    double s = 0.0;
    if (width >= 1 && width <= ELL_MAXW) {
        int c[ELL_MAXW];
        double v[ELL_MAXW], xv[ELL_MAXW];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = ell_slot(k, width, n, row);
            c[k] = col[o];
            v[k] = val[o];
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const double t = s + v[k] * (scaled ? xv[k] * scale : xv[k]);
            s = k < width ? t : s;
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            s += val[o] * (scaled ? x[col[o]] * scale : x[col[o]]);
        }
    }
    return s;
}

// FP32 vectors, values stored as float or bfloat16, fused multiply-add.
template <typename MT>
__device__ __forceinline__ float ell_row_f32(int n, int width, int row,
                                             const int* __restrict__ col,
                                             const MT* __restrict__ val,
                                             const float* __restrict__ x, float scale)
{
    float s = 0.0f;
    if (width >= 1 && width <= ELL_MAXW) {
        int c[ELL_MAXW];
        MT v[ELL_MAXW];
        float xv[ELL_MAXW];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = ell_slot(k, width, n, row);
            c[k] = col[o];
            v[k] = val[o];
        }
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const float t = fmaf(ell_val_f32(v[k]), xv[k] * scale, s);
            s = k < width ? t : s;
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            s = fmaf(ell_val_f32(val[o]), x[col[o]] * scale, s);
        }
    }
    return s;
}

}  // namespace phg
