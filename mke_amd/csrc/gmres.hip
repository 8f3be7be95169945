// Restarted GMRES(m) on gfx950, device resident.
//
// Replaces Solver<T>::solve of the reference's external solver library
// (call sites src/mesh.h:740-746, src/impl/slaplace_impl.h:171).  One restart
// cycle = m Arnoldi steps; each cycle begins from the true FP64 residual
// r = b - A x, so the same driver gives
//   * the plain FP64 form (vectors, matrix and Krylov basis in FP64), and
//   * the mixed form: Krylov basis, work vector and Arnoldi SpMV in FP32
//     (matrix values stored as FP32 or BF16), with FP64 kept for the
//     residual r = b - A x, every dot/norm accumulation, the Hessenberg
//     matrix and its Givens rotations, and the update x += V y.  Restarting
//     from the FP64 residual makes each cycle one step of iterative
//     refinement.
//
// Orthogonalisation: modified Gram-Schmidt (one kernel per basis vector, the
// reference form) or classical Gram-Schmidt with one re-orthogonalisation
// pass (CGS2: the j+1 dots of a pass are computed together).
//
// Kernels per Arnoldi step j: MGS j+2 (SpMV, j MGS passes, final pass with
// the norm); CGS2 3 (SpMV with the first-pass dots, and two update kernels,
// each of which first adds the previous kernel's dot partials itself).  The Givens rotation of step j-1 and the decision to
// continue are computed redundantly by every block at the start of the SpMV
// kernel of step j (all blocks read the same inputs and agree); block 0
// records them.  A whole restart cycle is captured once in a hipGraph and
// replayed; kernels of a finished cycle return at once.
//
// Reductions are deterministic: every block writes one partial per dot, and
// the consumer adds the partials in block order.  Results depend only on the
// inputs and the launch shape, never on timing.
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "ell.h"
#include "phg_internal.h"

namespace phg {

constexpr int GM_MAXM = 64;
constexpr int GM_BS = 256;
constexpr int GM_RPT = 4;     // max rows per thread
constexpr int GM_RG = 4;      // rows whose Arnoldi SpMV loads are issued together
constexpr int GM_MAXG = 1024; // max blocks

struct GmresState {
    double H[(GM_MAXM + 1) * GM_MAXM];  // Hessenberg columns as computed (unrotated)
    double R[GM_MAXM * GM_MAXM];        // rotated (upper triangular) columns
    double cs[GM_MAXM], sn[GM_MAXM];
    double g[GM_MAXM];                  // rotated right-hand side, final entries
    double carry[GM_MAXM + 1];          // carry[j+1] = last entry after step j
    double y[GM_MAXM];
    double hc[GM_MAXM + 1];             // CGS2 coefficients of the current pass
    double beta, bnorm, rel, rtol, inner_rtol;
    int m, active, k, iters, iters0, max_iters, converged, cycles;
    int abort;                          // persistent cycle: an exchange timed out
};

// The persistent cycle's exchange rows live in uncached device memory
// (hipDeviceMallocUncached): an agent-scope store of a partial and the polls
// that see it cost 2.3 instead of 2.8 us per sum across 161 workgroups in
// coarse-grained memory, whose accesses go through each XCD's L2.  Uncached
// memory that a kernel has read is never freed while the process runs (an
// allocation made later at its address may read stale data): a workspace
// hands its rows back to this pool when it is destroyed, after its last
// solve has synchronised, and the next workspace on the device takes them.
struct XchChunk {
    unsigned long long* p = nullptr;
    size_t n = 0;
    int dev = -1;
};
inline std::mutex& xch_pool_mutex()
{
    static std::mutex* m = new std::mutex();  // kept until exit, like the pool
    return *m;
}
inline std::vector<XchChunk>& xch_pool()
{
    static std::vector<XchChunk>* v = new std::vector<XchChunk>();  // kept until exit
    return *v;
}
struct XchRows {
    XchChunk c;
    XchRows() = default;
    XchRows(const XchRows&) = delete;
    XchRows& operator=(const XchRows&) = delete;
    ~XchRows() { give_back(); }
    unsigned long long* p() const { return c.p; }
    void give_back()
    {
        if (!c.p) return;
        std::lock_guard<std::mutex> g(xch_pool_mutex());
        xch_pool().push_back(c);
        c = XchChunk();
    }
    // at least count slots on the current device: the smallest pooled chunk
    // that holds them, else a new allocation
    hipError_t alloc(size_t count)
    {
        give_back();
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        {
            std::lock_guard<std::mutex> g(xch_pool_mutex());
            std::vector<XchChunk>& pool = xch_pool();
            size_t best = pool.size();
            for (size_t i = 0; i < pool.size(); ++i)
                if (pool[i].dev == dev && pool[i].n >= count &&
                    (best == pool.size() || pool[i].n < pool[best].n))
                    best = i;
            if (best < pool.size()) {
                c = pool[best];
                pool.erase(pool.begin() + (std::ptrdiff_t)best);
                return hipSuccess;
            }
        }
        void* q = nullptr;
        e = hipExtMallocWithFlags(&q, count * sizeof(unsigned long long), hipDeviceMallocUncached);
        if (e != hipSuccess) return e;
        c.p = (unsigned long long*)q;
        c.n = count;
        c.dev = dev;
        return hipSuccess;
    }
};

struct GmresWork {
    int n = 0, m = 0, G = 0, rpt = 1, prec = 0;
    size_t ldv = 0;
    DevBuf<char> V, W;     // (m+1) x ldv and 2 x ldv vectors of the Krylov type
    DevBuf<double> part;   // 2 x (m+2) x G partial sums (two regions for CGS2)
    DevBuf<GmresState> st;
    XchRows xch;           // persistent cycle: exchange rows x G slots
    int xch_rows = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};  // around each persistent cycle launch
    GmresState* host = nullptr;  // pinned mirror
    // captured restart cycle
    hipStream_t cap = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
    int key_m = 0, key_ortho = -1;
    ~GmresWork()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (cap) (void)hipStreamDestroy(cap);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (host) (void)hipHostFree(host);
    }
};

void destroy_gmres_work(GmresWork* w) { delete w; }

namespace {

// ---------------------------------------------------------------- helpers --
// A double moved between lanes by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// v of the even 16-lane row (ROWS 16) or lower half (32) plus v of the odd
// row or upper half, in every lane of the pair (v_permlane16/32_swap).
template <int ROWS>
__device__ __forceinline__ double swap_add(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned l = (unsigned)u, h = (unsigned)(u >> 32);
    const auto lo = ROWS == 16 ? __builtin_amdgcn_permlane16_swap(l, l, false, false)
                               : __builtin_amdgcn_permlane32_swap(l, l, false, false);
    const auto hi = ROWS == 16 ? __builtin_amdgcn_permlane16_swap(h, h, false, false)
                               : __builtin_amdgcn_permlane32_swap(h, h, false, false);
    const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double b = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return a + b;
}

// Sum over the 64 lanes of a wavefront, all active: a butterfly of VALU
// lane moves (no LDS round trip) in which each step adds a lane and its
// partner (lane ^ 1, lane ^ 2, the mirror lane within 8 and within 16, then
// the two rows of a 32-lane half and the two halves), so every lane ends
// with the same, order-fixed bits.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    v += dpp_d<0x140>(v);  // row_mirror
    v = swap_add<16>(v);
    return swap_add<32>(v);
}

// A lane's partials b = lane + 64 k (+0.0 past G), added in k order over the
// groups that hold any of the G partials.
__device__ __forceinline__ double lane_partials(const double (&v)[GM_MAXG / 64], int G)
{
    double s = v[0];
#pragma unroll
    for (int k = 1; k < GM_MAXG / 64; ++k)
        if (64 * k < G) s += v[k];
    return s;
}

// Sum of a block's per-thread values, identical in every thread.
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    v = wave_sum(v);
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < GM_BS / 64; ++w) r += sh[w];
    __syncthreads();
    return r;
}

// Sum of G <= GM_MAXG block partials in a fixed order, computed by each
// wavefront.  The partials were written by other XCDs in the previous
// kernel, so every load is a round trip to memory: all of a lane's loads are
// issued before the first add.
__device__ __forceinline__ double sum_partials(const double* __restrict__ p, int G)
{
    const int lane = threadIdx.x & 63;
    double v[GM_MAXG / 64];
#pragma unroll
    for (int k = 0; k < GM_MAXG / 64; ++k) {
        const int b = lane + 64 * k;
        v[k] = (b < G) ? p[b] : 0.0;
    }
    return wave_sum(lane_partials(v, G));
}

template <typename T>
__device__ __forceinline__ double to_d(T v) { return (double)v; }
template <typename V>
__device__ __forceinline__ float mval(V v);
template <>
__device__ __forceinline__ float mval<float>(float v) { return v; }
template <>
__device__ __forceinline__ float mval<uint16_t>(uint16_t v) { return __uint_as_float(((uint32_t)v) << 16); }

// y_row of the Arnoldi SpMV with the source vector scaled on the fly.
template <typename VT, typename MT>
struct RowDot;
template <>
struct RowDot<double, double> {
    __device__ __forceinline__ static double run(int n, int width, int row, const int* col,
                                                 const double* val, const double* x,
                                                 double inv)
    {
        return ell_row_f64(n, width, row, col, val, x, inv, true);
    }
};
template <typename MT>
struct RowDot<float, MT> {
    __device__ __forceinline__ static float run(int n, int width, int row, const int* col,
                                                const MT* val, const float* x, float inv)
    {
        return ell_row_f32<MT>(n, width, row, col, val, x, inv);
    }
};

template <typename T>
__device__ __forceinline__ T ld_agent(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_agent(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ell_row_f64 / ell_row_f32 with x gathered by agent-scope loads (same order
// of operations)
__device__ __forceinline__ double ell_row_f64_agent(int n, int width, int row,
                                                    const int* __restrict__ col,
                                                    const double* __restrict__ val,
                                                    const double* x, double scale, bool scaled)
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
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = ld_agent(x + c[k]);
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const double t = s + v[k] * (scaled ? xv[k] * scale : xv[k]);
            s = k < width ? t : s;
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            const double xk = ld_agent(x + col[o]);
            s += val[o] * (scaled ? xk * scale : xk);
        }
    }
    return s;
}

template <typename MT>
__device__ __forceinline__ float ell_row_f32_agent(int n, int width, int row,
                                                   const int* __restrict__ col,
                                                   const MT* __restrict__ val, const float* x,
                                                   float scale)
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
        for (int k = 0; k < ELL_MAXW; ++k) xv[k] = ld_agent(x + c[k]);
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const float t = fmaf(ell_val_f32(v[k]), xv[k] * scale, s);
            s = k < width ? t : s;
        }
    } else {
        for (int k = 0; k < width; ++k) {
            const size_t o = (size_t)k * n + row;
            s = fmaf(ell_val_f32(val[o]), ld_agent(x + col[o]) * scale, s);
        }
    }
    return s;
}

template <typename VT, typename MT>
struct RowDotAgent;
template <>
struct RowDotAgent<double, double> {
    __device__ __forceinline__ static double run(int n, int width, int row, const int* col,
                                                 const double* val, const double* x, double inv)
    {
        return ell_row_f64_agent(n, width, row, col, val, x, inv, true);
    }
};
template <typename MT>
struct RowDotAgent<float, MT> {
    __device__ __forceinline__ static float run(int n, int width, int row, const int* col,
                                                const MT* val, const float* x, float inv)
    {
        return ell_row_f32_agent<MT>(n, width, row, col, val, x, inv);
    }
};

// One term of RowDotAgent's sum: s + val * (x * inv) in FP64 (no fused
// multiply-add), fmaf(val, x * inv, s) in FP32, as ell_row_f64_agent and
// ell_row_f32_agent add them.
template <typename VT, typename MT>
struct RowTerm;
template <>
struct RowTerm<double, double> {
    __device__ __forceinline__ static double add(double s, double v, double x, double inv)
    {
        return s + v * (x * inv);
    }
};
template <typename MT>
struct RowTerm<float, MT> {
    __device__ __forceinline__ static float add(float s, MT v, float x, float inv)
    {
        return fmaf(ell_val_f32(v), x * inv, s);
    }
};


// Loads of values other workgroups of the same launch may have written
// (agent scope: they bypass this XCD's L2 copy) and plain loads, chosen at
// compile time.
template <bool AGENT, typename T>
__device__ __forceinline__ T ld_x(const T* p)
{
    if constexpr (AGENT) return ld_agent(p);
    else return *p;
}

// The Arnoldi SpMV of a thread's rows (rows base + r * GM_BS, r < rpt):
// xo[r] = x[row] and w[r] = RowDot (or RowDotAgent) of the row with x
// scaled by inv, 0 for rows at or past n.  For rows of at most ELL_MAXW
// entries, in groups of RG rows: every row's own entry and column/value
// slots are loaded first, then every row's gathers, then the sums in
// RowDot's order: two memory round trips per group instead of three per
// row.
template <typename VT, typename MT, bool AGENT, int RG = GM_RG>
__device__ __forceinline__ void arnoldi_rows(int n, int rpt, int width, int base,
                                             const int* __restrict__ col,
                                             const MT* __restrict__ val, const VT* x, VT inv,
                                             VT (&xo)[GM_RPT], VT (&w)[GM_RPT])
{
    if (width >= 1 && width <= ELL_MAXW) {
#pragma unroll
        for (int g = 0; g < GM_RPT; g += RG) {
            if (g >= rpt) {
#pragma unroll
                for (int r = g; r < GM_RPT; ++r) {
                    xo[r] = VT(0);
                    w[r] = VT(0);
                }
                break;
            }
            int c[RG][ELL_MAXW];
            MT v[RG][ELL_MAXW];
            VT xv[RG][ELL_MAXW];
#pragma unroll
            for (int q = 0; q < RG; ++q) {
                if (g + q < rpt) {
                    const int rr = min(base + (g + q) * GM_BS, n - 1);  // in bounds; unused past n
                    xo[g + q] = ld_x<AGENT>(x + rr);
#pragma unroll
                    for (int k = 0; k < ELL_MAXW; ++k) {
                        const size_t o = ell_slot(k, width, n, rr);
                        c[q][k] = col[o];
                        v[q][k] = val[o];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < RG; ++q) {
                if (g + q < rpt) {
#pragma unroll
                    for (int k = 0; k < ELL_MAXW; ++k) xv[q][k] = ld_x<AGENT>(x + c[q][k]);
                }
            }
#pragma unroll
            for (int q = 0; q < RG; ++q) {
                const int r = g + q;
                VT sum = VT(0);
                if (r < rpt && base + r * GM_BS < n) {
#pragma unroll
                    for (int k = 0; k < ELL_MAXW; ++k) {
                        const VT t = RowTerm<VT, MT>::add(sum, v[q][k], xv[q][k], inv);
                        sum = k < width ? t : sum;
                    }
                } else {
                    xo[r] = VT(0);
                }
                w[r] = sum;
            }
        }
    } else {
#pragma unroll
        for (int r = 0; r < GM_RPT; ++r) {
            const int rr = base + r * GM_BS;
            xo[r] = VT(0);
            w[r] = VT(0);
            if (r < rpt && rr < n) {
                xo[r] = ld_x<AGENT>(x + rr);
                if constexpr (AGENT) w[r] = RowDotAgent<VT, MT>::run(n, width, rr, col, val, x, inv);
                else w[r] = RowDot<VT, MT>::run(n, width, rr, col, val, x, inv);
            }
        }
    }
}

// A thread's row slots (ELL columns and values of its rows; past a row's
// width the last slot again, and rows past n clamped to row n - 1, so every
// load is in bounds), loaded once per launch of the persistent cycle: the
// matrix does not change within a cycle, so each step's SpMV is one round
// trip (the gathers) instead of two.
template <typename MT>
struct RowSlots {
    int c[GM_RPT][ELL_MAXW];
    MT v[GM_RPT][ELL_MAXW];
};
template <typename MT>
__device__ __forceinline__ void load_slots(int n, int width, int base, const int* __restrict__ col,
                                           const MT* __restrict__ val, RowSlots<MT>& s)
{
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int rr = min(base + r * GM_BS, n - 1);
#pragma unroll
        for (int k = 0; k < ELL_MAXW; ++k) {
            const size_t o = ell_slot(k, width, n, rr);
            s.c[r][k] = col[o];
            s.v[r][k] = val[o];
        }
    }
}
// arnoldi_rows (agent-scope gathers, rows of 1..ELL_MAXW entries) on
// preloaded slots: every row's own entry and gathers issued together, then
// the sums in RowDot's order (the same bits).
template <typename VT, typename MT>
__device__ __forceinline__ void arnoldi_rows_slots(int n, int rpt, int width, int base,
                                                   const RowSlots<MT>& s, const VT* x, VT inv,
                                                   VT (&xo)[GM_RPT], VT (&w)[GM_RPT])
{
    VT xv[GM_RPT][ELL_MAXW];
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        if (r < rpt) {
            xo[r] = ld_agent(x + min(base + r * GM_BS, n - 1));
#pragma unroll
            for (int k = 0; k < ELL_MAXW; ++k) xv[r][k] = ld_agent(x + s.c[r][k]);
        }
    }
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        VT sum = VT(0);
        if (r < rpt && base + r * GM_BS < n) {
#pragma unroll
            for (int k = 0; k < ELL_MAXW; ++k) {
                const VT t = RowTerm<VT, MT>::add(sum, s.v[r][k], xv[r][k], inv);
                sum = k < width ? t : sum;
            }
        } else {
            xo[r] = VT(0);
        }
        w[r] = sum;
    }
}

constexpr int GM_LB = 8;  // basis vectors loaded per batch

// Block partials of the dots <wv, v_i>, i < nd, over this thread's rows
// (v_skip is taken from `own` instead of memory); nd = 0 gives ||wv||^2.
// The basis loads are issued GM_LB vectors at a time so that their memory
// latency overlaps; one barrier for all dots.  Writes
// part[(out0 + i) * G + block].
template <typename VT>
__device__ __forceinline__ void block_dots(int nd, int skip, const VT* own, const VT* __restrict__ V,
                                           size_t ldv, int base, int rpt, int n, const VT* wv,
                                           double* __restrict__ part, int G, int out0)
{
    __shared__ double shd[GM_BS / 64][GM_MAXM + 1];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (nd == 0) {
        double acc = 0.0;
#pragma unroll
        for (int r = 0; r < GM_RPT; ++r) {
            const int row = base + r * GM_BS;
            if (r < rpt && row < n) acc += to_d(wv[r]) * to_d(wv[r]);
        }
        acc = wave_sum(acc);
        if (lane == 0) shd[wid][0] = acc;
    }
    for (int i0 = 0; i0 < nd; i0 += GM_LB) {
        VT vb[GM_LB][GM_RPT];
#pragma unroll
        for (int q = 0; q < GM_LB; ++q) {
            const int i = i0 + q;
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) {
                const int row = base + r * GM_BS;
                vb[q][r] = (i < nd && i != skip && r < rpt && row < n) ? V[(size_t)i * ldv + row]
                                                                      : VT(0);
            }
        }
#pragma unroll
        for (int q = 0; q < GM_LB; ++q) {
            const int i = i0 + q;
            if (i >= nd) break;
            double acc = 0.0;
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) {
                const int row = base + r * GM_BS;
                if (r < rpt && row < n) acc += to_d(wv[r]) * to_d(i == skip ? own[r] : vb[q][r]);
            }
            acc = wave_sum(acc);
            if (lane == 0) shd[wid][i] = acc;
        }
    }
    __syncthreads();
    const int outs = nd == 0 ? 1 : nd;
    if ((int)threadIdx.x < outs) {
        double s = 0.0;
        for (int q = 0; q < GM_BS / 64; ++q) s += shd[q][threadIdx.x];
        part[(size_t)(out0 + threadIdx.x) * G + lblock()] = s;
    }
}

// Givens rotation of step j (j < 64) from the unrotated column H(0..j, j)
// and the norm partials.  Executed by all 64 lanes of one wavefront, which
// compute identical values (lane i loads row i, shuffles broadcast them);
// lane 0 stores the results when `write`.  Returns whether the cycle goes
// on, and hn = ||w_j||.
struct GivensOut {
    bool cont;
    double hn;
};

// Value of v in lane `lane` (wave-uniform index): two v_readlane_b32.
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ GivensOut givens_step(GmresState* __restrict__ st, int j,
                                 const double* __restrict__ part, int G, bool write)
{
    const int lane = threadIdx.x & 63;
    const int m = st->m;
    const double s = sum_partials(part + (size_t)(j + 1) * G, G);
    const double hreg = lane <= j ? st->H[lane * m + j] : 0.0;
    const double creg = lane < j ? st->cs[lane] : 0.0;
    const double sreg = lane < j ? st->sn[lane] : 0.0;
    const double hn = sqrt(s);
    double cur = readlane_d(hreg, 0);
    double tv = 0.0;  // lane i keeps R(i, j): one store per lane after the loop
    for (int i = 0; i < j; ++i) {
        const double c = readlane_d(creg, i), sn = readlane_d(sreg, i);
        const double nxt = readlane_d(hreg, i + 1);
        const double t = c * cur + sn * nxt;
        cur = -sn * cur + c * nxt;
        tv = lane == i ? t : tv;
    }
    if (write && lane < j) st->R[lane * m + j] = tv;
    const double r = hypot(cur, hn);
    const double cj = cur / r, sj = hn / r;
    const double carry = st->carry[j];
    const double gj = cj * carry, cn = -sj * carry;
    const int iters = st->iters0 + j + 1;
    const double stop = fmax(st->rtol * st->bnorm, st->inner_rtol * st->beta);
    const bool cont = !(fabs(cn) <= stop || !(hn > 0.0) || iters >= st->max_iters || j + 1 == m);
    if (write && lane == 0) {
        st->R[j * m + j] = r;
        st->H[(j + 1) * m + j] = hn;
        st->cs[j] = cj;
        st->sn[j] = sj;
        st->g[j] = gj;
        st->carry[j + 1] = cn;
        st->k = j + 1;
        st->iters = iters;
        if (!cont) st->active = 0;
    }
    return GivensOut{cont, hn};
}

// ---------------------------------------------------------------- kernels --
// r = b - A x in FP64 (the refinement residual); W0 = r; partial ||r||^2.
template <typename VT>
__global__ void __launch_bounds__(GM_BS)
k_residual(int n, int rpt, int width, const int* __restrict__ col,
           const double* __restrict__ val, const double* __restrict__ b,
           const double* __restrict__ x, VT* __restrict__ W0, double* __restrict__ part,
           const GmresState* __restrict__ st)
{
    __shared__ double sh[GM_BS / 64];
    if (st->converged) return;
    double acc = 0.0;
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    for (int r = 0; r < rpt; ++r) {
        const int row = base + r * GM_BS;
        if (row >= n) break;
        const double s = ell_row_f64(n, width, row, col, val, x, 1.0, false);
        const double rr = b[row] - s;
        W0[row] = (VT)rr;
        acc += rr * rr;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[lblock()] = acc;
}

__global__ void __launch_bounds__(GM_BS)
k_sumsq(int n, int rpt, const double* __restrict__ b, double* __restrict__ part)
{
    __shared__ double sh[GM_BS / 64];
    double acc = 0.0;
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    for (int r = 0; r < rpt; ++r) {
        const int row = base + r * GM_BS;
        if (row >= n) break;
        acc += b[row] * b[row];
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[lblock()] = acc;
}

__global__ void k_set_bnorm(GmresState* st, const double* __restrict__ part, int G)
{
    const double s = sum_partials(part, G);
    if (threadIdx.x == 0) {
        double bn = sqrt(s);
        st->bnorm = bn > 0.0 ? bn : 1.0;
        st->iters = 0;
        st->cycles = 0;
        st->converged = 0;
        st->active = 0;
        st->k = 0;
    }
}

// Start of a restart cycle from the residual partials.
__global__ void k_cycle_start(GmresState* st, const double* __restrict__ part, int G)
{
    if (st->converged) return;
    const double s = sum_partials(part, G);
    if (threadIdx.x == 0) {
        const double beta = sqrt(s);
        st->beta = beta;
        st->rel = beta / st->bnorm;
        st->k = 0;
        st->iters0 = st->iters;
        if (st->rel <= st->rtol || beta == 0.0) {
            st->converged = 1;
            st->active = 0;
        } else if (st->iters >= st->max_iters) {
            st->active = 0;
        } else {
            st->active = 1;
            st->carry[0] = beta;
            st->cycles += 1;
        }
    }
}

// Arnoldi step j.  For j >= 1 first the Givens rotation of step j-1 (and
// the decision to go on); then v_j = W_src / h (stored), W_dst = A v_j, and
// the partial of <W_dst, v_0> (MGS, ndots = 1) or of <W_dst, v_i> for all
// i <= j (CGS2 first pass, ndots = j + 1).  The norm partials of step j - 1
// are read from nrm slot j: for MGS that is part itself (this kernel writes
// only slot 0 there), for CGS2 a region of its own, since this kernel writes
// part slots 0..j while slower blocks may still be reading the norm (with
// more blocks than fit on the device at once, a late block would otherwise
// read early blocks' new dots in place of the norm).
template <typename VT, typename MT>
__global__ void __launch_bounds__(GM_BS)
k_step_spmv(int n, int rpt, int width, const int* __restrict__ col, const MT* __restrict__ val,
            const VT* __restrict__ Wsrc, VT* __restrict__ V, size_t ldv, VT* __restrict__ Wdst,
            int j, int ndots, double* __restrict__ part, const double* nrm, int G,
            GmresState* __restrict__ st)
{
    __shared__ double sh_h;
    __shared__ int sh_go;
    if (!st->active) return;
    if (j == 0) {
        if (threadIdx.x == 0) {
            sh_h = st->beta;
            sh_go = 1;
        }
    } else if (threadIdx.x < 64) {
        const GivensOut g = givens_step(st, j - 1, nrm, G, lblock() == 0);
        if (threadIdx.x == 0) {
            sh_h = g.hn;
            sh_go = g.cont;
        }
    }
    __syncthreads();
    if (!sh_go) return;
    const VT inv = (VT)(1.0 / sh_h);
    VT* Vj = V + (size_t)j * ldv;
    VT w[GM_RPT], vj[GM_RPT], xo[GM_RPT];
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    // two rows per load group: 133 registers, so three workgroups fit a CU
    // and level 8's 641 workgroups are resident at once (four rows: 174)
    arnoldi_rows<VT, MT, false, 2>(n, rpt, width, base, col, val, Wsrc, inv, xo, w);
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int row = base + r * GM_BS;
        vj[r] = VT(0);
        if (r < rpt && row < n) {
            vj[r] = xo[r] * inv;
            Vj[row] = vj[r];
            Wdst[row] = w[r];
        }
    }
    block_dots<VT>(ndots, j, vj, V, ldv, base, rpt, n, w, part, G, 0);
}

// MGS pass i (1 <= i <= j+1): h_{i-1,j} from the partials, w -= h v_{i-1},
// then the partial of <w, v_i> (i <= j) or of ||w||^2 (i = j + 1).
template <typename VT>
__global__ void __launch_bounds__(GM_BS)
k_mgs(int n, int rpt, VT* __restrict__ w, const VT* __restrict__ V, size_t ldv, int i, int j,
      double* __restrict__ part, int G, GmresState* __restrict__ st)
{
    __shared__ double sh[GM_BS / 64];
    // every load of the kernel is issued before the first use: the flag, the
    // G partials and the row data are independent round trips
    const int active = st->active;
    const double h = sum_partials(part + (size_t)(i - 1) * G, G);
    const VT* Vp = V + (size_t)(i - 1) * ldv;
    const VT* Vi = V + (size_t)i * ldv;
    const bool last = (i == j + 1);
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    VT wv[GM_RPT], pv[GM_RPT], iv[GM_RPT];
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int row = base + r * GM_BS;
        const bool in = r < rpt && row < n;
        wv[r] = in ? w[row] : VT(0);
        pv[r] = in ? Vp[row] : VT(0);
        iv[r] = (in && !last) ? Vi[row] : VT(0);
    }
    if (!active) return;
    if (lblock() == 0 && threadIdx.x == 0) st->H[(i - 1) * st->m + j] = h;
    const VT hv = (VT)h;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int row = base + r * GM_BS;
        if (r < rpt && row < n) {
            const VT wr = wv[r] - hv * pv[r];
            w[row] = wr;
            const double o = last ? to_d(wr) : to_d(iv[r]);
            acc += to_d(wr) * o;
        }
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(size_t)i * G + lblock()] = acc;
}


// CGS2 update: hc_i = the sum of the G partials of dot i (i = 0..j, read
// from pin; every block sums them itself, in the same fixed order, so the
// pass needs no separate reduction launch; block 0 records H(i, j)), then
// w -= sum_i hc_i v_i; then partial dots with all v_i (first pass, to pout)
// or the partial of ||w||^2 (final pass, to pout slot j + 1).  pin and pout
// are different regions: blocks still read pin while others write pout.
template <typename VT>
__global__ void __launch_bounds__(GM_BS)
k_cgs_update(int n, int rpt, VT* __restrict__ w, const VT* __restrict__ V, size_t ldv, int j,
             int final_pass, const double* __restrict__ pin, double* __restrict__ pout, int G,
             GmresState* __restrict__ st)
{
    __shared__ double hs[GM_MAXM + 1];
    __shared__ double hw[GM_BS / 64][GM_MAXM + 1];
    if (!st->active) return;
    VT wv[GM_RPT];
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int row = base + r * GM_BS;
        wv[r] = (r < rpt && row < n) ? w[row] : VT(0);
    }
    {
        // thread t adds partials t, t + 256, ... of 16 dots at a time (the
        // 16 loads issued first), then per dot a wavefront sum; below, the
        // four wavefronts in order: the same order in every block
        const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int i0 = 0; i0 <= j; i0 += 16) {
            double acc[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[q] = 0.0;
            for (int g0 = 0; g0 < G; g0 += GM_BS) {
                const int g = g0 + (int)threadIdx.x;
                double v[16];
#pragma unroll
                for (int q = 0; q < 16; ++q)
                    v[q] = (i0 + q <= j && g < G) ? pin[(size_t)(i0 + q) * G + g] : 0.0;
#pragma unroll
                for (int q = 0; q < 16; ++q) acc[q] += v[q];
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                if (i0 + q > j) break;
                const double t = wave_sum(acc[q]);
                if (lane == 0) hw[wid][i0 + q] = t;
            }
        }
        __syncthreads();
        if ((int)threadIdx.x <= j) {
            const int i = threadIdx.x;
            double t = 0.0;
            for (int ww = 0; ww < GM_BS / 64; ++ww) t += hw[ww][i];
            hs[i] = t;
            if (lblock() == 0) {
                st->hc[i] = t;
                if (!final_pass) st->H[i * st->m + j] = t;
                else st->H[i * st->m + j] += t;
            }
        }
        __syncthreads();
    }
    // w -= sum_i hc_i v_i in order i = 0..j, basis loads batched
    for (int i0 = 0; i0 <= j; i0 += GM_LB) {
        VT vb[GM_LB][GM_RPT];
#pragma unroll
        for (int q = 0; q < GM_LB; ++q)
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) {
                const int row = base + r * GM_BS;
                vb[q][r] = (i0 + q <= j && r < rpt && row < n)
                               ? V[(size_t)(i0 + q) * ldv + row] : VT(0);
            }
#pragma unroll
        for (int q = 0; q < GM_LB; ++q) {
            if (i0 + q > j) break;
            const VT h = (VT)hs[i0 + q];
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) wv[r] = wv[r] - h * vb[q][r];
        }
    }
#pragma unroll
    for (int r = 0; r < GM_RPT; ++r) {
        const int row = base + r * GM_BS;
        if (r < rpt && row < n) w[row] = wv[r];
    }
    if (final_pass) block_dots<VT>(0, -1, wv, V, ldv, base, rpt, n, wv, pout, G, j + 1);
    else block_dots<VT>(j + 1, -1, wv, V, ldv, base, rpt, n, wv, pout, G, 0);
}

// End of a cycle: the pending Givens rotation (if the cycle ran to its last
// step) and y = R^{-1} g, one block.
__global__ void __launch_bounds__(GM_BS)
k_solve_y(const double* __restrict__ nrm, int G, GmresState* __restrict__ st)
{
    __shared__ double Rs[GM_MAXM * GM_MAXM];
    __shared__ double gs[GM_MAXM];
    if (st->converged) return;
    if (st->active && threadIdx.x < 64) givens_step(st, st->k, nrm, G, true);
    __syncthreads();
    const int k = st->k, m = st->m;
    // stage R and g in LDS with all threads (parallel loads), then solve
    for (int e = threadIdx.x; e < k * k; e += blockDim.x) Rs[e] = st->R[(e / k) * m + (e % k)];
    for (int e = threadIdx.x; e < k; e += blockDim.x) gs[e] = st->g[e];
    __syncthreads();
    if (threadIdx.x != 0) return;
    st->active = 0;
    double y[GM_MAXM];
    for (int i = k - 1; i >= 0; --i) {
        double s = gs[i];
        for (int l = i + 1; l < k; ++l) s -= Rs[i * k + l] * y[l];
        y[i] = s / Rs[i * k + i];
    }
    for (int i = 0; i < k; ++i) st->y[i] = y[i];
}

// x += sum_i y_i v_i, accumulated in FP64.
template <typename VT>
__global__ void __launch_bounds__(GM_BS)
k_xupdate(int n, int rpt, const VT* __restrict__ V, size_t ldv, double* __restrict__ x,
          const GmresState* __restrict__ st)
{
    __shared__ double ys[GM_MAXM];
    if (st->converged) return;
    const int k = st->k;
    if (k == 0) return;
    if ((int)threadIdx.x < k) ys[threadIdx.x] = st->y[threadIdx.x];
    __syncthreads();
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    for (int r = 0; r < rpt; ++r) {
        const int row = base + r * GM_BS;
        if (row >= n) break;
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += to_d(V[(size_t)i * ldv + row]) * ys[i];
        x[row] += s;
    }
}

__global__ void k_fill0(int n, double* __restrict__ x)
{
    const int i = lblock() * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// ------------------------------------------------- persistent MGS cycle --
// One cooperative launch per restart cycle: every workgroup is resident, and
// the global sums of the MGS passes go through device memory instead of
// kernel boundaries.  Exchange e of the cycle owns row e of xch (G 64-bit
// slots, set to XCH_EMPTY before the launch): each workgroup stores its
// block partial into its slot with an agent-scope relaxed atomic store, and
// every workgroup waits until no slot of the row holds XCH_EMPTY, then adds
// the G partials in sum_partials' fixed order.  A row is written once per
// launch, so no slot is reused while another workgroup may still read it.
// The vectors other workgroups gather (the next Arnoldi vector, x before the
// residual) are written with agent-scope atomic stores that complete
// (s_waitcnt vmcnt(0)) before the partial announcing them is stored, and are
// gathered with agent-scope atomic loads; no cache-wide writeback or
// invalidation is needed.  Every wait is bounded: after XCH_SPIN polls a
// workgroup raises GmresState::abort and leaves, and the others leave at
// their next poll; the host reports the abort as an error.
//
// Each workgroup keeps the cycle's Givens state (current Hessenberg column,
// rotations, rotated right-hand side, R) in LDS and computes it from the
// same exchanged sums as every other workgroup; at the cycle's end workgroup
// 0 records y, the step count and the next cycle's start in GmresState (no
// global stores during the steps: a workgroup's wait for its loads would
// also wait for them, and every exchange waits for the slowest workgroup).  The arithmetic is that of k_step_spmv, k_mgs, k_solve_y,
// k_xupdate, k_residual and k_cycle_start, in the same order over the same
// block partials, so the solve gives the same bits as the kernel-per-pass
// form.
typedef unsigned long long xch_t;
constexpr xch_t XCH_EMPTY = 0x7FF4DEADBEEF0001ull;  // a signalling-NaN pattern no partial takes
constexpr long XCH_SPIN = 1L << 22;                  // polls before a wait gives up
// (PHG_XCH_SPIN overrides it: the failure-path test forces a give-up with
// a limit of a few polls; every workgroup still exits normally)

// Wave 0: once every slot of the row is written, the sum of its G partials
// in sum_partials' order (the same value in every lane).  false when the
// wait gives up or another workgroup gave up.
__device__ bool xch_sum(const xch_t* row, int G, int* abort_flag, long spin, double* out)
{
    const int lane = threadIdx.x & 63;
    for (long it = 0; it < spin; ++it) {
        double v[GM_MAXG / 64];
        bool all = true;
#pragma unroll
        for (int k = 0; k < GM_MAXG / 64; ++k) {
            const int b = lane + 64 * k;
            xch_t u = 0;  // the bits of +0.0, as sum_partials adds beyond G
            if (b < G) u = ld_agent(row + b);
            all = all && (u != XCH_EMPTY);
            v[k] = __longlong_as_double((long long)u);
        }
        if (__all(all)) {
            *out = wave_sum(lane_partials(v, G));
            return true;
        }
        if ((it & 63) == 63 && ld_agent(abort_flag)) return false;
    }
    if (lane == 0) st_agent(abort_flag, 1);
    return false;
}

struct XchLds {
    double val;
    int ok;
};

// All threads: the workgroup's partial (block_sum of acc: wave sums, then
// the four waves in order) into its slot of row, and the row's sum in *sum.
// With publish, every thread's earlier stores complete first (the vector
// the partial announces); without it, loads already issued (the next
// pass's basis rows) stay in flight while wave 0 waits.
__device__ __forceinline__ bool xch_block(xch_t* row, int G, double acc, bool publish,
                                          int* abort_flag, long spin, double* sum, double* sh,
                                          XchLds* l)
{
    const double v = wave_sum(acc);
    if (publish) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this thread's stores are complete
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        if (threadIdx.x == 0) {
            double r = 0.0;
            for (int q = 0; q < GM_BS / 64; ++q) r += sh[q];
            st_agent(row + lblock(), (xch_t)__double_as_longlong(r));
        }
        double s = 0.0;
        const bool ok = xch_sum(row, G, abort_flag, spin, &s);
        if (threadIdx.x == 0) {
            l->val = s;
            l->ok = ok;
        }
    }
    __syncthreads();
    *sum = l->val;
    return l->ok;
}

// givens_step on the LDS copy of the cycle's state: same arithmetic, stored
// in LDS only (a global store here would hold up this workgroup's next wait
// for its loads, since the load counter also counts stores).
__device__ GivensOut givens_lds(int j, double s, int m, double stop, int iters0, int max_iters,
                                const double* sH, double* s_cs, double* s_sn, double* s_g,
                                double* s_carry, double* s_R)
{
    const int lane = threadIdx.x & 63;
    const double hreg = lane <= j ? sH[lane] : 0.0;
    const double creg = lane < j ? s_cs[lane] : 0.0;
    const double sreg = lane < j ? s_sn[lane] : 0.0;
    const double hn = sqrt(s);
    double cur = readlane_d(hreg, 0);
    double tv = 0.0;  // lane i keeps R(i, j): one store per lane after the loop
    for (int i = 0; i < j; ++i) {
        const double c = readlane_d(creg, i), sn = readlane_d(sreg, i);
        const double nxt = readlane_d(hreg, i + 1);
        const double t = c * cur + sn * nxt;
        cur = -sn * cur + c * nxt;
        tv = lane == i ? t : tv;
    }
    if (lane < j) s_R[lane * m + j] = tv;
    const double r = hypot(cur, hn);
    const double cj = cur / r, sj = hn / r;
    const double carry = s_carry[j];
    const double gj = cj * carry, cn = -sj * carry;
    const int iters = iters0 + j + 1;
    const bool cont = !(fabs(cn) <= stop || !(hn > 0.0) || iters >= max_iters || j + 1 == m);
    if (lane == 0) {
        s_R[j * m + j] = r;
        s_cs[j] = cj;
        s_sn[j] = sj;
        s_g[j] = gj;
        s_carry[j + 1] = cn;
    }
    return GivensOut{cont, hn};
}

// One restart cycle (Arnoldi steps with MGS, y, x += V y, r = b - A x and
// the decision of the next cycle) as one cooperative launch of G workgroups.
template <typename VT, typename MT>
__global__ void __launch_bounds__(GM_BS)
k_mgs_cycle(int n, int rpt, int width, const int* __restrict__ col, const MT* __restrict__ val,
            int width64, const int* __restrict__ col64, const double* __restrict__ val64,
            const double* __restrict__ b, double* x, VT* V, size_t ldv, VT* W0, VT* W1,
            xch_t* xch, int G, long spin, GmresState* __restrict__ st)
{
    __shared__ double sh[GM_BS / 64];
    __shared__ XchLds xl;
    __shared__ double sH[GM_MAXM + 1];
    __shared__ double s_cs[GM_MAXM], s_sn[GM_MAXM], s_g[GM_MAXM], s_carry[GM_MAXM + 1];
    __shared__ double s_R[GM_MAXM * GM_MAXM];
    __shared__ double s_y[GM_MAXM];
    __shared__ double s_hn;
    __shared__ int s_cont;
    if (!st->active) return;  // the previous launch's decision, the same for all
    int* abort_flag = &st->abort;
    const int m = st->m, iters0 = st->iters0, max_iters = st->max_iters;
    const double stop = fmax(st->rtol * st->bnorm, st->inner_rtol * st->beta);
    const bool rec = lblock() == 0;
    if (threadIdx.x == 0) s_carry[0] = st->carry[0];
    const int base = lblock() * GM_BS * rpt + threadIdx.x;
    xch_t* row = xch;
    double h = st->beta;
    int k = 0;
    bool ok = true;
    const bool pre = width >= 1 && width <= ELL_MAXW;
    RowSlots<MT> slots;
    if (pre) load_slots<MT>(n, width, base, col, val, slots);
    VT v0r[GM_RPT], v1r[GM_RPT];  // this thread's rows of v_0 and v_1, kept for the cycle
    for (int j = 0; j < m; ++j) {
        VT* Wsrc = (j & 1) ? W1 : W0;
        VT* Wdst = (j & 1) ? W0 : W1;
        const VT inv = (VT)(1.0 / h);
        VT* Vj = V + (size_t)j * ldv;
        VT w[GM_RPT], vj[GM_RPT], xo[GM_RPT];
        if (pre) arnoldi_rows_slots<VT, MT>(n, rpt, width, base, slots, Wsrc, inv, xo, w);
        else arnoldi_rows<VT, MT, true>(n, rpt, width, base, col, val, Wsrc, inv, xo, w);
        // pv = v_{i-1}, iv = v_i of the coming pass i (own rows; v_0 and v_1
        // from registers, kept since steps 0 and 1, so the first two passes
        // wait for no load); pass i + 1 reuses iv as its pv, so each later
        // pass loads one basis vector, issued before the exchange it does not
        // depend on
        VT pv[GM_RPT], iv[GM_RPT];
#pragma unroll
        for (int r = 0; r < GM_RPT; ++r) {
            const int rr = base + r * GM_BS;
            vj[r] = VT(0);
            if (r < rpt && rr < n) {
                vj[r] = xo[r] * inv;
                Vj[rr] = vj[r];
            }
            if (j == 0) v0r[r] = vj[r];
            if (j == 1) v1r[r] = vj[r];
            pv[r] = v0r[r];
            iv[r] = j >= 1 ? v1r[r] : VT(0);
        }
        double acc = 0.0;
        // partial of <w, v_0> (block_dots with one dot)
#pragma unroll
        for (int r = 0; r < GM_RPT; ++r) {
            const int rr = base + r * GM_BS;
            if (r < rpt && rr < n) acc += to_d(w[r]) * to_d(pv[r]);
        }
        double hsum = 0.0;
        ok = xch_block(row, G, acc, false, abort_flag, spin, &hsum, sh, &xl);
        row += G;
        if (!ok) break;
        if (threadIdx.x == 0) sH[0] = hsum;
        // MGS passes i = 1..j+1: w -= h_{i-1} v_{i-1}, then <w, v_i> or ||w||^2
        bool cont = false;
        for (int i = 1; i <= j + 1 && ok; ++i) {
            const bool last = (i == j + 1);
            const VT hv = (VT)hsum;
            double a2 = 0.0;
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) {
                const int rr = base + r * GM_BS;
                if (r < rpt && rr < n) {
                    const VT wr = w[r] - hv * pv[r];
                    w[r] = wr;
                    const double o = last ? to_d(wr) : to_d(iv[r]);
                    a2 += to_d(wr) * o;
                    if (last) st_agent(Wdst + rr, wr);
                }
            }
            // v_{i+1} for the next pass
            VT nv[GM_RPT];
#pragma unroll
            for (int r = 0; r < GM_RPT; ++r) {
                const int rr = base + r * GM_BS;
                nv[r] = (r < rpt && rr < n && i + 1 <= j)
                            ? (i + 1 == j ? vj[r] : V[(size_t)(i + 1) * ldv + rr]) : VT(0);
            }
            double s = 0.0;
            ok = xch_block(row, G, a2, last, abort_flag, spin, &s, sh, &xl);
            row += G;
            if (!ok) break;
            if (!last) {
                hsum = s;
                if (threadIdx.x == 0) sH[i] = s;
#pragma unroll
                for (int r = 0; r < GM_RPT; ++r) {
                    pv[r] = iv[r];
                    iv[r] = nv[r];
                }
            } else {
                if (threadIdx.x < 64) {
                    const GivensOut g = givens_lds(j, s, m, stop, iters0, max_iters, sH, s_cs, s_sn,
                                                   s_g, s_carry, s_R);
                    if (threadIdx.x == 0) {
                        s_hn = g.hn;
                        s_cont = g.cont;
                    }
                }
                __syncthreads();
                h = s_hn;
                cont = s_cont;
            }
        }
        if (!ok) break;
        k = j + 1;
        if (!cont) break;
    }
    if (!ok) return;
    // y = R^-1 g (k_solve_y), in every workgroup
    if (threadIdx.x == 0) {
        for (int i = k - 1; i >= 0; --i) {
            double s = s_g[i];
            for (int l = i + 1; l < k; ++l) s -= s_R[i * m + l] * s_y[l];
            s_y[i] = s / s_R[i * m + i];
        }
        if (rec) {
            for (int i = 0; i < k; ++i) st->y[i] = s_y[i];
            st->k = k;
            st->iters = iters0 + k;
            st->active = 0;
        }
    }
    __syncthreads();
    // x += V y (k_xupdate), stored for the residual's gathers
    for (int r = 0; r < rpt; ++r) {
        const int rr = base + r * GM_BS;
        if (rr >= n) break;
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += to_d(V[(size_t)i * ldv + rr]) * s_y[i];
        st_agent(x + rr, x[rr] + s);
    }
    double dummy = 0.0;
    ok = xch_block(row, G, 0.0, true, abort_flag, spin, &dummy, sh, &xl);
    row += G;
    if (!ok) return;
    // r = b - A x (k_residual), W0 = r, ||r||^2
    double acc = 0.0;
    for (int r = 0; r < rpt; ++r) {
        const int rr = base + r * GM_BS;
        if (rr >= n) break;
        const double sx = ell_row_f64_agent(n, width64, rr, col64, val64, x, 1.0, false);
        const double res = b[rr] - sx;
        W0[rr] = (VT)res;
        acc += res * res;
    }
    double s2 = 0.0;
    ok = xch_block(row, G, acc, false, abort_flag, spin, &s2, sh, &xl);
    if (!ok) return;
    // the next cycle's start (k_cycle_start), recorded by workgroup 0
    if (rec && threadIdx.x == 0) {
        const double beta = sqrt(s2);
        st->beta = beta;
        st->rel = beta / st->bnorm;
        st->k = 0;
        st->iters0 = st->iters;
        if (st->rel <= st->rtol || beta == 0.0) {
            st->converged = 1;
            st->active = 0;
        } else if (st->iters >= st->max_iters) {
            st->active = 0;
        } else {
            st->active = 1;
            st->carry[0] = beta;
            st->cycles += 1;
        }
    }
}

// Sets the first `count` exchange slots to XCH_EMPTY and clears the abort flag.
__global__ void k_xch_reset(xch_t* xch, size_t count, GmresState* st)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) xch[i] = XCH_EMPTY;
    if (i == 0) st->abort = 0;
}

// Exchange rows one cycle of restart m uses: j + 2 per Arnoldi step, one
// before and one after the residual.
inline int xch_rows_for(int m) { return m * (m + 3) / 2 + 2; }

// Set once a cooperative launch was refused by the runtime: later solves of
// the process take the kernel-per-pass form directly.
inline std::atomic<bool>& coop_refused()
{
    static std::atomic<bool> refused{false};
    return refused;
}

// Whether every workgroup of the cycle kernel can be resident at once.
template <typename VT, typename MT>
bool coop_fits(int G)
{
    // per template instance, computed once (thread-safe static initialisation)
    static const int capacity = [] {
        int dev = 0, cus = 0, coop = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, (const void*)k_mgs_cycle<VT, MT>, GM_BS, 0) != hipSuccess)
            return 0;
        return coop ? cus * per : 0;
    }();
    return G <= capacity;
}

// ---------------------------------------------------------------- driver --
int prepare_work(phg_matrix_s* A, int prec, int m, int ortho, GmresWork** out)
{
    const bool mgs = ortho == PHG_MGS;
    auto& slot = A->work[2 * prec + (mgs ? 0 : 1)];
    const int n = A->n_rows;
    if (!slot || slot->m < m) {
        std::unique_ptr<GmresWork> w(new GmresWork());
        w->n = n;
        w->m = m;
        w->prec = prec;
        // target block count.  MGS: the fewest blocks that hold the rows at
        // GM_RPT rows per thread (161 at level 7): each of an Arnoldi step's
        // j + 2 dependent sums then gathers fewer partials, and the
        // persistent cycle ran 35.3 instead of 33.3 solves/s at level 7
        // (the kernel-per-pass form, which shares the shape so that both
        // give the same bits, 32.2 instead of 33.6).  CGS2: 256, one per
        // CU, since its passes stream the basis.  PHG_GMRES_BLOCKS overrides
        // both (launch-shape experiments; it changes the grouping of the
        // dot partials).
        int target = mgs ? std::max(1, (n + GM_BS * GM_RPT - 1) / (GM_BS * GM_RPT)) : 256;
        if (const char* e = getenv("PHG_GMRES_BLOCKS")) target = std::max(1, std::min(atoi(e), GM_MAXG));
        int G = std::max(1, std::min((n + GM_BS - 1) / GM_BS, target));
        int rpt = (n + G * GM_BS - 1) / (G * GM_BS);
        if (rpt > GM_RPT) {
            rpt = GM_RPT;
            G = (n + GM_BS * GM_RPT - 1) / (GM_BS * GM_RPT);
        }
        if (G > GM_MAXG) {
            set_error("gmres: matrix too large for one device workspace");
            return PHG_ERR_ARG;
        }
        w->G = G;
        w->rpt = std::max(rpt, 1);
        w->ldv = ((size_t)n + 63) / 64 * 64;
        const size_t es = prec == PHG_F64 ? 8 : 4;
        PHG_HIP(w->V.alloc((size_t)(m + 1) * w->ldv * es));
        PHG_HIP(w->W.alloc(2 * w->ldv * es));
        PHG_HIP(w->part.alloc((size_t)3 * (m + 2) * G));  // regions A, B, C (CGS2 passes)
        PHG_HIP(w->st.alloc(1));
        PHG_HIP(hipHostMalloc((void**)&w->host, sizeof(GmresState), hipHostMallocDefault));
        slot.reset(w.release());
    }
    *out = slot.get();
    return PHG_OK;
}

template <typename VT, typename MT>
void enqueue_cycle(phg_matrix_s* A, const Ell<MT>& E, GmresWork* w, const double* b, double* x,
                   int m, bool cgs, hipStream_t s)
{
    const int n = w->n, G = w->G, rpt = w->rpt;
    const size_t ldv = w->ldv;
    VT* V = (VT*)w->V.p;
    VT* W0 = (VT*)w->W.p;
    VT* W1 = W0 + ldv;
    double* part = w->part.p;
    // ||w||^2 partials: MGS leaves them in part (slot j + 1), CGS2 in region C
    double* nrm = cgs ? part + (size_t)2 * (w->m + 2) * G : part;
    GmresState* st = w->st.p;
    const dim3 grid(G), block(GM_BS);
    for (int j = 0; j < m; ++j) {
        VT* src = (j & 1) ? W1 : W0;
        VT* dst = (j & 1) ? W0 : W1;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_step_spmv<VT, MT>), grid, block, 0, s, n, rpt,
                           E.width, E.col.p, E.val.p, src, V, ldv, dst, j, cgs ? j + 1 : 1,
                           part, nrm, G, st);
        if (!cgs) {
            for (int i = 1; i <= j + 1; ++i)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mgs<VT>), grid, block, 0, s, n, rpt, dst, V,
                                   ldv, i, j, part, G, st);
        } else {
            // first pass: dots from the step kernel (region A) -> second-pass
            // dots (region B); second pass: region B -> ||w||^2 (region C)
            double* partB = part + (size_t)(w->m + 2) * G;
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cgs_update<VT>), grid, block, 0, s, n, rpt, dst,
                               V, ldv, j, 0, part, partB, G, st);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cgs_update<VT>), grid, block, 0, s, n, rpt, dst,
                               V, ldv, j, 1, partB, nrm, G, st);
        }
    }
    hipLaunchKernelGGL(k_solve_y, dim3(1), dim3(GM_BS), 0, s, nrm, G, st);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_xupdate<VT>), grid, block, 0, s, n, rpt, V, ldv, x, st);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_residual<VT>), grid, block, 0, s, n, rpt,
                       A->ell64.width, A->ell64.col.p, A->ell64.val.p, b, x, W0, part, st);
    hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(64), 0, s, st, part, G);
}

template <typename VT, typename MT>
int gmres_run(phg_matrix_s* A, const Ell<MT>& E, GmresWork* w, const double* b, double* x,
              const phg_gmres_opts& o, phg_gmres_info* info, hipStream_t s)
{
    const int n = w->n, m = o.restart, G = w->G, rpt = w->rpt;
    VT* W0 = (VT*)w->W.p;
    double* part = w->part.p;
    GmresState* st = w->st.p;
    const dim3 grid(G), block(GM_BS);
    const bool cgs = o.ortho == PHG_CGS2;
    // MGS cycles as one cooperative kernel when every workgroup fits at once
    bool coop = !cgs && o.persistent && !coop_refused().load() && coop_fits<VT, MT>(G);
    if (coop && w->xch_rows < xch_rows_for(m)) {
        PHG_HIP(w->xch.alloc((size_t)xch_rows_for(m) * G));
        w->xch_rows = xch_rows_for(m);
    }
    if (coop && !w->ev[0]) {
        PHG_HIP(hipEventCreate(&w->ev[0]));
        PHG_HIP(hipEventCreate(&w->ev[1]));
    }
    int launches = 0;
    double cycle_ms = 0.0;

    // parameters of the solve
    GmresState init;
    memset(&init, 0, sizeof init);
    init.m = m;
    init.rtol = o.rtol;
    init.inner_rtol = (o.precision == PHG_F64) ? 0.0 : o.inner_rtol;
    init.max_iters = o.max_iters;
    PHG_HIP(hipMemcpyAsync(st, &init, sizeof init, hipMemcpyHostToDevice, s));
    if (!o.use_x0) hipLaunchKernelGGL(k_fill0, dim3((n + 255) / 256), dim3(256), 0, s, n, x);
    hipLaunchKernelGGL(k_sumsq, grid, block, 0, s, n, rpt, b, part);
    hipLaunchKernelGGL(k_set_bnorm, dim3(1), dim3(64), 0, s, st, part, G);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_residual<VT>), grid, block, 0, s, n, rpt,
                       A->ell64.width, A->ell64.col.p, A->ell64.val.p, b, x, W0, part, st);
    hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(64), 0, s, st, part, G);
    PHG_HIP(hipGetLastError());

    // the captured cycle; replayed only when built (or found current) here
    bool graph_ready = false;
    if (o.graph && !coop) {
        graph_ready = true;
        const void* key[4] = {b, x, E.val.p, A->ell64.val.p};
        const bool same = w->exec && w->key_m == m && w->key_ortho == o.ortho &&
                          !memcmp(key, w->key, sizeof key);
        if (!same) {
            if (w->exec) {
                PHG_HIP(hipGraphExecDestroy(w->exec));
                w->exec = nullptr;
            }
            // recorded on the caller's stream; the null stream cannot be
            // captured, so for it on a stream of the workspace (a stream
            // per workspace would also take a share of the process's few
            // hardware queues from the caller's own streams)
            hipStream_t cs = s;
            if (!cs) {
                if (!w->cap) PHG_HIP(hipStreamCreateWithFlags(&w->cap, hipStreamNonBlocking));
                cs = w->cap;
            }
            hipGraph_t g = nullptr;
            PHG_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            enqueue_cycle<VT, MT>(A, E, w, b, x, m, cgs, cs);
            PHG_HIP(hipStreamEndCapture(cs, &g));
            hipError_t e = hipGraphInstantiate(&w->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            PHG_HIP(e);
            memcpy(w->key, key, sizeof key);
            w->key_m = m;
            w->key_ortho = o.ortho;
        }
    }

    for (;;) {
        PHG_HIP(hipMemcpyAsync(w->host, st, sizeof(GmresState), hipMemcpyDeviceToHost, s));
        PHG_HIP(hipStreamSynchronize(s));
        if (coop && launches > 0) {
            float ms = 0.0f;
            PHG_HIP(hipEventElapsedTime(&ms, w->ev[0], w->ev[1]));
            cycle_ms += ms;
        }
        const GmresState& h = *w->host;
        if (h.abort) {
            set_error("gmres: a workgroup of the persistent MGS cycle timed out waiting for "
                      "the others");
            return PHG_ERR_HIP;
        }
        if (h.converged || !h.active) break;
        if (coop) {
            const size_t cnt = (size_t)xch_rows_for(m) * G;
            hipLaunchKernelGGL(k_xch_reset, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                               w->xch.p(), cnt, st);
            PHG_HIP(hipGetLastError());
            int n_ = n, rpt_ = rpt, width = E.width, width64 = A->ell64.width, G_ = G;
            const int* col = E.col.p;
            const MT* val = E.val.p;
            const int* col64 = A->ell64.col.p;
            const double* val64 = A->ell64.val.p;
            VT* V = (VT*)w->V.p;
            VT* W1 = W0 + w->ldv;
            size_t ldv = w->ldv;
            xch_t* xch = w->xch.p();
            long spin = XCH_SPIN;
            if (const char* e = getenv("PHG_XCH_SPIN")) spin = std::max(1L, atol(e));
            void* args[] = {&n_, &rpt_, &width, &col, &val, &width64, &col64, &val64, &b, &x,
                            &V, &ldv, &W0, &W1, &xch, &G_, &spin, &st};
            PHG_HIP(hipEventRecord(w->ev[0], s));
            const hipError_t e = hipLaunchCooperativeKernel((const void*)k_mgs_cycle<VT, MT>,
                                                            dim3(G), dim3(GM_BS), args, 0, s);
            if (e == hipSuccess) {
                PHG_HIP(hipEventRecord(w->ev[1], s));
                ++launches;
            } else {
                // a runtime that refuses the cooperative launch here: the
                // kernel-per-pass form (same bits) for this solve and, from
                // now on, for this process
                (void)hipGetLastError();
                coop_refused().store(true);
                coop = false;
                enqueue_cycle<VT, MT>(A, E, w, b, x, m, cgs, s);
                PHG_HIP(hipGetLastError());
            }
        } else if (graph_ready) {
            PHG_HIP(hipGraphLaunch(w->exec, s));
        } else {
            enqueue_cycle<VT, MT>(A, E, w, b, x, m, cgs, s);
            PHG_HIP(hipGetLastError());
        }
    }
    const GmresState& h = *w->host;
    if (info) {
        info->iterations = h.iters;
        info->restarts = h.cycles;
        info->refinements = o.precision == PHG_F64 ? 0 : h.cycles;
        info->rel_residual = h.rel;
        info->converged = h.converged;
        info->cycle_launches = launches;
        info->cycle_ms = cycle_ms;
    }
    return h.converged ? PHG_OK : PHG_ERR_NOCONV;
}

}  // namespace

int gmres_solve(phg_matrix_s* A, const double* b, double* x, const phg_gmres_opts* opts,
                phg_gmres_info* info, hipStream_t s)
{
    phg_gmres_opts o = *opts;
    if (o.restart < 1 || o.restart > GM_MAXM) {
        set_error("gmres: restart must be in 1..64");
        return PHG_ERR_ARG;
    }
    if (o.ortho != PHG_MGS && o.ortho != PHG_CGS2) {
        set_error("gmres: ortho must be PHG_MGS or PHG_CGS2");
        return PHG_ERR_ARG;
    }
    if (A->n_rows != A->n_cols) {
        set_error("gmres: matrix is not square");
        return PHG_ERR_ARG;
    }
    if (A->n_rows == 0) {
        if (info) memset(info, 0, sizeof *info), info->converged = 1;
        return PHG_OK;
    }
    int rc = ensure_ell64(A, s);
    if (rc) return rc;
    GmresWork* w = nullptr;
    rc = prepare_work(A, o.precision, o.restart, o.ortho, &w);
    if (rc) return rc;
    if (o.precision == PHG_F64) return gmres_run<double, double>(A, A->ell64, w, b, x, o, info, s);
    if (o.precision == PHG_F32) {
        rc = ensure_ell32(A, s);
        if (rc) return rc;
        return gmres_run<float, float>(A, A->ell32, w, b, x, o, info, s);
    }
    rc = ensure_ell16(A, s);
    if (rc) return rc;
    return gmres_run<float, uint16_t>(A, A->ell16, w, b, x, o, info, s);
}

}  // namespace phg
