// Streaming vector kernels, deterministic reductions and the device-side CG
// scalar recurrences of include/solverCG.h:61-117.
//
// Every reduction is two-level: each block writes one partial per quantity to
// a fixed slot, then a single block sums the partials in a fixed order.  The
// launch geometry of every streaming kernel is fixed per problem size, so all
// sums are bitwise reproducible from run to run.
#pragma once
#include "common.h"

namespace fans {

// device scalar slots (double) -- layout shared with the host
enum {
    SC_BETA     = 0,
    SC_ALPHA    = 1,
    SC_DELTA    = 2,
    SC_DELTA0   = 3,
    SC_DELTAMID = 4,
    SC_ERR      = 5,
    SC_ERR0     = 6,
    SC_DN       = 7,
    SC_TOL      = 8,
    SC_RES0     = 16,  // generic reduction results [16, 32)
    SC_ARING    = 32,  // linear CG: alpha of iteration j at SC_ARING + j % kDirRing
    SC_COUNT    = kScSlot
};
// linear CG: search directions kept per iteration before the solution update
// u -= sum alpha_j d_j is applied in one pass
constexpr int kDirRing = 8;
// device integer slots: IS_* (common.h)

enum {
    RED_ERR_INIT = 0,  // q0 = norm partials      -> err0, iter = 0, delta = 1, deltamid = 0
    RED_DELTA    = 1,  // q0 = <r,s> partials     -> delta, beta
    RED_ALPHA    = 2,  // q0 = <d,rnew>, q1 = <d,rnew_prev> (nq 2) -> alpha (and deltamid)
    RED_ERR_NEXT = 4,  // q0 = norm partials      -> err, iter++
    RED_SUM      = 5,  // q0..q(nq-1) sums        -> SC_RES0 + q
    RED_ERR_SET  = 6   // q0 = norm partials      -> err, stop test (iteration and CG scalars kept)
};

__device__ __forceinline__ double finish_norm(double v, int measure) { return measure == 1 ? sqrt(v) : v; }

// The scalars a reduction stage reads, loaded before its first store (the
// compiler cannot tell the stores to sc / is / hist from those loads, so it
// would keep each load behind the stores before it: one round trip each).
struct StageScalars {
    int    iter, errtype, nit;
    double delta, deltamid, beta, err0, tol;
};
__device__ __forceinline__ StageScalars load_stage_scalars(const double *sc, const int *is)
{
    return StageScalars{is[IS_ITER],     is[IS_ERRTYPE], is[IS_NIT], sc[SC_DELTA], sc[SC_DELTAMID],
                        sc[SC_BETA],     sc[SC_ERR0],    sc[SC_TOL]};
}

// records the error of iteration it and sets done if the loop condition of
// solverCG.h:84 / solverFP.h:221 fails
__device__ __forceinline__ void record_error(double *sc, int *is, double *hist, const StageScalars &S, int it,
                                             double err)
{
    hist[it] = err;
    if (it == 0)
        sc[SC_ERR0] = err;
    sc[SC_ERR]           = err;
    const double err_rel = it == 0 ? 100.0 : err / S.err0;
    const double e       = S.errtype == 0 ? err : err_rel;
    is[IS_DONE]          = !(it < S.nit && e > S.tol);
}

// one block of kReduceThreads; each thread folds its strided partials into
// four independent accumulators (fixed order) before the block reduction
constexpr int kReduceThreads = 1024;

// partial_b (nullable): quantity 1 comes from partial_b[0, nblk_b) instead of
// partial[nblk, 2 nblk) (two producing kernels of different grids)
// One block per slot of the launch (grid z); every buffer at its slot's offset.
__global__ __launch_bounds__(kReduceThreads) void k_reduce(const double *__restrict__ partial, int nblk, int nq,
                                                           int stage, double *sc, int *is, double *hist,
                                                           const double *__restrict__ partial_b, int nblk_b,
                                                           Slots sl)
{
    __shared__ double sm[2 * (kReduceThreads / 64)];
    const int         sb = slot_of(sl, false);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        partial   = partial + sb * sl.part;
        partial_b = slot_ptr(partial_b, sb * sl.part);
        sc += (int64_t) sb * SC_COUNT;
        is += (int64_t) sb * IS_COUNT;
        hist += sb * sl.hist;
    }
    const int measure = is[IS_MEASURE];
    if (stage != RED_ERR_INIT && stage != RED_SUM && stage != RED_ERR_SET && is[IS_DONE])
        return;
    double v[2] = {0.0, 0.0};
    bool   mx[2];
    for (int q = 0; q < nq; q += 2) {
        const int nqq = (nq - q) < 2 ? 1 : 2;
        for (int k = 0; k < 2; ++k) {
            const bool norm_slot = (k + q == 0) && (stage == RED_ERR_INIT || stage == RED_ERR_NEXT || stage == RED_ERR_SET);
            mx[k]                = norm_slot && measure == 2;
            double a[4] = {0.0, 0.0, 0.0, 0.0};
            if (k < nqq) {
                const bool    alt = partial_b && q + k == 1;
                const double *src = alt ? partial_b : partial + (int64_t) (q + k) * nblk;
                const int     nb  = alt ? nblk_b : nblk;
                int           b   = threadIdx.x;
                for (; b + 3 * kReduceThreads < nb; b += 4 * kReduceThreads)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const double p = src[b + j * kReduceThreads];
                        a[j]           = mx[k] ? fmax(a[j], p) : a[j] + p;
                    }
                for (; b < nb; b += kReduceThreads)
                    a[0] = mx[k] ? fmax(a[0], src[b]) : a[0] + src[b];
            }
            v[k] = mx[k] ? fmax(fmax(a[0], a[1]), fmax(a[2], a[3])) : (a[0] + a[1]) + (a[2] + a[3]);
        }
        block_reduce<2>(v, mx, sm);
        if (threadIdx.x == 0) {
            if (stage == RED_SUM) {
                sc[SC_RES0 + q] = v[0];
                if (nqq == 2)
                    sc[SC_RES0 + q + 1] = v[1];
            } else {
                const StageScalars S = load_stage_scalars(sc, is);
                if (stage == RED_ERR_INIT) {
                    is[IS_ITER]     = 0;
                    sc[SC_DELTA]    = 1.0;
                    sc[SC_DELTAMID] = 0.0;
                    record_error(sc, is, hist, S, 0, finish_norm(v[0], measure));
                } else if (stage == RED_DELTA) {
                    sc[SC_DELTA0] = S.delta;
                    sc[SC_DELTA]  = v[0];
                    sc[SC_BETA]   = fmax(0.0, (v[0] - S.deltamid) / S.delta);
                } else if (stage == RED_ALPHA) {
                    const double alpha                = S.delta / v[0];
                    sc[SC_DN]                         = v[0];
                    sc[SC_ALPHA]                      = alpha;
                    sc[SC_ARING + S.iter % kDirRing] = alpha;
                    // deltamid of the next iteration, <r - alpha K d, s> with
                    // s = d - beta d_prev (solverCG.h:87, 91 and 104):
                    // delta - alpha (<d, K d> - beta <K d_prev, d>), K symmetric
                    if (nq == 2)
                        sc[SC_DELTAMID] = S.delta - alpha * (v[0] - S.beta * v[1]);
                } else if (stage == RED_ERR_NEXT) {
                    is[IS_ITER] = S.iter + 1;
                    record_error(sc, is, hist, S, S.iter + 1, finish_norm(v[0], measure));
                } else if (stage == RED_ERR_SET) {
                    record_error(sc, is, hist, S, S.iter, finish_norm(v[0], measure));
                }
            }
        }
        __syncthreads();
    }
}

// Statistics of per-voxel isotropic moduli for the reference stiffness and the
// input check, one partial per block and quantity ([6][gridDim.x]):
// sum k, -min lambda, max lambda, -min g, max g (lambda = k - 2/3 g) and the
// count of voxels with k <= 0 or g <= 0 (g = k for conduction, H = 1)
__global__ __launch_bounds__(256) void k_moduli_stats(const double *__restrict__ k, const double *__restrict__ g,
                                                      int64_t n, double *__restrict__ partial)
{
    __shared__ double sm[6 * 4];
    double            v[6] = {0.0, -INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const double ki = k[i], gi = g[i], lam = ki - (2.0 / 3.0) * gi;
        v[0] += ki;
        v[1] = fmax(v[1], -lam);
        v[2] = fmax(v[2], lam);
        v[3] = fmax(v[3], -gi);
        v[4] = fmax(v[4], gi);
        v[5] += (ki > 0.0 && gi > 0.0) ? 0.0 : 1.0;
    }
    const bool mx[6] = {false, true, true, true, true, false};
    block_reduce<6>(v, mx, sm);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q)
            partial[(int64_t) q * gridDim.x + blockIdx.x] = v[q];
}

// partial <a, b> (slot 0) and optionally the norm of a (slot 1)
template <typename T>
__global__ __launch_bounds__(256) void k_dot(const T *__restrict__ a, const T *__restrict__ b, int64_t n, int measure,
                                             double *__restrict__ partial)
{
    __shared__ double sm[8];
    double            a0 = 0.0, a1 = 0.0;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        if (b)
            a0 += (double) a[i] * (double) b[i];
        acc_norm(a1, a[i], measure);
    }
    double     v[2]  = {a0, a1};
    const bool mx[2] = {false, measure == 2};
    block_reduce<2>(v, mx, sm);
    if (threadIdx.x == 0) {
        partial[blockIdx.x]             = v[0];
        partial[gridDim.x + blockIdx.x] = v[1];
    }
}

// start of an inner FP32 solve of the mixed form in one pass: r32 = (float) r,
// x = 0, d = 0 (solverCG's zero search direction; d null: kept) and per block
// the norm partial of r32 (slot 0); same grid and accumulation order as
// k_dot's norm
__global__ __launch_bounds__(256) void k_inner_start(float *__restrict__ r32, float *__restrict__ x,
                                                     float *__restrict__ d, const double *__restrict__ r, int64_t n,
                                                     int measure, double *__restrict__ partial, Slots sl)
{
    __shared__ double sm[8];
    const int         sb = slot_of(sl, false);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        r32 += sb * sl.field;
        x += sb * sl.field;
        d = slot_ptr(d, sb * sl.field);
        r += sb * sl.field;
        partial += sb * sl.part;
    }
    double            a0 = 0.0;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const float v = (float) r[i];
        r32[i]        = v;
        x[i]          = 0.f;
        if (d)
            d[i] = 0.f;
        acc_norm(a0, v, measure);
    }
    double     v[1]  = {a0};
    const bool mx[1] = {measure == 2};
    block_reduce<1>(v, mx, sm);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = v[0];
}

// Linear CG solution update over the direction ring: u -= alpha_j d_j for the
// ring slots 0 .. count-1 in order (iteration order, each step rounded as the
// reference's per-iteration update v_u_real -= alpha * d_real,
// solverCG.h:105), alpha_j from the device scalars
// elements per step of the ring update (256^3, interleaved runs: FP64 solve
// 70.06 -> 69.64 ms with 2, 69.92 with 4; mixed 38.48 -> 38.09 ms; the
// traced duration of the update itself stays at 0.70 ms)
#ifndef FANS_RING_GROUP
#define FANS_RING_GROUP 2
#endif
constexpr int kRingGroup = FANS_RING_GROUP;
template <typename T>
struct DirRing {
    const T *d[kDirRing];
};

// FULL: all kDirRing slots (count ignored; every load unconditional, so all of
// them are in flight together)
// count < 0: each slot's pending count, its iteration count mod kDirRing.
// gate: skip finished slots.
template <typename T, bool FULL>
__global__ __launch_bounds__(256) void k_ring_update(T *__restrict__ u, DirRing<T> ring,
                                                    const double *__restrict__ alphas, int count, int64_t n,
                                                    Slots sl, int gate)
{
    const int sb = slot_of(sl, gate != 0);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        u += sb * sl.field;
#pragma unroll
        for (int m = 0; m < kDirRing; ++m)
            ring.d[m] += sb * sl.field;
        alphas += (int64_t) sb * SC_COUNT;
    }
    if (count < 0)
        count = sl.is[sb * IS_COUNT + IS_ITER] % kDirRing;
    T a[kDirRing];
#pragma unroll
    for (int m = 0; m < kDirRing; ++m)
        a[m] = (FULL || m < count) ? (T) alphas[m] : (T) 0;
    // kRingGroup elements per step, all their loads before their stores (the
    // store of u may alias the next loads as far as the compiler knows); the
    // arithmetic of each element is unchanged
    const int64_t stride = (int64_t) gridDim.x * blockDim.x;
    int64_t       i      = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int G      = kRingGroup;
    for (; i + (G - 1) * stride < n; i += G * stride) {
        T dv[G][kDirRing], v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int m = 0; m < kDirRing; ++m)
                dv[g][m] = (FULL || m < count) ? __builtin_nontemporal_load(&ring.d[m][i + g * stride]) : (T) 0;
            v[g] = __builtin_nontemporal_load(&u[i + g * stride]);
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int m = 0; m < kDirRing; ++m)
                if (FULL || m < count)
                    v[g] -= a[m] * dv[g][m];
            __builtin_nontemporal_store(v[g], &u[i + g * stride]);
        }
    }
    for (; i < n; i += stride) {
        T dv[kDirRing];
#pragma unroll
        for (int m = 0; m < kDirRing; ++m)
            dv[m] = (FULL || m < count) ? __builtin_nontemporal_load(&ring.d[m][i]) : (T) 0;
        T v = __builtin_nontemporal_load(&u[i]);
#pragma unroll
        for (int m = 0; m < kDirRing; ++m)
            if (FULL || m < count)
                v -= a[m] * dv[m];
        __builtin_nontemporal_store(v, &u[i]);
    }
}

// Streaming copy, 16 bytes per lane per access, each block owning a
// contiguous chunk (the bandwidth reference of the passes above: the same
// access width and no arithmetic)
__global__ __launch_bounds__(256) void k_stream_copy(double2 *__restrict__ dst, const double2 *__restrict__ src,
                                                     int64_t n, int64_t per_block)
{
    const int64_t b0 = (int64_t) blockIdx.x * per_block, b1 = min(b0 + per_block, n);
    for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
        dst[i] = src[i];
}

// y = a * x + b * y  (a, b from host), n per slot
template <typename T, typename TX>
__global__ __launch_bounds__(256) void k_axpby(T *__restrict__ y, const TX *__restrict__ x, double a, double b, int64_t n,
                                               Slots sl)
{
    const int sb = slot_of(sl, false);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        y += sb * sl.field;
        x += sb * sl.field;
    }
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        y[i] = (T) (a * (double) x[i] + b * (double) y[i]);
}

// d = s + beta d with beta from the device scalars (nonlinear CG, solverCG.h:94)
template <typename T>
__global__ __launch_bounds__(256) void k_direction(T *__restrict__ d, const T *__restrict__ s, const double *__restrict__ sc,
                                                   int64_t n)
{
    const T beta = (T) sc[SC_BETA];
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        d[i] = s[i] + beta * d[i];
}

// start of the secant line search in one pass (solverCG.h:94, :127-129): the
// new direction d = s + beta d (beta from the device scalars), per block the
// partial of <r, d> (slot 0; same grid and accumulation order as k_dot) and the
// first step u += alpha0 d (the expression of k_axpby with b = 1)
template <typename T>
__global__ __launch_bounds__(256) void k_ls_start(T *__restrict__ d, const T *__restrict__ s,
                                                  const T *__restrict__ r, T *__restrict__ u,
                                                  const double *__restrict__ sc, double alpha0, int64_t n,
                                                  double *__restrict__ partial)
{
    __shared__ double sm[8];
    const T           beta = (T) sc[SC_BETA];
    double            a0   = 0.0;
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const T dv = s[i] + beta * d[i];
        d[i]       = dv;
        a0 += (double) r[i] * (double) dv;
        u[i] = (T) (alpha0 * (double) dv + 1.0 * (double) u[i]);
    }
    double     v[1]  = {a0};
    const bool mx[1] = {false};
    block_reduce<1>(v, mx, sm);
    if (threadIdx.x == 0)
        partial[blockIdx.x] = v[0];
}

template <typename T, typename TS>
__global__ __launch_bounds__(256) void k_convert(T *__restrict__ y, const TS *__restrict__ x, int64_t n)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        y[i] = (T) x[i];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ y, T v, int64_t n)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
        y[i] = v;
}

// Solver::extrapolateDisplacement (solver.h:302-311)
__global__ __launch_bounds__(256) void k_extrapolate(double *__restrict__ u, double *__restrict__ up, int64_t n)
{
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const double delta = u[i] - up[i];
        up[i]              = u[i];
        u[i] += delta;
    }
}

}  // namespace fans
