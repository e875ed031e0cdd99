// Shared device/host helpers for the FANS MI355X kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace fans {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T a, T b)
{
    return cplx<T>{a, b};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b)
{
    return {a.x + b.x, a.y + b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b)
{
    return {a.x - b.x, a.y - b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s)
{
    return {a.x * s, a.y * s};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conjg(cplx<T> a)
{
    return {a.x, -a.y};
}
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV>
__host__ __device__ __forceinline__ cplx<T> mul_mi(cplx<T> a)
{
    return INV ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

// 64-lane wavefront reductions (deterministic butterfly order)
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide reduction of NV doubles; result valid in thread 0.  `sm` must hold
// NV * (blockDim.x / 64) doubles.  is_max selects max instead of sum per slot.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const bool (&is_max)[NV], double *sm)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k)
        v[k] = is_max[k] ? wave_max(v[k]) : wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
            sm[k * nw + wid] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double a = sm[k * nw];
            for (int w = 1; w < nw; ++w)
                a = is_max[k] ? fmax(a, sm[k * nw + w]) : a + sm[k * nw + w];
            v[k] = a;
        }
    }
}

// phase id of voxel i: the device microstructure holds bytes when every id of
// the grid is below 256 and 16-bit ids (the reference's unsigned short,
// reader.h:53) otherwise
__device__ __forceinline__ int phase_id(const uint8_t *ms, int wide, int64_t i)
{
    return wide ? (int) reinterpret_cast<const uint16_t *>(ms)[i] : (int) ms[i];
}

// accumulate |v| into a norm: measure 0 = L1, 1 = L2 (squared), 2 = Linf
template <typename T>
__device__ __forceinline__ void acc_norm(double &acc, T v, int measure)
{
    const double a = fabs((double) v);
    if (measure == 2)
        acc = fmax(acc, a);
    else if (measure == 1)
        acc += a * a;
    else
        acc += a;
}

// device integer scalars of one RVE slot (the double scalars are laid out in
// vec_kernels.h, kScSlot per slot)
enum { IS_DONE = 0, IS_ITER = 1, IS_NIT = 2, IS_ERRTYPE = 3, IS_MEASURE = 4, IS_COUNT = 8 };
constexpr int kScSlot = 40;

// RVE slots of a launch.  A context holds one RVE or a batch of same-shape
// RVEs whose per-RVE buffers lie at a fixed stride from slot 0's.  A batched
// launch enumerates its slots in grid z through `map` (slot = map[blockIdx.z],
// entries < 0: none) and the kernels move every buffer to the slot's offset;
// without a map the launch has one z block and its buffers are the RVE's own
// (slot 0, no offsets).  With `is`, a slot whose done flag is set returns at
// entry.
struct Slots {
    const int *map   = nullptr;
    const int *is    = nullptr;  // IS_COUNT ints per slot
    int64_t    field = 0;        // elements of a field array (u, r, d, K d ...)
    int64_t    spec  = 0;        // complex entries of a spectrum
    int64_t    green = 0;        // entries of a Green operator
    int64_t    part  = 0;        // doubles of a partial-sum buffer
    int64_t    ms    = 0;        // bytes of a phase-id array
    int64_t    ph    = 0;        // entries of a phase table
    int64_t    hist  = 0;        // doubles of an error history
};

// slot of this block, or -1 if it has none or (check_done) is finished
__device__ __forceinline__ int slot_of(const Slots &s, bool check_done = true)
{
    const int b = s.map ? s.map[blockIdx.z] : (int) blockIdx.z;
    if (b < 0 || (check_done && s.is && s.is[b * IS_COUNT + IS_DONE]))
        return -1;
    return b;
}

// p + off unless p is null (optional buffers stay null)
template <typename P>
__device__ __forceinline__ P *slot_ptr(P *p, int64_t off)
{
    return p ? p + off : p;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map the hardware block id so that each XCD gets a contiguous range
// of logical blocks: neighbouring blocks that share cache lines (halo nodes,
// partial lines of narrow column groups) then meet in the same L2.
__device__ __forceinline__ int xcd_tile_index(int bid, int nb)
{
    const int q = nb / 8, r = nb % 8, k = bid % 8, p = bid / 8;
    return k * q + min(k, r) + p;
}

}  // namespace fans
