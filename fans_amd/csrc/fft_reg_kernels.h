// Register four-step FFT kernels for the power-of-two transform lengths
// 16..256 and, along y, for 48 = 12 x 4, 80 = 20 x 4 and 96 = 24 x 4.
//
// A length N = N1 * N2 transform of one column is done by N2 cooperating
// threads, thread j holding the N1 values x[j + N2 m]:
//   1. an N1-point DFT in registers over m,
//   2. the twiddle W_N^(j k1),
//   3. one transpose through LDS (rows swizzled so neither the writes nor the
//      reads of a wave collide in LDS banks),
//   4. N1/N2 N2-point DFTs in registers.
// Thread j ends with X[j + N2 q + N1 k2] in v[q N2 + k2] -- the same index set
// an inverse transform of the same shape takes as input, so a forward
// transform, a pointwise operation and an inverse transform chain without a
// second exchange (used by the x pass around the Green operator).
#pragma once
#include <type_traits>
#include "fft_butterflies.h"
#include "fft_kernels.h"

namespace fans {

// in-register DFT of length 16 as 4 x 4 with constant internal twiddles
template <typename T, bool INV>
__device__ __forceinline__ void dft16(cplx<T> (&v)[16])
{
    const T c = (T) 0.92387953251128675613, s = (T) 0.38268343236508977173, r = (T) 0.70710678118654752440;
    cplx<T> y[4][4];
#pragma unroll
    for (int i2 = 0; i2 < 4; ++i2) {
        cplx<T> a[4] = {v[i2], v[i2 + 4], v[i2 + 8], v[i2 + 12]};
        dft4<T, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
            y[i2][k1] = a[k1];
    }
    // y[i2][k1] *= W16^(i2 k1); forward W16^p = (cos, -sin)(2 pi p / 16)
    auto tw = [&](cplx<T> a, T wr, T wi) { return INV ? cplx<T>{a.x * wr + a.y * wi, a.y * wr - a.x * wi}
                                                       : cplx<T>{a.x * wr - a.y * wi, a.x * wi + a.y * wr}; };
    // (wr, wi) below are the forward twiddles; the lambda conjugates for INV
    y[1][1] = tw(y[1][1], c, -s);
    y[1][2] = tw(y[1][2], r, -r);
    y[1][3] = tw(y[1][3], s, -c);
    y[2][1] = tw(y[2][1], r, -r);
    y[2][2] = tw(y[2][2], (T) 0, (T) -1);
    y[2][3] = tw(y[2][3], -r, -r);
    y[3][1] = tw(y[3][1], s, -c);
    y[3][2] = tw(y[3][2], -r, -r);
    y[3][3] = tw(y[3][3], -c, s);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cplx<T> a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<T, INV>(a);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2)
            v[k1 + 4 * k2] = a[k2];
    }
}

// exp(-2 pi i k / n) for compile-time k, n: the angle reduced to a quadrant
// and a Taylor series there (terms below 1e-20 of the result)
constexpr double ct_taylor(double x, bool sine)
{
    // sin: x - x^3/3! + ...; cos: 1 - x^2/2! + ...
    double t = sine ? x : 1.0, s = t;
    for (int m = 1; m < 14; ++m) {
        const double a = sine ? 2.0 * m : 2.0 * m - 1.0;
        t *= -x * x / (a * (a + 1.0));
        s += t;
    }
    return s;
}
struct CtW {
    double re, im;
};
constexpr CtW ct_twiddle(int k, int n)
{
    k %= n;
    const int    q  = (4 * k) / n;  // quadrant of 2 pi k / n
    const double r  = 1.57079632679489661923 * (double) (4 * k - q * n) / (double) n;
    const double cr = ct_taylor(r, false), sr = ct_taylor(r, true);
    // angle q pi/2 + r: rotate (cos r, sin r) by q quarter turns
    const double c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
    const double s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
    return CtW{c, -s};
}

template <typename T, int N, bool INV>
__device__ __forceinline__ void dft_reg(cplx<T> (&v)[N]);

// in-register DFT of length P Q by Cooley-Tukey: P-point DFTs over the
// stride-Q subsequences, the twiddles W_PQ^(i2 k1), Q-point DFTs; input
// index i2 + Q m, output k1 + P k2 (dft16's order for P = Q = 4)
template <typename T, int P, int Q, bool INV>
__device__ __forceinline__ void dft_pq(cplx<T> (&v)[P * Q])
{
    cplx<T> y[Q][P];
#pragma unroll
    for (int i2 = 0; i2 < Q; ++i2) {
        cplx<T> a[P];
#pragma unroll
        for (int m = 0; m < P; ++m)
            a[m] = v[i2 + Q * m];
        dft_reg<T, P, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) {
            if (i2 * k1 == 0) {
                y[i2][k1] = a[k1];
            } else {
                constexpr int n = P * Q;
                const CtW     w = ct_twiddle(i2 * k1, n);
                const T       wr = (T) w.re, wi = INV ? (T) -w.im : (T) w.im;
                y[i2][k1]        = cplx<T>{a[k1].x * wr - a[k1].y * wi, a[k1].x * wi + a[k1].y * wr};
            }
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        cplx<T> b[Q];
#pragma unroll
        for (int i2 = 0; i2 < Q; ++i2)
            b[i2] = y[i2][k1];
        dft_reg<T, Q, INV>(b);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2)
            v[k1 + P * k2] = b[k2];
    }
}

template <typename T, int N, bool INV>
__device__ __forceinline__ void dft_reg(cplx<T> (&v)[N])
{
    if constexpr (N == 2 || N == 3 || N == 5)
        dft<T, N, INV>(v);
    else if constexpr (N == 4)
        dft4<T, INV>(v);
    else if constexpr (N == 8)
        dft8<T, INV>(v);
    else if constexpr (N == 16)
        dft16<T, INV>(v);
    else if constexpr (N == 12)
        dft_pq<T, 3, 4, INV>(v);
    else if constexpr (N == 20)
        dft_pq<T, 5, 4, INV>(v);
    else {
        static_assert(N == 24, "unsupported register DFT length");
        dft_pq<T, 3, 8, INV>(v);
    }
}

// Four-step transform of one column (see the file comment).  `tile` is the
// block's LDS tile with row stride `ld` (complex), `col` this column; `tw` the
// length-N table W_N^e = exp(-2 pi i e / N).  Ends synchronised.
// The pointer to the twiddle table, made opaque to the optimiser: a forward
// and an inverse transform in one kernel read the same twiddles, and without
// this the compiler keeps the forward pass's N1 - 1 complex twiddles live in
// registers across everything in between (60 VGPRs in FP64), halving the
// occupancy of the x+Green pass.
template <typename T>
__device__ __forceinline__ const cplx<T> *opaque_tw(const cplx<T> *tw)
{
    asm volatile("" : "+s"(tw));
    return tw;
}

template <typename T, int N1, int N2, bool INV>
__device__ __forceinline__ void fourstep(cplx<T> (&v)[N1], int j, cplx<T> *tile, int ld, int col,
                                         const cplx<T> *__restrict__ tw_in)
{
    const cplx<T> *__restrict__ tw = opaque_tw(tw_in);
    dft_reg<T, N1, INV>(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        const cplx<T> w = tw[j * k1];
        v[k1]           = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1)
        tile[(k1 * N2 + ((j + k1) & (N2 - 1))) * ld + col] = v[k1];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        const int k1 = j + N2 * q;
        cplx<T>   w[N2];
#pragma unroll
        for (int i2 = 0; i2 < N2; ++i2)
            w[i2] = tile[(k1 * N2 + ((i2 + k1) & (N2 - 1))) * ld + col];
        dft_reg<T, N2, INV>(w);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            v[q * N2 + k2] = w[k2];
    }
    __syncthreads();
}

// fourstep() with the transpose done in two halves (real parts, then imaginary
// parts) through a real-valued tile of half the size: the same data movement
// with twice the barriers, for kernels whose occupancy the full complex tile
// limits.  `tile` holds N1 * N2 rows of `ld` reals.
template <typename T, int N1, int N2, bool INV>
__device__ __forceinline__ void fourstep_split(cplx<T> (&v)[N1], int j, T *tile, int ld, int col,
                                               const cplx<T> *__restrict__ tw_in)
{
    const cplx<T> *__restrict__ tw = opaque_tw(tw_in);
    dft_reg<T, N1, INV>(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
        const cplx<T> w = tw[j * k1];
        v[k1]           = INV ? cmulc(v[k1], w) : cmul(v[k1], w);
    }
    cplx<T> w[N1 / N2][N2];
#pragma unroll
    for (int part = 0; part < 2; ++part) {
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1)
            tile[(k1 * N2 + ((j + k1) & (N2 - 1))) * ld + col] = part ? v[k1].y : v[k1].x;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < N1 / N2; ++q) {
            const int k1 = j + N2 * q;
#pragma unroll
            for (int i2 = 0; i2 < N2; ++i2) {
                const T x = tile[(k1 * N2 + ((i2 + k1) & (N2 - 1))) * ld + col];
                if (part)
                    w[q][i2].y = x;
                else
                    w[q][i2].x = x;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q) {
        dft_reg<T, N2, INV>(w[q]);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            v[q * N2 + k2] = w[q][k2];
    }
}

// register order of the four-step output -> input order of the next transform
// of the same shape: v_in[m] = v_out[q N2 + k2] with m = q + (N1/N2) k2
template <typename T, int N1, int N2>
__device__ __forceinline__ void fourstep_out_to_in(cplx<T> (&v)[N1])
{
    cplx<T> t[N1];
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2)
            t[q + (N1 / N2) * k2] = v[q * N2 + k2];
#pragma unroll
    for (int m = 0; m < N1; ++m)
        v[m] = t[m];
}

// ---- plain strided pass (y axis) ------------------------------------------
// data[o*ostride + i*sn + b], i < N1*N2, b in [c0, c1); C columns per block,
// block = C * N2 threads, column fastest in the lane order.  1-D grid of
// nbx * n_outer blocks, XCD-mapped so that neighbouring column groups (which
// share cache lines) run on the same XCD.
template <typename T, int N1, int N2, int C, bool INV>
__global__ __launch_bounds__(C * N2) void k_fft_strided_reg(cplx<T> *__restrict__ data,
                                                            const cplx<T> *__restrict__ tw, int64_t sn, int c0,
                                                            int c1, int64_t ostride, int nbx, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        data += sb * sl.spec;
    }
    constexpr int N  = N1 * N2;
    constexpr int LD = C + 1;
    __shared__ cplx<T> tile[N * LD];
    const int          t = threadIdx.x, col = t % C, j = t / C;
    const int          lb = xcd_tile_index(blockIdx.x, gridDim.x), bx = lb % nbx, bo = lb / nbx;
    const int          b  = c0 + bx * C + col;
    const bool         ok = b < c1;
    cplx<T>           *base = data + (int64_t) bo * ostride + b;
    cplx<T>            v[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m)
        v[m] = ok ? base[(int64_t) (j + N2 * m) * sn] : cplx<T>{(T) 0, (T) 0};
    fourstep<T, N1, N2, INV>(v, j, tile, LD, col, tw);
    if (ok)
#pragma unroll
        for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2)
                base[(int64_t) (j + N2 * q + N1 * k2) * sn] = v[q * N2 + k2];
}

// lane <- lane (quad base + s) for every lane of a quad (DPP quad_perm broadcast)
template <int S>
__device__ __forceinline__ double quad_bcast(double x)
{
    constexpr int ctrl = S * 0x55;
    const int     lo   = __builtin_amdgcn_mov_dpp(__double2loint(x), ctrl, 0xf, 0xf, true);
    const int     hi   = __builtin_amdgcn_mov_dpp(__double2hiint(x), ctrl, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
template <int S>
__device__ __forceinline__ float quad_bcast(float x)
{
    constexpr int ctrl = S * 0x55;
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), ctrl, 0xf, 0xf, true));
}

// lane l <- lane l + S inside each row of 16 lanes (S > 0: row_shl, S < 0:
// row_shr); lanes whose source leaves the row get 0 (bound_ctrl)
template <int S>
__device__ __forceinline__ int row_shift_b32(int x)
{
    constexpr int ctrl = S > 0 ? 0x100 + S : 0x110 - S;
    return __builtin_amdgcn_mov_dpp(x, ctrl, 0xf, 0xf, true);
}
template <int S>
__device__ __forceinline__ double row_shift(double x)
{
    return __hiloint2double(row_shift_b32<S>(__double2hiint(x)), row_shift_b32<S>(__double2loint(x)));
}
template <int S>
__device__ __forceinline__ float row_shift(float x)
{
    return __int_as_float(row_shift_b32<S>(__float_as_int(x)));
}

#ifndef FANS_XG_LDS_GREEN
#define FANS_XG_LDS_GREEN 0
#endif
constexpr bool kXGreenLdsGreen = FANS_XG_LDS_GREEN;
#ifndef FANS_XG_KEEP_INPUT
#define FANS_XG_KEEP_INPUT 1
#endif
constexpr bool kXGreenKeepInput = FANS_XG_KEEP_INPUT;

// ---- x pass: forward FFT, Green multiply, inverse FFT ---------------------
// spectrum [x][y][kz][H]; the C columns of a block are (kz, c).  For H = 3
// they are packed in threes when C = 3 KB + 1 (C = 16: five frequencies per row
// of 16 lanes, one idle lane; the Green multiply gathers the other two
// components with row shifts by 1 and 2), else padded to 4 (the components of
// a frequency in a lane quad, exchanged with quad DPP; a quarter of the lanes
// idle).  Packing in threes took the 256^3 pass from 0.313 to 0.293 ms (FP64)
// and 0.186 to 0.163 ms (FP32).
// Logical block (bk, y) of the XCD-mapped 1-D grid (nbk * ny blocks) owns kz
// in [kz_lo + bk*KB, ... + KB).  partial (nullable): the block's partial of the
// spectral <r, s>, s = -G r (kz_weight), at partial[blockIdx.x].
// SPLIT: the transposes go through a real-valued tile of half the size
// (fourstep_split), which lifts the LDS limit on resident workgroups.
template <typename T, int H, int N1, int N2, int KB, int C, bool SPLIT>
__global__ __launch_bounds__(C *N2) void k_fft_x_green_reg(cplx<T> *__restrict__ data, const cplx<T> *__restrict__ tw,
                                                          GreenView<T> gv, int kz_lo, int kz_hi, int nbk,
                                                          double *__restrict__ partial, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        data += sb * sl.spec;
        gv.G += sb * sl.green;
        partial = slot_ptr(partial, sb * sl.part);
    }
    __shared__ double red[(C * N2) / 64];
    double            acc = 0.0;
    constexpr int  N    = N1 * N2;
    constexpr bool GRP3 = H == 3 && C == 3 * KB + 1;
    constexpr int  HP   = H == 3 ? (GRP3 ? 3 : 4) : 1;
    constexpr int  LD   = C + 1;
    static_assert(GRP3 || C == KB * HP, "column layout");
    __shared__ cplx<T> tile_c[SPLIT ? 1 : N * LD];
    __shared__ T       tile_r[SPLIT ? N * LD : 1];
    const int          ny = gv.ny, Mc = gv.Mc;
    const int          t = threadIdx.x, col = t % C, j = t / C;
    const int          kk = col / HP, c = col - HP * kk;
    const int          lb = xcd_tile_index(blockIdx.x, gridDim.x);
    const int          kz = kz_lo + (lb % nbk) * KB + kk;
    // ky and its mirror ny - ky read the same quarter-stored Green values:
    // order the rows 0, 1, ny-1, 2, ny-2, ... so that both run back to back
    // on one XCD and the second finds them in that XCD's L2
    const int          yt = lb / nbk;
    const int          y  = yt == 0 ? 0 : ((yt & 1) ? (yt + 1) >> 1 : ny - (yt >> 1));
    const bool         ok = kz < kz_hi && c < H && kk < KB;
    const int64_t      sn = (int64_t) ny * Mc * H;
    cplx<T>           *base = data + ((int64_t) y * Mc + kz) * H + c;
    cplx<T>            v[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m)
        v[m] = ok ? base[(int64_t) (j + N2 * m) * sn] : cplx<T>{(T) 0, (T) 0};
    cplx<T> rkeep[kXGreenKeepInput ? N1 : 1];
    if constexpr (kXGreenKeepInput)
#pragma unroll
        for (int m = 0; m < N1; ++m)
            rkeep[m] = v[m];
    if constexpr (SPLIT)
        fourstep_split<T, N1, N2, false>(v, j, tile_r, LD, col, tw);
    else
        fourstep<T, N1, N2, false>(v, j, tile_c, LD, col, tw);
    // Parseval weight of this lane's frequency; 0 on idle lanes (which hold
    // zeros anyway) and past kz_hi
    const double wk = ok ? kz_weight(kz, Mc) : 0.0;
    // The Green values of this block -- stored rows kx' < gxq of its ky and
    // its KB frequencies kz, P = H(H+1)/2 each -- are staged into the transpose
    // tile, idle between the two transforms: one coalesced cooperative load
    // instead of 3 dependent loads per frequency and lane.
    constexpr int P   = H * (H + 1) / 2;
    const int     kz0 = kz_lo + (lb % nbk) * KB, nk = min(KB, kz_hi - kz0);
    const bool    fy  = y >= gv.gyq;
    const T       sy  = fy ? (T) -1 : (T) 1;
    const int     kyq = fy ? ny - y : y;
    T *const      Gs  = SPLIT ? tile_r : reinterpret_cast<T *>(tile_c);
    static_assert((SPLIT ? 1 : 2) * N * LD >= N * KB * P, "the Green block fits the tile");
    if constexpr (kXGreenLdsGreen) {
        const int      per = nk * P, tot = gv.gxq * per;
        const T *const src = gv.G + ((int64_t) kyq * Mc + kz0) * P;
        const int64_t  rs  = (int64_t) gv.gyq * Mc * P;  // stride between stored kx rows
        for (int e = threadIdx.x; e < tot; e += C * N2) {
            const int r = e / per, m = e - r * per;
            Gs[r * (KB * P) + m] = src[r * rs + m];
        }
    }
    if constexpr (kXGreenLdsGreen)
        __syncthreads();
    const int kkg = kk < nk ? kk : nk - 1;  // idle lanes read a valid slot
#pragma unroll
    for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) {
            const int  row = j + N2 * q + N1 * k2;  // x frequency
            const bool fx  = row >= gv.gxq;
            const T    sx  = fx ? (T) -1 : (T) 1;
            const T   *g   = kXGreenLdsGreen
                                 ? Gs + (fx ? gv.nx - row : row) * (KB * P) + kkg * P
                                 : gv.G + (((int64_t) (fx ? gv.nx - row : row) * gv.gyq + kyq) * Mc + kz0 + kkg) * P;
            cplx<T>   &a   = v[q * N2 + k2];
            if constexpr (H == 1) {
                a = cscale(a, g[0]);
            } else {
                // row c of the packed symmetric 3x3 (00 01 02 11 12 22) with the
                // mirror signs of the xy (sx sy), xz (sx) and yz (sy) entries
                const int i0 = c == 0 ? 0 : (c == 1 ? 1 : 2);
                const int i1 = c == 0 ? 1 : (c == 1 ? 3 : 4);
                const int i2 = c == 0 ? 2 : (c == 1 ? 4 : 5);
                const T   s0 = c == 0 ? (T) 1 : (c == 1 ? sx * sy : sx);
                const T   s1 = c == 0 ? sx * sy : (c == 1 ? (T) 1 : sy);
                const T   s2 = c == 0 ? sx : (c == 1 ? sy : (T) 1);
                const T   ga = g[c < 3 ? i0 : 0] * s0, gb = g[c < 3 ? i1 : 0] * s1, gc = g[c < 3 ? i2 : 0] * s2;
                cplx<T>   v0, v1, v2;
                if constexpr (GRP3) {
                    // components 0 1 2 of this frequency at lanes l - c .. l - c + 2
                    const cplx<T> m1{row_shift<-1>(a.x), row_shift<-1>(a.y)};
                    const cplx<T> m2{row_shift<-2>(a.x), row_shift<-2>(a.y)};
                    const cplx<T> p1{row_shift<1>(a.x), row_shift<1>(a.y)};
                    const cplx<T> p2{row_shift<2>(a.x), row_shift<2>(a.y)};
                    v0 = c == 0 ? a : (c == 1 ? m1 : m2);
                    v1 = c == 0 ? p1 : (c == 1 ? a : m1);
                    v2 = c == 0 ? p2 : (c == 1 ? p1 : a);
                } else {
                    v0 = cplx<T>{quad_bcast<0>(a.x), quad_bcast<0>(a.y)};
                    v1 = cplx<T>{quad_bcast<1>(a.x), quad_bcast<1>(a.y)};
                    v2 = cplx<T>{quad_bcast<2>(a.x), quad_bcast<2>(a.y)};
                }
                a = cplx<T>{ga * v0.x + gb * v1.x + gc * v2.x, ga * v0.y + gb * v1.y + gc * v2.y};
            }
        }
    if constexpr (kXGreenLdsGreen)
        __syncthreads();  // the tile is the inverse transform's again
    fourstep_out_to_in<T, N1, N2>(v);
    if constexpr (SPLIT)
        fourstep_split<T, N1, N2, true>(v, j, tile_r, LD, col, tw);
    else
        fourstep<T, N1, N2, true>(v, j, tile_c, LD, col, tw);
    if (ok) {
        // <r, s> by Parseval along x: with G r^ = S^ and the unnormalised
        // inverse S = sum_kx S^ e^(+i..), sum_kx conj(r^) S^ = sum_x conj(r) S
        // for the x line's input r (this pass's own input, re-read just before
        // the output overwrites it -- a cache hit, the block read it at entry)
        // and s = -S
        if (partial)
#pragma unroll
            for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
                for (int k2 = 0; k2 < N2; ++k2) {
                    // input element x = j + N2 q + N1 k2 is v[m] with m = q + (N1/N2) k2 at entry
                    const cplx<T> rin = kXGreenKeepInput ? rkeep[q + (N1 / N2) * k2]
                                                         : base[(int64_t) (j + N2 * q + N1 * k2) * sn];
                    const cplx<T> &o  = v[q * N2 + k2];
                    acc -= wk * ((double) rin.x * o.x + (double) rin.y * o.y);
                }
#pragma unroll
        for (int q = 0; q < N1 / N2; ++q)
#pragma unroll
            for (int k2 = 0; k2 < N2; ++k2)
                base[(int64_t) (j + N2 * q + N1 * k2) * sn] = v[q * N2 + k2];
    }
    if (partial) {
        double     v1[1] = {acc};
        const bool mx[1] = {false};
        block_reduce<1>(v1, mx, red);
        if (threadIdx.x == 0)
            partial[blockIdx.x] = v1[0];
    }
}

}  // namespace fans
