// LDS-staged Stockham FFT kernels for 3-D r2c / c2r of nodal fields.
//
// Layouts (all row-major, components fastest):
//   real field      [x][y][z][h]            (T)
//   spectrum        [x][y][kz][h], kz<=nz/2 (cplx<T>)
// The r2c/c2r along z use the half-length complex transform (pairs of reals
// packed into one complex number) and a post/pre-processing twiddle.  The y
// and x passes are batched strided transforms: one workgroup owns a tile of
// contiguous (kz,h) columns for every y (or x) and transforms it in LDS.  The x
// pass carries the Green operator multiply between the forward and inverse
// transforms, so the spectrum is read and written once for that stage.
#pragma once
#include "common.h"

namespace fans {

struct FftLen {
    int n;        // transform length
    int nst;      // number of radix stages
    int tws;      // stride into the twiddle table (table length n * tws)
    int rad[14];  // radices in application order
};

// ---- small DFT butterflies ------------------------------------------------
template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cplx<T> (&v)[R]);

template <>
__device__ __forceinline__ void dft<double, 2, false>(cplx<double> (&v)[2])
{
    cplx<double> a = v[0];
    v[0]           = a + v[1];
    v[1]           = a - v[1];
}
template <>
__device__ __forceinline__ void dft<double, 2, true>(cplx<double> (&v)[2])
{
    dft<double, 2, false>(v);
}
template <>
__device__ __forceinline__ void dft<float, 2, false>(cplx<float> (&v)[2])
{
    cplx<float> a = v[0];
    v[0]          = a + v[1];
    v[1]          = a - v[1];
}
template <>
__device__ __forceinline__ void dft<float, 2, true>(cplx<float> (&v)[2])
{
    dft<float, 2, false>(v);
}

template <typename T, bool INV>
__device__ __forceinline__ void dft3(cplx<T> (&v)[3])
{
    const T      s3 = (T) 0.86602540378443864676;
    cplx<T>      t1 = v[1] + v[2];
    cplx<T>      t2 = v[0] - cscale(t1, (T) 0.5);
    cplx<T>      t3 = mul_mi<T, INV>(cscale(v[1] - v[2], s3));
    v[0]            = v[0] + t1;
    v[1]            = t2 + t3;
    v[2]            = t2 - t3;
}
template <>
__device__ __forceinline__ void dft<double, 3, false>(cplx<double> (&v)[3]) { dft3<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 3, true>(cplx<double> (&v)[3]) { dft3<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 3, false>(cplx<float> (&v)[3]) { dft3<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 3, true>(cplx<float> (&v)[3]) { dft3<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx<T> (&v)[4])
{
    cplx<T> t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3];
    cplx<T> t3 = mul_mi<T, INV>(v[1] - v[3]);
    v[0]       = t0 + t2;
    v[2]       = t0 - t2;
    v[1]       = t1 + t3;
    v[3]       = t1 - t3;
}
template <>
__device__ __forceinline__ void dft<double, 4, false>(cplx<double> (&v)[4]) { dft4<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 4, true>(cplx<double> (&v)[4]) { dft4<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 4, false>(cplx<float> (&v)[4]) { dft4<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 4, true>(cplx<float> (&v)[4]) { dft4<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft5(cplx<T> (&v)[5])
{
    const T c1 = (T) 0.30901699437494742410, c2 = (T) -0.80901699437494742410;
    const T s1 = (T) 0.95105651629515357212, s2 = (T) 0.58778525229247312917;
    cplx<T> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
    cplx<T> p1 = v[0] + cscale(a1, c1) + cscale(a2, c2);
    cplx<T> p2 = v[0] + cscale(a1, c2) + cscale(a2, c1);
    cplx<T> q1 = mul_mi<T, INV>(cscale(b1, s1) + cscale(b2, s2));
    cplx<T> q2 = mul_mi<T, INV>(cscale(b1, s2) - cscale(b2, s1));
    v[0]       = v[0] + a1 + a2;
    v[1]       = p1 + q1;
    v[4]       = p1 - q1;
    v[2]       = p2 + q2;
    v[3]       = p2 - q2;
}
template <>
__device__ __forceinline__ void dft<double, 5, false>(cplx<double> (&v)[5]) { dft5<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 5, true>(cplx<double> (&v)[5]) { dft5<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 5, false>(cplx<float> (&v)[5]) { dft5<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 5, true>(cplx<float> (&v)[5]) { dft5<float, true>(v); }

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cplx<T> (&v)[8])
{
    const T r = (T) 0.70710678118654752440;
    cplx<T> a[4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        a[t] = v[t] + v[t + 4];
        b[t] = v[t] - v[t + 4];
    }
    // b[1] *= W8, b[2] *= W8^2, b[3] *= W8^3  (W8 = exp(-+ 2 pi i / 8))
    {
        cplx<T> x = b[1];
        b[1]      = INV ? cplx<T>{(x.x - x.y) * r, (x.x + x.y) * r} : cplx<T>{(x.x + x.y) * r, (x.y - x.x) * r};
        b[2]      = mul_mi<T, INV>(b[2]);
        x         = b[3];
        b[3]      = INV ? cplx<T>{(-x.x - x.y) * r, (x.x - x.y) * r} : cplx<T>{(x.y - x.x) * r, (-x.x - x.y) * r};
    }
    dft4<T, INV>(a);
    dft4<T, INV>(b);
    v[0] = a[0];
    v[2] = a[1];
    v[4] = a[2];
    v[6] = a[3];
    v[1] = b[0];
    v[3] = b[1];
    v[5] = b[2];
    v[7] = b[3];
}
template <>
__device__ __forceinline__ void dft<double, 8, false>(cplx<double> (&v)[8]) { dft8<double, false>(v); }
template <>
__device__ __forceinline__ void dft<double, 8, true>(cplx<double> (&v)[8]) { dft8<double, true>(v); }
template <>
__device__ __forceinline__ void dft<float, 8, false>(cplx<float> (&v)[8]) { dft8<float, false>(v); }
template <>
__device__ __forceinline__ void dft<float, 8, true>(cplx<float> (&v)[8]) { dft8<float, true>(v); }

// ---- one Stockham radix-R stage over `ncols` independent columns ----------
// element (i, col) lives at buf[i * ncols + col]
template <typename T, int R, bool INV>
__device__ __forceinline__ void stockham_stage(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst, int n,
                                               int ncols, int Ns, const cplx<T> *__restrict__ tw, int tws)
{
    const int nb    = n / R;
    const int tstep = (n / (Ns * R)) * tws;
    for (int w = threadIdx.x; w < nb * ncols; w += blockDim.x) {
        const int col = w % ncols, j = w / ncols;
        const int k   = j % Ns;
        cplx<T>   v[R];
#pragma unroll
        for (int t = 0; t < R; ++t)
            v[t] = src[(j + t * nb) * ncols + col];
        if (Ns > 1) {
#pragma unroll
            for (int t = 1; t < R; ++t) {
                const cplx<T> wv = tw[t * k * tstep];
                v[t]             = INV ? cmulc(v[t], wv) : cmul(v[t], wv);
            }
        }
        dft<T, R, INV>(v);
        const int o = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t)
            dst[(o + t * Ns) * ncols + col] = v[t];
    }
}

// full transform; returns the buffer that holds the result (a or b)
template <typename T, bool INV>
__device__ cplx<T> *lds_fft(cplx<T> *a, cplx<T> *b, const FftLen &L, const cplx<T> *__restrict__ tw, int ncols)
{
    int Ns = 1;
    for (int s = 0; s < L.nst; ++s) {
        const int R = L.rad[s];
        switch (R) {
        case 2: stockham_stage<T, 2, INV>(a, b, L.n, ncols, Ns, tw, L.tws); break;
        case 3: stockham_stage<T, 3, INV>(a, b, L.n, ncols, Ns, tw, L.tws); break;
        case 4: stockham_stage<T, 4, INV>(a, b, L.n, ncols, Ns, tw, L.tws); break;
        case 5: stockham_stage<T, 5, INV>(a, b, L.n, ncols, Ns, tw, L.tws); break;
        default: stockham_stage<T, 8, INV>(a, b, L.n, ncols, Ns, tw, L.tws); break;
        }
        __syncthreads();
        cplx<T> *t = a;
        a          = b;
        b          = t;
        Ns *= R;
    }
    return a;
}

// ---- z pass, forward r2c ----------------------------------------------------
// LM: complex length M = nz/2 with tws = 2 into the length-nz table `tw`.
// Block b handles lines [b*lpb, b*lpb + lpb) of nz*H reals each.
template <typename T, int H>
__global__ __launch_bounds__(256) void k_fft_z_r2c(const T *__restrict__ in, cplx<T> *__restrict__ out, FftLen LM,
                                                   const cplx<T> *__restrict__ tw, int nz, int64_t nlines, int lpb,
                                                   const int *__restrict__ done)
{
    if (done && *done)
        return;
    extern __shared__ char smem_raw[];
    cplx<T>      *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int     M     = nz / 2;
    const int64_t line0 = (int64_t) blockIdx.x * lpb;
    const int     nl    = (int) ((nlines - line0) < lpb ? (nlines - line0) : lpb);
    const int     ncols = nl * H;
    cplx<T>      *B     = A + M * ncols;
    const T      *src   = in + line0 * nz * H;
    const int     nzh   = nz * H;
    for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x) {
        const int l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
        T        *d = reinterpret_cast<T *>(&A[(z >> 1) * ncols + l * H + c]);
        d[z & 1]    = src[e];
    }
    __syncthreads();
    cplx<T>  *Z   = lds_fft<T, false>(A, B, LM, tw, ncols);
    const int Mc  = M + 1;
    cplx<T>  *dst = out + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * Mc * H; e += blockDim.x) {
        const int     l = e / (Mc * H), rem = e - l * (Mc * H), k = rem / H, c = rem - k * H;
        const int     col = l * H + c;
        const cplx<T> zk  = Z[(k == M ? 0 : k) * ncols + col];
        const cplx<T> zc  = conjg(Z[(k == 0 ? 0 : M - k) * ncols + col]);
        const cplx<T> E   = cscale(zk + zc, (T) 0.5);
        const cplx<T> D   = cscale(zk - zc, (T) 0.5);
        const cplx<T> O   = cplx<T>{D.y, -D.x};  // (zk - zc) / (2i)
        dst[e]            = E + cmul(tw[k], O);
    }
}

// ---- z pass, inverse c2r with fused epilogues -------------------------------
enum { EPI_STORE = 0, EPI_CG = 1, EPI_FP = 2 };

template <typename T>
struct EpiArgs {
    T            *out;      // EPI_STORE: result; EPI_CG: s = -result
    const T      *r;        // EPI_CG: partial <r, s>
    double       *u;        // EPI_FP: u -= result (FP64 displacement)
    double       *partial;  // EPI_CG: one partial sum per block
};

template <typename T, int H, int EPI>
__global__ __launch_bounds__(256) void k_fft_z_c2r(const cplx<T> *__restrict__ in, FftLen LM,
                                                   const cplx<T> *__restrict__ tw, int nz, int64_t nlines, int lpb,
                                                   EpiArgs<T> ep, const int *__restrict__ done)
{
    if (done && *done)
        return;
    extern __shared__ char smem_raw[];
    __shared__ double red[8];
    cplx<T>          *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int         M     = nz / 2, Mc = M + 1;
    const int64_t     line0 = (int64_t) blockIdx.x * lpb;
    const int         nl    = (int) ((nlines - line0) < lpb ? (nlines - line0) : lpb);
    const int         ncols = nl * H;
    cplx<T>          *B     = A + M * ncols;
    const cplx<T>    *src   = in + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * M * H; e += blockDim.x) {
        const int     l = e / (M * H), rem = e - l * (M * H), k = rem / H, c = rem - k * H;
        const cplx<T> xk = src[(l * Mc + k) * H + c];
        const cplx<T> xm = conjg(src[(l * Mc + (M - k)) * H + c]);
        const cplx<T> D  = xk - xm;
        const cplx<T> W  = conjg(tw[k]);
        const cplx<T> WD = cmul(W, D);
        A[k * ncols + l * H + c] = (xk + xm) + cplx<T>{-WD.y, WD.x};  // + i * W * D
    }
    __syncthreads();
    cplx<T>      *Z   = lds_fft<T, true>(A, B, LM, tw, ncols);
    const int     nzh = nz * H;
    const int64_t g0  = line0 * nzh;
    double        acc = 0.0;
    for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x) {
        const int     l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
        const cplx<T> v   = Z[(z >> 1) * ncols + l * H + c];
        const T       val = (z & 1) ? v.y : v.x;
        const int64_t gi  = g0 + e;
        if constexpr (EPI == EPI_STORE) {
            ep.out[gi] = val;
        } else if constexpr (EPI == EPI_CG) {
            const T sv = -val;
            ep.out[gi] = sv;
            acc += (double) ep.r[gi] * (double) sv;
        } else {
            ep.u[gi] -= (double) val;
        }
    }
    if constexpr (EPI == EPI_CG) {
        double     v[1]  = {acc};
        const bool mx[1] = {false};
        block_reduce<1>(v, mx, red);
        if (threadIdx.x == 0)
            ep.partial[blockIdx.x] = v[0];
    }
}

// ---- batched strided complex pass (y or x axis) ----------------------------
// data[o*ostride + i*sn + b] for i < L.n, b < Kc; block (bx, o) owns columns
// [bx*bc, bx*bc + bc).
template <typename T, bool INV>
__global__ __launch_bounds__(256) void k_fft_strided(cplx<T> *__restrict__ data, FftLen L,
                                                     const cplx<T> *__restrict__ tw, int64_t sn, int Kc,
                                                     int64_t ostride, int bc, const int *__restrict__ done)
{
    if (done && *done)
        return;
    extern __shared__ char smem_raw[];
    cplx<T>  *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int b0    = blockIdx.x * bc;
    const int ncols = (Kc - b0) < bc ? (Kc - b0) : bc;
    cplx<T>  *B     = A + L.n * ncols;
    cplx<T>  *base  = data + (int64_t) blockIdx.y * ostride + b0;
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i = w / ncols, c = w - i * ncols;
        A[w]        = base[i * sn + c];
    }
    __syncthreads();
    cplx<T> *R = lds_fft<T, INV>(A, B, L, tw, ncols);
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i     = w / ncols, c = w - i * ncols;
        base[i * sn + c] = R[w];
    }
}

// ---- x pass with the Green operator between forward and inverse ------------
// spectrum [x][y][kz][H]; block (bx, y) owns kz in [bx*bk, bx*bk+bk) and all H.
// G packed per frequency [x][y][kz][H(H+1)/2] = (00 01 02 11 12 22) or (0).
template <typename T, int H>
__global__ __launch_bounds__(256) void k_fft_x_green(cplx<T> *__restrict__ data, FftLen L,
                                                     const cplx<T> *__restrict__ tw, const T *__restrict__ G, int ny,
                                                     int Mc, int bk, const int *__restrict__ done)
{
    if (done && *done)
        return;
    constexpr int NP = H * (H + 1) / 2;
    extern __shared__ char smem_raw[];
    cplx<T>      *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int     y     = blockIdx.y;
    const int     kz0   = blockIdx.x * bk;
    const int     nk    = (Mc - kz0) < bk ? (Mc - kz0) : bk;
    const int     ncols = nk * H;
    cplx<T>      *B     = A + L.n * ncols;
    const int64_t sn    = (int64_t) ny * Mc * H;
    cplx<T>      *base  = data + ((int64_t) y * Mc + kz0) * H;
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i = w / ncols, c = w - i * ncols;
        A[w]        = base[i * sn + c];
    }
    __syncthreads();
    cplx<T> *R = lds_fft<T, false>(A, B, L, tw, ncols);
    cplx<T> *S = (R == A) ? B : A;
    for (int w = threadIdx.x; w < L.n * nk; w += blockDim.x) {
        const int i = w / nk, kk = w - i * nk;
        const T  *g = G + (((int64_t) i * ny + y) * Mc + kz0 + kk) * NP;
        cplx<T>  *v = R + i * ncols + kk * H;
        if constexpr (H == 1) {
            v[0] = cscale(v[0], g[0]);
        } else {
            const cplx<T> a = v[0], b = v[1], c = v[2];
            const T       g00 = g[0], g01 = g[1], g02 = g[2], g11 = g[3], g12 = g[4], g22 = g[5];
            v[0]                = cplx<T>{g00 * a.x + g01 * b.x + g02 * c.x, g00 * a.y + g01 * b.y + g02 * c.y};
            v[1]                = cplx<T>{g01 * a.x + g11 * b.x + g12 * c.x, g01 * a.y + g11 * b.y + g12 * c.y};
            v[2]                = cplx<T>{g02 * a.x + g12 * b.x + g22 * c.x, g02 * a.y + g12 * b.y + g22 * c.y};
        }
    }
    __syncthreads();
    cplx<T> *Q = lds_fft<T, true>(R, S, L, tw, ncols);
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i     = w / ncols, c = w - i * ncols;
        base[i * sn + c] = Q[w];
    }
}

}  // namespace fans
