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
//
// Every stage runs in place in a single LDS tile: each thread first reads all
// inputs of its butterflies into registers, the block synchronises, then the
// outputs are written.  Rows of the tile are padded by one complex value so
// the power-of-two strides of the first stages do not collide in LDS banks.
#pragma once
#include "common.h"
#include "fft_butterflies.h"

namespace fans {

struct FftLen {
    int n;        // transform length
    int nst;      // number of radix stages
    int tws;      // stride into the twiddle table (table length n * tws)
    int rad[14];  // radices in application order
};

// Register budget of the in-place stages: a thread holds at most FFT_MAXV
// complex values; callers keep n * ncols <= FFT_MAXV * 240 so every radix
// (including 3 and 5, where MAXV/R rounds down) fits.
constexpr int FFT_MAXV = 16;
constexpr int FFT_MAX_ELEMS = FFT_MAXV * 240;

// ---- one in-place Stockham radix-R stage over `ncols` independent columns --
// element (i, col) lives at buf[i * ld + col]
template <typename T, int R, bool INV>
__device__ __forceinline__ void stockham_stage(cplx<T> *__restrict__ buf, int n, int ncols, int ld, int Ns,
                                               const cplx<T> *__restrict__ tw, int tws)
{
    constexpr int MB    = FFT_MAXV / R;
    const int     nb    = n / R;
    const int     total = nb * ncols;
    const int     tstep = (n / (Ns * R)) * tws;
    cplx<T>       v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int w = threadIdx.x + b * blockDim.x;
        if (w < total) {
            const int col = w % ncols, j = w / ncols;
#pragma unroll
            for (int t = 0; t < R; ++t)
                v[b][t] = buf[(j + t * nb) * ld + col];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int w = threadIdx.x + b * blockDim.x;
        if (w < total) {
            const int col = w % ncols, j = w / ncols, k = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    const cplx<T> wv = tw[t * k * tstep];
                    v[b][t]          = INV ? cmulc(v[b][t], wv) : cmul(v[b][t], wv);
                }
            }
            dft<T, R, INV>(v[b]);
            const int o = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t)
                buf[(o + t * Ns) * ld + col] = v[b][t];
        }
    }
    __syncthreads();
}

// full in-place transform of the LDS tile (ends synchronised)
template <typename T, bool INV>
__device__ void lds_fft(cplx<T> *buf, const FftLen &L, const cplx<T> *__restrict__ tw, int ncols, int ld)
{
    int Ns = 1;
    for (int s = 0; s < L.nst; ++s) {
        const int R = L.rad[s];
        switch (R) {
        case 2: stockham_stage<T, 2, INV>(buf, L.n, ncols, ld, Ns, tw, L.tws); break;
        case 3: stockham_stage<T, 3, INV>(buf, L.n, ncols, ld, Ns, tw, L.tws); break;
        case 4: stockham_stage<T, 4, INV>(buf, L.n, ncols, ld, Ns, tw, L.tws); break;
        case 5: stockham_stage<T, 5, INV>(buf, L.n, ncols, ld, Ns, tw, L.tws); break;
        default: stockham_stage<T, 8, INV>(buf, L.n, ncols, ld, Ns, tw, L.tws); break;
        }
        Ns *= R;
    }
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
    const int     ncols = nl * H, ld = ncols + 1;
    const T      *src   = in + line0 * nz * H;
    const int     nzh   = nz * H;
    for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x) {
        const int l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
        T        *d = reinterpret_cast<T *>(&A[(z >> 1) * ld + l * H + c]);
        d[z & 1]    = src[e];
    }
    __syncthreads();
    lds_fft<T, false>(A, LM, tw, ncols, ld);
    const int Mc  = M + 1;
    cplx<T>  *dst = out + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * Mc * H; e += blockDim.x) {
        const int     l = e / (Mc * H), rem = e - l * (Mc * H), k = rem / H, c = rem - k * H;
        const int     col = l * H + c;
        const cplx<T> zk  = A[(k == M ? 0 : k) * ld + col];
        const cplx<T> zc  = conjg(A[(k == 0 ? 0 : M - k) * ld + col]);
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
    const int         ncols = nl * H, ld = ncols + 1;
    const cplx<T>    *src   = in + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * M * H; e += blockDim.x) {
        const int     l = e / (M * H), rem = e - l * (M * H), k = rem / H, c = rem - k * H;
        const cplx<T> xk = src[(l * Mc + k) * H + c];
        const cplx<T> xm = conjg(src[(l * Mc + (M - k)) * H + c]);
        const cplx<T> D  = xk - xm;
        const cplx<T> W  = conjg(tw[k]);
        const cplx<T> WD = cmul(W, D);
        A[k * ld + l * H + c] = (xk + xm) + cplx<T>{-WD.y, WD.x};  // + i * W * D
    }
    __syncthreads();
    lds_fft<T, true>(A, LM, tw, ncols, ld);
    const int     nzh = nz * H;
    const int64_t g0  = line0 * nzh;
    double        acc = 0.0;
    for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x) {
        const int     l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
        const cplx<T> v   = A[(z >> 1) * ld + l * H + c];
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
    const int ncols = (Kc - b0) < bc ? (Kc - b0) : bc, ld = ncols + 1;
    cplx<T>  *base  = data + (int64_t) blockIdx.y * ostride + b0;
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i = w / ncols, c = w - i * ncols;
        A[i * ld + c] = base[i * sn + c];
    }
    __syncthreads();
    lds_fft<T, INV>(A, L, tw, ncols, ld);
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i     = w / ncols, c = w - i * ncols;
        base[i * sn + c] = A[i * ld + c];
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
    const int     ncols = nk * H, ld = ncols + 1;
    const int64_t sn    = (int64_t) ny * Mc * H;
    cplx<T>      *base  = data + ((int64_t) y * Mc + kz0) * H;
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i = w / ncols, c = w - i * ncols;
        A[i * ld + c] = base[i * sn + c];
    }
    __syncthreads();
    lds_fft<T, false>(A, L, tw, ncols, ld);
    for (int w = threadIdx.x; w < L.n * nk; w += blockDim.x) {
        const int i = w / nk, kk = w - i * nk;
        const T  *g = G + (((int64_t) i * ny + y) * Mc + kz0 + kk) * NP;
        cplx<T>  *v = A + i * ld + kk * H;
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
    lds_fft<T, true>(A, L, tw, ncols, ld);
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i     = w / ncols, c = w - i * ncols;
        base[i * sn + c] = A[i * ld + c];
    }
}

}  // namespace fans
