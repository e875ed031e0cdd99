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

// Green operator storage, G[(kxq * gyq + kyq) * Mc + kz][H(H+1)/2] packed
// (00 01 02 11 12 22), 1/N folded in.  Full storage: gxq = nx, gyq = ny.
// Quarter storage (the reference element stiffness is symmetric under the x
// and y mirrors and n_x, n_y are even): only kx <= nx/2 and ky <= ny/2 are
// stored, and G(-kx) = Rx G(kx) Rx with Rx = diag(-1, 1, 1) flips the xy and
// xz entries (G(-ky): the xy and yz entries).
template <typename T>
struct GreenView {
    const T *G;
    int      nx, ny, gxq, gyq, Mc;
};

// stored block of frequency (kx, ky, kz) and the mirror signs sx, sy
template <typename T, int H>
__device__ __forceinline__ const T *green_at(const GreenView<T> &gv, int kx, int ky, int kz, T &sx, T &sy)
{
    const bool fx = kx >= gv.gxq, fy = ky >= gv.gyq;
    sx            = fx ? (T) -1 : (T) 1;
    sy            = fy ? (T) -1 : (T) 1;
    const int kxq = fx ? gv.nx - kx : kx, kyq = fy ? gv.ny - ky : ky;
    return gv.G + (((int64_t) kxq * gv.gyq + kyq) * gv.Mc + kz) * (H * (H + 1) / 2);
}

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

// ---- CG update fused into the forward z pass --------------------------------
// PRO_CG: the forward z pass of CG iteration i first finishes the residual
// update of iteration i-1 (solverCG.h:106-117): r -= alpha rnew, and emits per
// block the partial |r| (norm of the measure), which the reduction turns into
// the error.  Only the updated r enters the transform.  deltamid = <r, s_prev>
// needs no s here: the CG scalars form it from <d, K d> (matvec) and
// <d, K d_prev> (inverse z pass).  The matching solution update u -= alpha d
// runs in the inverse z pass of the next iteration, which reads that d anyway.
enum { PRO_NONE = 0, PRO_CG = 1 };

template <typename T>
struct ProArgs {
    const T      *rnew;     // K d of the previous iteration
    const double *alpha;    // device scalar
    int           measure;  // 0 L1, 1 L2, 2 Linf
    double       *partial;  // [gridDim.x]
};

// Loads a thread issues together before the stores that follow them in the
// fused prologues / epilogues of the z passes.  Element by element, a store may
// alias the next element's loads, so the compiler keeps program order: one
// memory round trip per element.  kZGroup for the compile-time passes,
// kZGroupRt for the runtime-length ones (bounds-checked groups).
#ifndef FANS_Z_GROUP
#define FANS_Z_GROUP 16
#endif
constexpr int kZGroup = FANS_Z_GROUP;
#ifndef FANS_Z_GROUP_RT
#define FANS_Z_GROUP_RT 8
#endif
constexpr int kZGroupRt = FANS_Z_GROUP_RT;

template <typename T, int PRO>
__device__ __forceinline__ T z_load(T *__restrict__ in, int64_t gi, const ProArgs<T> &pa, T alpha, double &a0)
{
    if constexpr (PRO == PRO_CG) {
        // r and rnew are not read again before the spectrum passes that
        // follow: streaming (non-temporal) accesses keep them from displacing
        // the freshly written spectrum in the Infinity Cache
        const T rv = __builtin_nontemporal_load(&in[gi]) - alpha * __builtin_nontemporal_load(&pa.rnew[gi]);
        __builtin_nontemporal_store(rv, &in[gi]);
        acc_norm(a0, rv, pa.measure);
        return rv;
    } else {
        return in[gi];
    }
}

template <int PRO, typename T>
__device__ __forceinline__ void z_pro_partials(const ProArgs<T> &pa, double a0, double *red)
{
    if constexpr (PRO == PRO_CG) {
        double     v[1]  = {a0};
        const bool mx[1] = {pa.measure == 2};
        block_reduce<1>(v, mx, red);
        if (threadIdx.x == 0)
            pa.partial[blockIdx.x] = v[0];
    }
}

// r2c post-processing of the half-length transform Z (in LDS, column col):
// X[k] = (Z[k] + conj Z[M-k]) / 2 + W^k (Z[k] - conj Z[M-k]) / (2i)
template <typename T>
__device__ __forceinline__ cplx<T> r2c_untangle(const cplx<T> *A, int ld, int col, int k, int M,
                                                const cplx<T> *__restrict__ tw)
{
    const cplx<T> zk = A[(k == M ? 0 : k) * ld + col];
    const cplx<T> zc = conjg(A[(k == 0 ? 0 : M - k) * ld + col]);
    const cplx<T> E  = cscale(zk + zc, (T) 0.5);
    const cplx<T> D  = cscale(zk - zc, (T) 0.5);
    return E + cmul(tw[k], cplx<T>{D.y, -D.x});
}

// ---- z pass, forward r2c ----------------------------------------------------
// LM: complex length M = nz/2 with tws = 2 into the length-nz table `tw`.
// Block b handles lines [b*lpb, b*lpb + lpb) of nz*H reals each.
template <typename T, int H, int PRO>
__global__ __launch_bounds__(256) void k_fft_z_r2c(T *__restrict__ in, cplx<T> *__restrict__ out, FftLen LM,
                                                   const cplx<T> *__restrict__ tw, int nz, int64_t nlines, int lpb,
                                                   ProArgs<T> pa, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        in += sb * sl.field;
        out += sb * sl.spec;
        slot_pro(pa, sb, sl);
    }
    extern __shared__ char smem_raw[];
    __shared__ double red[8];
    cplx<T>      *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int     M     = nz / 2;
    const int64_t line0 = (int64_t) blockIdx.x * lpb;
    const int     nl    = (int) ((nlines - line0) < lpb ? (nlines - line0) : lpb);
    const int     ncols = nl * H, ld = ncols + 1;
    const int     nzh   = nz * H;
    const int64_t g0    = line0 * nzh;
    T             alpha = (T) 0;
    double        a0    = 0.0;
    if constexpr (PRO == PRO_CG)
        alpha = (T) *pa.alpha;
    if constexpr (PRO == PRO_CG && kZGroupRt > 0) {
        // r -= alpha K d in groups (same arithmetic and order as z_load)
        constexpr int CH = kZGroupRt;
        const int     tot = nl * nzh, stp = blockDim.x;
        for (int e0 = threadIdx.x; e0 < tot; e0 += CH * stp) {
            T rv[CH], qv[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int e = e0 + i * stp;
                rv[i] = qv[i] = (T) 0;
                if (e < tot) {
                    rv[i] = __builtin_nontemporal_load(&in[g0 + e]);
                    qv[i] = __builtin_nontemporal_load(&pa.rnew[g0 + e]);
                }
            }
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int e = e0 + i * stp;
                if (e < tot) {
                    const int l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
                    const T   r = rv[i] - alpha * qv[i];
                    __builtin_nontemporal_store(r, &in[g0 + e]);
                    acc_norm(a0, r, pa.measure);
                    T *d     = reinterpret_cast<T *>(&A[(z >> 1) * ld + l * H + c]);
                    d[z & 1] = r;
                }
            }
        }
    } else {
        for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x) {
            const int l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
            T        *d = reinterpret_cast<T *>(&A[(z >> 1) * ld + l * H + c]);
            d[z & 1]    = z_load<T, PRO>(in, g0 + e, pa, alpha, a0);
        }
    }
    __syncthreads();
    lds_fft<T, false>(A, LM, tw, ncols, ld);
    const int Mc  = M + 1;
    cplx<T>  *dst = out + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * Mc * H; e += blockDim.x) {
        const int l = e / (Mc * H), rem = e - l * (Mc * H), k = rem / H, c = rem - k * H;
        dst[e]      = r2c_untangle(A, ld, l * H + c, k, M, tw);
    }
    z_pro_partials<PRO>(pa, a0, red);
}

// ---- z pass, inverse c2r with fused epilogues -------------------------------
// EPI_CGD (linear CG, solverCG.h:88-92): the new search direction
// d = s + beta d_prev with s = -result (beta from the spectral <r, s> of the x
// pass); s itself is not stored.  With qprev = K d_prev (the previous matvec's
// output) it emits per block the partial <d, K d_prev>, from which the CG
// scalars form the next deltamid.
enum { EPI_STORE = 0, EPI_CG = 1, EPI_FP = 2, EPI_CGD = 3 };

template <typename T>
struct EpiArgs {
    T            *out;      // EPI_STORE: result; EPI_CG: s = -result
    const T      *r;        // EPI_CG: partial <r, s>
    double       *u;        // EPI_FP: u -= result (FP64 displacement)
    double       *partial;  // EPI_CG / EPI_CGD: one partial sum per block
    const T      *dprev;    // EPI_CGD: previous direction (read)
    const T      *qprev;    // EPI_CGD: K d_prev (nullable), partial <d, K d_prev>
    T            *dnext;    // EPI_CGD: new direction (written)
    const double *scal;     // EPI_CGD: device scalars (beta at [0])
};

// the argument blocks at RVE slot sb
template <typename T>
__device__ __forceinline__ void slot_pro(ProArgs<T> &pa, int sb, const Slots &sl)
{
    pa.rnew    = slot_ptr(pa.rnew, sb * sl.field);
    pa.alpha   = slot_ptr(pa.alpha, (int64_t) sb * kScSlot);
    pa.partial = slot_ptr(pa.partial, sb * sl.part);
}
template <typename T>
__device__ __forceinline__ void slot_epi(EpiArgs<T> &ep, int sb, const Slots &sl)
{
    ep.out     = slot_ptr(ep.out, sb * sl.field);
    ep.r       = slot_ptr(ep.r, sb * sl.field);
    ep.u       = slot_ptr(ep.u, sb * sl.field);
    ep.partial = slot_ptr(ep.partial, sb * sl.part);
    ep.dprev   = slot_ptr(ep.dprev, sb * sl.field);
    ep.qprev   = slot_ptr(ep.qprev, sb * sl.field);
    ep.dnext   = slot_ptr(ep.dnext, sb * sl.field);
    ep.scal    = slot_ptr(ep.scal, (int64_t) sb * kScSlot);
}

// c2r pre-processing of one (line, component): `line` points at kz = 0 of the
// component (stride H); the half-length transform input is
// Z[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k])
template <typename T, int H>
__device__ __forceinline__ cplx<T> c2r_tangle(const cplx<T> *__restrict__ line, int k, int M,
                                              const cplx<T> *__restrict__ tw)
{
    const cplx<T> xk = line[k * H];
    const cplx<T> xm = conjg(line[(M - k) * H]);
    const cplx<T> WD = cmul(conjg(tw[k]), xk - xm);
    return (xk + xm) + cplx<T>{-WD.y, WD.x};
}

template <typename T, int EPI>
__device__ __forceinline__ void z_epilogue(const EpiArgs<T> &ep, int64_t gi, T val, double &acc, T beta)
{
    if constexpr (EPI == EPI_STORE) {
        ep.out[gi] = val;
    } else if constexpr (EPI == EPI_CGD) {
        // d_next is read by the very next pass (the matvec) and stays cached;
        // d_prev and u stream past the caches
        const T dp = __builtin_nontemporal_load(&ep.dprev[gi]);
        const T dn = beta * dp - val;
        ep.dnext[gi] = dn;
        if (ep.qprev)
            acc += (double) dn * (double) __builtin_nontemporal_load(&ep.qprev[gi]);
    } else if constexpr (EPI == EPI_CG) {
        // s is read by the very next pass (the matvec) and stays cached; r is
        // not read again this iteration
        const T sv = -val;
        ep.out[gi] = sv;
        acc += (double) __builtin_nontemporal_load(&ep.r[gi]) * (double) sv;
    } else {
        ep.u[gi] -= (double) val;
    }
}

// z_epilogue over this thread's elements e = threadIdx.x + k stp < tot in
// groups of CH, each group's global operands requested before its stores (the
// stores never alias the operands of another element: d_next and d_prev, s
// and r are distinct buffers and u is updated per element); same arithmetic
// and accumulation order as z_epilogue.  val_of(e) is element e's result.
template <typename T, int EPI, int CH, typename ValOf>
__device__ __forceinline__ void z_epilogue_grouped(const EpiArgs<T> &ep, int64_t g0, int tot, int stp, ValOf val_of,
                                                   double &acc, T beta)
{
    for (int e0 = threadIdx.x; e0 < tot; e0 += CH * stp) {
        T      pa[CH], pb[CH];
        double pu[CH];
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int     e  = e0 + i * stp;
            const int64_t gi = g0 + e;
            pa[i] = pb[i] = (T) 0;
            pu[i]         = 0.0;
            if (e < tot) {
                if constexpr (EPI == EPI_CGD) {
                    pa[i] = __builtin_nontemporal_load(&ep.dprev[gi]);
                    if (ep.qprev)
                        pb[i] = __builtin_nontemporal_load(&ep.qprev[gi]);
                } else if constexpr (EPI == EPI_CG) {
                    pb[i] = __builtin_nontemporal_load(&ep.r[gi]);
                } else if constexpr (EPI == EPI_FP) {
                    pu[i] = ep.u[gi];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < CH; ++i) {
            const int     e  = e0 + i * stp;
            const int64_t gi = g0 + e;
            if (e < tot) {
                const T val = val_of(e);
                if constexpr (EPI == EPI_STORE) {
                    ep.out[gi] = val;
                } else if constexpr (EPI == EPI_CGD) {
                    const T dn   = beta * pa[i] - val;
                    ep.dnext[gi] = dn;
                    if (ep.qprev)
                        acc += (double) dn * (double) pb[i];
                } else if constexpr (EPI == EPI_CG) {
                    const T sv = -val;
                    ep.out[gi] = sv;
                    acc += (double) pb[i] * (double) sv;
                } else {
                    ep.u[gi] = pu[i] - (double) val;
                }
            }
        }
    }
}

template <int EPI, typename T>
__device__ __forceinline__ void z_epi_partials(const EpiArgs<T> &ep, double acc, double *red)
{
    if (EPI == EPI_CG || (EPI == EPI_CGD && ep.qprev)) {
        double     v[1]  = {acc};
        const bool mx[1] = {false};
        block_reduce<1>(v, mx, red);
        if (threadIdx.x == 0)
            ep.partial[blockIdx.x] = v[0];
    }
}

template <typename T, int H, int EPI>
__global__ __launch_bounds__(256) void k_fft_z_c2r(const cplx<T> *__restrict__ in, FftLen LM,
                                                   const cplx<T> *__restrict__ tw, int nz, int64_t nlines, int lpb,
                                                   EpiArgs<T> ep, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        in += sb * sl.spec;
        slot_epi(ep, sb, sl);
    }
    extern __shared__ char smem_raw[];
    __shared__ double red[8];
    cplx<T>          *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int         M     = nz / 2, Mc = M + 1;
    const int64_t     line0 = (int64_t) blockIdx.x * lpb;
    const int         nl    = (int) ((nlines - line0) < lpb ? (nlines - line0) : lpb);
    const int         ncols = nl * H, ld = ncols + 1;
    const cplx<T>    *src   = in + line0 * Mc * H;
    for (int e = threadIdx.x; e < nl * M * H; e += blockDim.x) {
        const int l = e / (M * H), rem = e - l * (M * H), k = rem / H, c = rem - k * H;
        A[k * ld + l * H + c] = c2r_tangle<T, H>(src + l * Mc * H + c, k, M, tw);
    }
    __syncthreads();
    lds_fft<T, true>(A, LM, tw, ncols, ld);
    const int     nzh = nz * H;
    const int64_t g0  = line0 * nzh;
    double        acc = 0.0;
    T             beta = (T) 0;
    if constexpr (EPI == EPI_CGD)
        beta = (T) ep.scal[0];
    auto val_of = [&](int e) {
        const int     l = e / nzh, rem = e - l * nzh, z = rem / H, c = rem - z * H;
        const cplx<T> v = A[(z >> 1) * ld + l * H + c];
        return (z & 1) ? v.y : v.x;
    };
    if constexpr (kZGroupRt > 0 && EPI != EPI_STORE) {
        z_epilogue_grouped<T, EPI, kZGroupRt>(ep, g0, nl * nzh, blockDim.x, val_of, acc, beta);
    } else {
        for (int e = threadIdx.x; e < nl * nzh; e += blockDim.x)
            z_epilogue<T, EPI>(ep, g0 + e, val_of(e), acc, beta);
    }
    z_epi_partials<EPI>(ep, acc, red);
}

// ---- z passes with compile-time shape ---------------------------------------
// operands of the linear CG's inverse z epilogue requested before the
// transform (0 none, 1 d_prev, 2 d_prev and K d_prev), per precision; the
// others are requested together right after it.  At 256^3: FP64 0.294 ms
// with d_prev ahead, 0.295 with none, 0.306 with both; FP32 0.134 ms with
// none, 0.136 with d_prev, 0.223 with both.
#ifndef FANS_C2R_AHEAD64
#define FANS_C2R_AHEAD64 1
#endif
#ifndef FANS_C2R_AHEAD32
#define FANS_C2R_AHEAD32 0
#endif
constexpr int kC2rAhead64 = FANS_C2R_AHEAD64;
constexpr int kC2rAhead32 = FANS_C2R_AHEAD32;

// For M = nz/2 in {16, 24, 32, 40, 48, 64, 128} the line count per block, the
// thread count and every index are compile-time constants (no runtime divisions), and the
// thread count equals the butterfly count of the first radix stage.
template <int M>
constexpr int zct_radix1()
{
    return M >= 32 ? 8 : 4;
}
template <int H, int M>
constexpr int zct_lpb()
{
    int l = 1;
    while ((M / zct_radix1<M>()) * (2 * l) * H <= 256)
        l *= 2;
    return l;
}
template <int H, int M>
constexpr int zct_threads()
{
    return (M / zct_radix1<M>()) * zct_lpb<H, M>() * H;
}

template <typename T, int R, bool INV, int N, int NC, int LD, int NS, int NT>
__device__ __forceinline__ void stockham_stage_ct(cplx<T> *__restrict__ buf, const cplx<T> *__restrict__ tw)
{
    constexpr int nb = N / R, total = nb * NC, MB = (total + NT - 1) / NT;
    constexpr int tstep = 2 * (N / (NS * R));  // into the length-2N table
    cplx<T>       v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int w = threadIdx.x + b * NT;
        if (total % NT == 0 || w < total) {
            const int col = w % NC, j = w / NC;
#pragma unroll
            for (int t = 0; t < R; ++t)
                v[b][t] = buf[(j + t * nb) * LD + col];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const int w = threadIdx.x + b * NT;
        if (total % NT == 0 || w < total) {
            const int col = w % NC, j = w / NC, k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    const cplx<T> wv = tw[t * k * tstep];
                    v[b][t]          = INV ? cmulc(v[b][t], wv) : cmul(v[b][t], wv);
                }
            }
            dft<T, R, INV>(v[b]);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t)
                buf[(o + t * NS) * LD + col] = v[b][t];
        }
    }
    __syncthreads();
}

// radix schedules: 128 = 8*4*4, 64 = 8*8, 48 = 8*3*2, 40 = 8*5, 32 = 8*4,
// 24 = 4*3*2, 16 = 4*4 (the first radix is zct_radix1)
template <typename T, bool INV, int N, int NC, int LD, int NT>
__device__ __forceinline__ void lds_fft_ct(cplx<T> *buf, const cplx<T> *__restrict__ tw)
{
    if constexpr (N == 48) {
        stockham_stage_ct<T, 8, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 3, INV, N, NC, LD, 8, NT>(buf, tw);
        stockham_stage_ct<T, 2, INV, N, NC, LD, 24, NT>(buf, tw);
    } else if constexpr (N == 40) {
        stockham_stage_ct<T, 8, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 5, INV, N, NC, LD, 8, NT>(buf, tw);
    } else if constexpr (N == 24) {
        stockham_stage_ct<T, 4, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 3, INV, N, NC, LD, 4, NT>(buf, tw);
        stockham_stage_ct<T, 2, INV, N, NC, LD, 12, NT>(buf, tw);
    } else if constexpr (N == 128) {
        stockham_stage_ct<T, 8, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 4, INV, N, NC, LD, 8, NT>(buf, tw);
        stockham_stage_ct<T, 4, INV, N, NC, LD, 32, NT>(buf, tw);
    } else if constexpr (N == 64) {
        stockham_stage_ct<T, 8, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 8, INV, N, NC, LD, 8, NT>(buf, tw);
    } else if constexpr (N == 32) {
        stockham_stage_ct<T, 8, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 4, INV, N, NC, LD, 8, NT>(buf, tw);
    } else {
        static_assert(N == 16, "unsupported compile-time length");
        stockham_stage_ct<T, 4, INV, N, NC, LD, 1, NT>(buf, tw);
        stockham_stage_ct<T, 4, INV, N, NC, LD, 4, NT>(buf, tw);
    }
}

// forward: lines [b*LPB, b*LPB + LPB); the host launches nlines / LPB blocks
template <typename T, int H, int M, int PRO>
__global__ __launch_bounds__((zct_threads<H, M>())) void k_fft_z_r2c_ct(T *__restrict__ in, cplx<T> *__restrict__ out,
                                                                      const cplx<T> *__restrict__ tw, ProArgs<T> pa,
                                                                      Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        in += sb * sl.field;
        out += sb * sl.spec;
        slot_pro(pa, sb, sl);
    }
    constexpr int NT = zct_threads<H, M>(), LPB = zct_lpb<H, M>();
    constexpr int NZH = 2 * M * H, NC = LPB * H, LD = NC + 1, MC = M + 1;
    __shared__ cplx<T> A[M * LD];
    __shared__ double  red[NT / 64 + 1];
    const int64_t      line0 = (int64_t) blockIdx.x * LPB;
    const int64_t      g0    = line0 * NZH;
    T                  alpha = (T) 0;
    double             a0    = 0.0;
    if constexpr (PRO == PRO_CG)
        alpha = (T) *pa.alpha;
    constexpr int PER = (LPB * NZH) / NT;
    constexpr int CH  = kZGroup < PER ? kZGroup : PER;
    if constexpr (PRO == PRO_CG && kZGroup > 0 && (LPB * NZH) % NT == 0 && PER % CH == 0) {
        // r -= alpha K d in groups of CH elements: the loads of a group are in
        // flight together (same arithmetic and accumulation order as z_load)
        for (int i0 = 0; i0 < PER; i0 += CH) {
            T rv[CH], qv[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int64_t gi = g0 + threadIdx.x + (i0 + i) * NT;
                rv[i]            = __builtin_nontemporal_load(&in[gi]);
                qv[i]            = __builtin_nontemporal_load(&pa.rnew[gi]);
            }
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                const int e = threadIdx.x + (i0 + i) * NT;
                const int l = e / NZH, rem = e - l * NZH, z = rem / H, c = rem - z * H;
                const T   r = rv[i] - alpha * qv[i];
                __builtin_nontemporal_store(r, &in[g0 + e]);
                acc_norm(a0, r, pa.measure);
                T *d     = reinterpret_cast<T *>(&A[(z >> 1) * LD + l * H + c]);
                d[z & 1] = r;
            }
        }
    } else {
#pragma unroll 4
        for (int e = threadIdx.x; e < LPB * NZH; e += NT) {
            const int l = e / NZH, rem = e - l * NZH, z = rem / H, c = rem - z * H;
            T        *d = reinterpret_cast<T *>(&A[(z >> 1) * LD + l * H + c]);
            d[z & 1]    = z_load<T, PRO>(in, g0 + e, pa, alpha, a0);
        }
    }
    __syncthreads();
    lds_fft_ct<T, false, M, NC, LD, NT>(A, tw);
    cplx<T> *dst = out + line0 * MC * H;
#pragma unroll 4
    for (int e = threadIdx.x; e < LPB * MC * H; e += NT) {
        const int l = e / (MC * H), rem = e - l * (MC * H), k = rem / H, c = rem - k * H;
        dst[e]      = r2c_untangle(A, LD, l * H + c, k, M, tw);
    }
    z_pro_partials<PRO>(pa, a0, red);
}

template <typename T, int H, int M, int EPI>
__global__ __launch_bounds__((zct_threads<H, M>())) void k_fft_z_c2r_ct(const cplx<T> *__restrict__ in,
                                                                      const cplx<T> *__restrict__ tw, EpiArgs<T> ep,
                                                                      Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        in += sb * sl.spec;
        slot_epi(ep, sb, sl);
    }
    constexpr int NT = zct_threads<H, M>(), LPB = zct_lpb<H, M>();
    constexpr int NZH = 2 * M * H, NC = LPB * H, LD = NC + 1, MC = M + 1;
    __shared__ cplx<T> A[M * LD];
    __shared__ double  red[NT / 64 + 1];
    const int64_t      line0 = (int64_t) blockIdx.x * LPB;
    const cplx<T>     *src   = in + line0 * MC * H;
    const int64_t      g0    = line0 * NZH;
    // EPI_CGD: this thread's values of d_prev and K d_prev, the ones of
    // kC2rAhead* requested before the transform so that their loads overlap it,
    // at the cost of registers during it
    constexpr int  PER   = (LPB * NZH) / NT;
    constexpr int  AHEAD = sizeof(T) == 8 ? kC2rAhead64 : kC2rAhead32;
    constexpr bool PRE   = EPI == EPI_CGD && (LPB * NZH) % NT == 0;
    constexpr bool PRED = PRE && AHEAD >= 1, PREQ = PRE && AHEAD >= 2;
    T              dpv[PRE ? PER : 1], qpv[PRE ? PER : 1];
    if constexpr (PRED) {
        const T *__restrict__ dp = ep.dprev;
        const T *__restrict__ qp = ep.qprev;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            dpv[i] = __builtin_nontemporal_load(&dp[g0 + threadIdx.x + i * NT]);
            if constexpr (PREQ)
                qpv[i] = qp ? __builtin_nontemporal_load(&qp[g0 + threadIdx.x + i * NT]) : (T) 0;
        }
    }
#pragma unroll 4
    for (int e = threadIdx.x; e < LPB * M * H; e += NT) {
        const int l = e / (M * H), rem = e - l * (M * H), k = rem / H, c = rem - k * H;
        A[k * LD + l * H + c] = c2r_tangle<T, H>(src + l * MC * H + c, k, M, tw);
    }
    __syncthreads();
    lds_fft_ct<T, true, M, NC, LD, NT>(A, tw);
    double acc  = 0.0;
    T      beta = (T) 0;
    if constexpr (EPI == EPI_CGD)
        beta = (T) ep.scal[0];
    if constexpr (PRE) {
        T *__restrict__ dn = ep.dnext;
        // the operands not requested ahead: all of this thread's values are
        // requested before the first store of d
        if constexpr (!PRED)
#pragma unroll
            for (int i = 0; i < PER; ++i)
                dpv[i] = __builtin_nontemporal_load(&ep.dprev[g0 + threadIdx.x + i * NT]);
        if constexpr (!PREQ)
            if (ep.qprev)
#pragma unroll
                for (int i = 0; i < PER; ++i)
                    qpv[i] = __builtin_nontemporal_load(&ep.qprev[g0 + threadIdx.x + i * NT]);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int     e = threadIdx.x + i * NT;
            const int     l = e / NZH, rem = e - l * NZH, z = rem / H, c = rem - z * H;
            const cplx<T> v = A[(z >> 1) * LD + l * H + c];
            const T       d = beta * dpv[i] - ((z & 1) ? v.y : v.x);
            dn[g0 + e]      = d;
            if (ep.qprev)
                acc += (double) d * (double) qpv[i];
        }
    } else {
        auto val_of = [&](int e) {
            const int     l = e / NZH, rem = e - l * NZH, z = rem / H, c = rem - z * H;
            const cplx<T> v = A[(z >> 1) * LD + l * H + c];
            return (z & 1) ? v.y : v.x;
        };
        if constexpr (kZGroup > 0 && EPI != EPI_STORE) {
            constexpr int CH = kZGroup < 8 ? kZGroup : 8;
            z_epilogue_grouped<T, EPI, CH>(ep, g0, LPB * NZH, NT, val_of, acc, beta);
        } else {
#pragma unroll 4
            for (int e = threadIdx.x; e < LPB * NZH; e += NT)
                z_epilogue<T, EPI>(ep, g0 + e, val_of(e), acc, beta);
        }
    }
    z_epi_partials<EPI>(ep, acc, red);
}

// ---- batched strided complex pass (y or x axis) ----------------------------
// data[o*ostride + i*sn + b] for i < L.n, b in [c0, c1); block (bx, o) owns
// columns [c0 + bx*bc, c0 + bx*bc + bc).
template <typename T, bool INV>
__global__ __launch_bounds__(256) void k_fft_strided(cplx<T> *__restrict__ data, FftLen L,
                                                     const cplx<T> *__restrict__ tw, int64_t sn, int c0, int c1,
                                                     int64_t ostride, int bc, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        data += sb * sl.spec;
    }
    extern __shared__ char smem_raw[];
    cplx<T>  *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int b0    = c0 + blockIdx.x * bc;
    const int ncols = (c1 - b0) < bc ? (c1 - b0) : bc, ld = ncols + 1;
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

// ---- spectral <r, s> --------------------------------------------------------
// With s = -G * r, the unnormalised r2c / c2r pair and G holding 1/N,
// sum_x r(x) s(x) = sum over the full spectrum of conj(r^) s^ (Parseval); the
// half spectrum kz in [0, nz/2] counts the frequencies 0 < kz < nz/2 twice
// (their mirrors -kz are the complex conjugates) and kz = 0, nz/2 once.
__device__ __forceinline__ double kz_weight(int kz, int Mc) { return (kz == 0 || kz == Mc - 1) ? 1.0 : 2.0; }

// ---- x pass with the Green operator between forward and inverse ------------
// spectrum [x][y][kz][H]; block (bx, y) owns kz in [kz_lo + bx*bk, ... + bk) and all H.
// G packed per frequency [x][y][kz][H(H+1)/2] = (00 01 02 11 12 22) or (0).
// partial (nullable): one partial of <r, s> (s = -G r) per block, block index
// blockIdx.y * gridDim.x + blockIdx.x.
template <typename T, int H>
__global__ __launch_bounds__(256) void k_fft_x_green(cplx<T> *__restrict__ data, FftLen L,
                                                     const cplx<T> *__restrict__ tw, GreenView<T> gv, int kz_lo,
                                                     int kz_hi, int bk, double *__restrict__ partial, Slots sl)
{
    const int sb = slot_of(sl);
    if (sb < 0)
        return;
    if (sl.map) {  // batched launch: this slot's buffers
        data += sb * sl.spec;
        gv.G += sb * sl.green;
        partial = slot_ptr(partial, sb * sl.part);
    }
    const int ny = gv.ny, Mc = gv.Mc;
    __shared__ double red[8];
    double            acc = 0.0;
    extern __shared__ char smem_raw[];
    cplx<T>      *A     = reinterpret_cast<cplx<T> *>(smem_raw);
    const int     y     = blockIdx.y;
    const int     kz0   = kz_lo + blockIdx.x * bk;
    const int     nk    = (kz_hi - kz0) < bk ? (kz_hi - kz0) : bk;
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
        T         sx, sy;
        const T  *g = green_at<T, H>(gv, i, y, kz0 + kk, sx, sy);
        cplx<T>  *v = A + i * ld + kk * H;
        const double wk = kz_weight(kz0 + kk, Mc);
        if constexpr (H == 1) {
            const cplx<T> a = v[0];
            v[0]            = cscale(a, g[0]);
            acc -= wk * ((double) a.x * v[0].x + (double) a.y * v[0].y);
        } else {
            const cplx<T> a = v[0], b = v[1], c = v[2];
            const T       g00 = g[0], g01 = g[1] * sx * sy, g02 = g[2] * sx, g11 = g[3], g12 = g[4] * sy, g22 = g[5];
            v[0]                = cplx<T>{g00 * a.x + g01 * b.x + g02 * c.x, g00 * a.y + g01 * b.y + g02 * c.y};
            v[1]                = cplx<T>{g01 * a.x + g11 * b.x + g12 * c.x, g01 * a.y + g11 * b.y + g12 * c.y};
            v[2]                = cplx<T>{g02 * a.x + g12 * b.x + g22 * c.x, g02 * a.y + g12 * b.y + g22 * c.y};
            acc -= wk * ((double) a.x * v[0].x + (double) a.y * v[0].y + (double) b.x * v[1].x +
                         (double) b.y * v[1].y + (double) c.x * v[2].x + (double) c.y * v[2].y);
        }
    }
    __syncthreads();
    lds_fft<T, true>(A, L, tw, ncols, ld);
    for (int w = threadIdx.x; w < L.n * ncols; w += blockDim.x) {
        const int i     = w / ncols, c = w - i * ncols;
        base[i * sn + c] = A[i * ld + c];
    }
    if (partial) {
        double     v1[1] = {acc};
        const bool mx[1] = {false};
        block_reduce<1>(v1, mx, red);
        if (threadIdx.x == 0)
            partial[blockIdx.y * gridDim.x + blockIdx.x] = v1[0];
    }
}

}  // namespace fans
