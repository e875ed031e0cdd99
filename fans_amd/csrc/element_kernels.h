// Element-level kernels: residual assembly (gather -> B -> stress -> B^T ->
// scatter) for trilinear hexahedra on a periodic voxel grid.
//
// Node (x,y,z) is the lower corner of element (x,y,z); element corner a has
// offsets (a&1, (a>>1)&1, (a>>2)&1) like include/solver.h:333-349.  The
// gradient at the Gauss points is evaluated in sum-factorised form: edge
// differences along one axis, then bilinear interpolation across the other two
// (the HEX8 B-matrix of include/matmodel.h:157-188 factorises this way), and
// B^T sigma is its exact transpose.  Mandel shear scaling follows
// include/matmodel.h:284-304.
//
// The assembly kernel marches along x.  A workgroup owns a 16x16 (y,z) tile of
// elements and emits the 15x15 nodes whose 4 surrounding elements in that
// plane are all in the tile; the x-neighbour contribution is carried in
// registers from one plane to the next.  Corner forces are combined through LDS
// in a fixed order, so results are deterministic and atomic-free.
#pragma once
#include "common.h"

namespace fans {

enum {
    MODEL_THERMAL_ISO    = 1,
    MODEL_THERMAL_TRI    = 2,
    MODEL_ELASTIC_ISO    = 3,
    MODEL_ELASTIC_TRI    = 4,
    MODEL_J2_NEW         = 5,
    MODEL_PSEUDO_PLASTIC = 6
};

// Per-phase material table.  p[] holds, by kind:
//   THERMAL_ISO  k                      THERMAL_TRI  K (3x3 row major)
//   ELASTIC_ISO  lambda mu              ELASTIC_TRI  C (6x6 row major)
//   J2_NEW       K G sigma_y H_iso      PSEUDO       K G sigma_y h eps_crit E_s
template <typename T>
struct Phase {
    int kind;
    int pad;
    T   p[36];
};

struct Geom {
    double ilx, ily, ilz;  // 1 / element edge lengths
    double w;              // v_e / n_gp
    double q0, q1;         // Gauss point coordinates 0.5 -+ sqrt(3)/6
};

constexpr double kSqrtHalf = 7.071067811865476e-01;  // literal of matmodel.h:288

// ---------------------------------------------------------------------------
// constitutive point updates
// ---------------------------------------------------------------------------

// J2PlasticityNew::get_sigma (J2PlasticityNew.h:43-109), always in FP64
__device__ __forceinline__ void j2_point(const double *p, const double (&e)[6], const double (&ep_in)[6], double q_in,
                                         double (&s)[6], double (&ep)[6], double &qo)
{
    const double c23 = 0.81649658092772603273;  // sqrt(2/3)
    const double G   = p[1];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        s[i] = 2.0 * G * (e[i] - ep_in[i]);
    const double lam = p[0] - 2.0 / 3.0 * G;
    const double tr  = e[0] + e[1] + e[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s[i] += lam * tr;
    const double mean = (s[0] + s[1] + s[2]) / 3.0;
    double       st[6], s_t = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        st[i] = s[i] - (i < 3 ? mean : 0.0);
        s_t += st[i] * st[i];
    }
    s_t                = sqrt(s_t);
    const double sy     = p[2] + q_in * p[3];
    const double phi    = s_t - c23 * sy;
    const double dgamma = fmax(0.0, phi / (2.0 * G + 2.0 / 3.0 * p[3]));
    const bool   ok     = s_t > 1e-12;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        ep[i] = ok ? ep_in[i] + dgamma * st[i] / s_t : ep_in[i];
        if (ok)
            s[i] = s[i] - dgamma * 2.0 * G * st[i] / s_t;
    }
    qo = q_in + c23 * dgamma;
}

// PseudoPlasticLinearHardening::get_sigma (PseudoPlastic.h:92-116)
template <typename T>
__device__ __forceinline__ void pseudo_point(const T *p, const T (&e)[6], T (&s)[6])
{
    const T a = (T) (2. / 3), b = (T) 0.81649658092772603273;
    const T tr = e[0] + e[1] + e[2];
    T       dev[6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        dev[i] = e[i] - (T) (1.0 / 3.0) * tr;
#pragma unroll
    for (int i = 3; i < 6; ++i)
        dev[i] = e[i];
    T nd = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        nd += dev[i] * dev[i];
    nd           = sqrt(nd);
    const T buf1 = p[0] * tr;
    const T buf2 = (nd <= p[4]) ? (T) 2.0 * p[1] : (b * p[2] + a * p[5] * p[3] * (nd - p[4])) / nd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s[i] = buf1 + buf2 * dev[i];
#pragma unroll
    for (int i = 3; i < 6; ++i)
        s[i] = buf2 * dev[i];
}

// State of J2 points, structure of arrays: comp c (0..5 plastic strain, 6 q)
// of Gauss point g of element e at st[(c * ngp + g) * nel + e].
struct J2State {
    double *st;   // committed (_t) state
    int64_t nel;  // number of elements
};

// sigma = model(eps) for one Gauss point; returns in s
template <typename T, int H>
__device__ __forceinline__ void material_point(const Phase<T> &P, const T (&e)[H == 1 ? 3 : 6], T (&s)[H == 1 ? 3 : 6],
                                               const J2State &js, int64_t el, int gp, int ngp)
{
    if constexpr (H == 1) {
        if (P.kind == MODEL_THERMAL_ISO) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[i] = P.p[0] * e[i];
        } else {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[i] = P.p[3 * i] * e[0] + P.p[3 * i + 1] * e[1] + P.p[3 * i + 2] * e[2];
        }
    } else {
        switch (P.kind) {
        case MODEL_ELASTIC_ISO: {
            const T b1 = P.p[0] * (e[0] + e[1] + e[2]), b2 = (T) 2 * P.p[1];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[i] = b1 + b2 * e[i];
#pragma unroll
            for (int i = 3; i < 6; ++i)
                s[i] = b2 * e[i];
        } break;
        case MODEL_ELASTIC_TRI:
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                T a = 0;
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    a += P.p[6 * i + j] * e[j];
                s[i] = a;
            }
            break;
        case MODEL_J2_NEW: {
            double ed[6], epin[6], sd[6], epo[6], qo, pd[4];
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                ed[i]   = (double) e[i];
                epin[i] = js.st[((int64_t) i * ngp + gp) * js.nel + el];
            }
            const double qin = js.st[((int64_t) 6 * ngp + gp) * js.nel + el];
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pd[i] = (double) P.p[i];
            j2_point(pd, ed, epin, qin, sd, epo, qo);
#pragma unroll
            for (int i = 0; i < 6; ++i)
                s[i] = (T) sd[i];
        } break;
        default: pseudo_point<T>(P.p, e, s); break;
        }
    }
}

// ---------------------------------------------------------------------------
// element kernel core: u[8][H] (corner values) -> f[8][H] = B^T sigma(B u + g0) w
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lerp1(T a, T b, T t)
{
    return a + t * (b - a);
}

// strain (Mandel) from displacement gradient g[c][d] = d u_c / d x_d
template <typename T, int H>
__device__ __forceinline__ void grad_to_strain(const T (&g)[H][3], const T *g0, T (&e)[H == 1 ? 3 : 6])
{
    if constexpr (H == 1) {
        e[0] = g[0][0] + g0[0];
        e[1] = g[0][1] + g0[1];
        e[2] = g[0][2] + g0[2];
    } else {
        const T r = (T) kSqrtHalf;
        e[0]      = g[0][0] + g0[0];
        e[1]      = g[1][1] + g0[1];
        e[2]      = g[2][2] + g0[2];
        e[3]      = r * (g[0][1] + g[1][0]) + g0[3];
        e[4]      = r * (g[0][2] + g[2][0]) + g0[4];
        e[5]      = r * (g[1][2] + g[2][1]) + g0[5];
    }
}

// traction-like tensor T[c][d] with f_a[c] = sum_d dN_a/dx_d T[c][d]
template <typename T, int H>
__device__ __forceinline__ void stress_to_T(const T (&s)[H == 1 ? 3 : 6], T (&Tt)[H][3])
{
    if constexpr (H == 1) {
        Tt[0][0] = s[0];
        Tt[0][1] = s[1];
        Tt[0][2] = s[2];
    } else {
        const T r = (T) kSqrtHalf;
        Tt[0][0]  = s[0];
        Tt[1][1]  = s[1];
        Tt[2][2]  = s[2];
        Tt[0][1] = Tt[1][0] = r * s[3];
        Tt[0][2] = Tt[2][0] = r * s[4];
        Tt[1][2] = Tt[2][1] = r * s[5];
    }
}

// Gradient factors.  HEX8: Gx[j][k] at (eta_j, zeta_k), Gy[i][k], Gz[i][j].
// HEX8R: single centre value stored in [0][0].
template <typename T, int H, int NGP>
__device__ __forceinline__ void grad_factors(const T (&u)[8][H], const Geom &g, T (&Gx)[2][2][H], T (&Gy)[2][2][H],
                                             T (&Gz)[2][2][H])
{
    const T ilx = (T) g.ilx, ily = (T) g.ily, ilz = (T) g.ilz;
#pragma unroll
    for (int c = 0; c < H; ++c) {
        // edge differences: Dx[ay][az], Dy[ax][az], Dz[ax][ay]
        T Dx[2][2], Dy[2][2], Dz[2][2];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                Dx[p][q] = (u[1 + 2 * p + 4 * q][c] - u[2 * p + 4 * q][c]) * ilx;
                Dy[p][q] = (u[p + 2 + 4 * q][c] - u[p + 4 * q][c]) * ily;
                Dz[p][q] = (u[p + 2 * q + 4][c] - u[p + 2 * q][c]) * ilz;
            }
        if constexpr (NGP == 8) {
            const T qq[2] = {(T) g.q0, (T) g.q1};
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                // interpolate first index at q_a
                const T x0 = lerp1(Dx[0][0], Dx[1][0], qq[a]), x1 = lerp1(Dx[0][1], Dx[1][1], qq[a]);
                const T y0 = lerp1(Dy[0][0], Dy[1][0], qq[a]), y1 = lerp1(Dy[0][1], Dy[1][1], qq[a]);
                const T z0 = lerp1(Dz[0][0], Dz[1][0], qq[a]), z1 = lerp1(Dz[0][1], Dz[1][1], qq[a]);
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    Gx[a][b][c] = lerp1(x0, x1, qq[b]);
                    Gy[a][b][c] = lerp1(y0, y1, qq[b]);
                    Gz[a][b][c] = lerp1(z0, z1, qq[b]);
                }
            }
        } else {
            Gx[0][0][c] = (T) 0.25 * ((Dx[0][0] + Dx[1][0]) + (Dx[0][1] + Dx[1][1]));
            Gy[0][0][c] = (T) 0.25 * ((Dy[0][0] + Dy[1][0]) + (Dy[0][1] + Dy[1][1]));
            Gz[0][0][c] = (T) 0.25 * ((Dz[0][0] + Dz[1][0]) + (Dz[0][1] + Dz[1][1]));
        }
    }
}

// transpose of grad_factors applied to S (already scaled): accumulates into f
template <typename T, int H, int NGP>
__device__ __forceinline__ void grad_factors_T(const T (&Sx)[2][2][H], const T (&Sy)[2][2][H], const T (&Sz)[2][2][H],
                                               const Geom &g, T (&f)[8][H])
{
    const T ilx = (T) g.ilx, ily = (T) g.ily, ilz = (T) g.ilz;
#pragma unroll
    for (int c = 0; c < H; ++c) {
        T Dx[2][2], Dy[2][2], Dz[2][2];
        if constexpr (NGP == 8) {
            const T q0 = (T) g.q0, q1 = (T) g.q1;
            // back-interpolate second index (N_0(q0)=q1, N_0(q1)=q0, N_1(q)=q)
            T ex[2][2], ey[2][2], ez[2][2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                ex[a][0] = q1 * Sx[a][0][c] + q0 * Sx[a][1][c];
                ex[a][1] = q0 * Sx[a][0][c] + q1 * Sx[a][1][c];
                ey[a][0] = q1 * Sy[a][0][c] + q0 * Sy[a][1][c];
                ey[a][1] = q0 * Sy[a][0][c] + q1 * Sy[a][1][c];
                ez[a][0] = q1 * Sz[a][0][c] + q0 * Sz[a][1][c];
                ez[a][1] = q0 * Sz[a][0][c] + q1 * Sz[a][1][c];
            }
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                Dx[0][b] = q1 * ex[0][b] + q0 * ex[1][b];
                Dx[1][b] = q0 * ex[0][b] + q1 * ex[1][b];
                Dy[0][b] = q1 * ey[0][b] + q0 * ey[1][b];
                Dy[1][b] = q0 * ey[0][b] + q1 * ey[1][b];
                Dz[0][b] = q1 * ez[0][b] + q0 * ez[1][b];
                Dz[1][b] = q0 * ez[0][b] + q1 * ez[1][b];
            }
        } else {
            const T sx = (T) 0.25 * Sx[0][0][c], sy = (T) 0.25 * Sy[0][0][c], sz = (T) 0.25 * Sz[0][0][c];
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    Dx[p][q] = sx;
                    Dy[p][q] = sy;
                    Dz[p][q] = sz;
                }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const T dx = Dx[p][q] * ilx, dy = Dy[p][q] * ily, dz = Dz[p][q] * ilz;
                f[1 + 2 * p + 4 * q][c] += dx;
                f[2 * p + 4 * q][c] -= dx;
                f[p + 2 + 4 * q][c] += dy;
                f[p + 4 * q][c] -= dy;
                f[p + 2 * q + 4][c] += dz;
                f[p + 2 * q][c] -= dz;
            }
    }
}

// Full element: f = sum_gp B^T sigma(B u + g0) * w.  MatF(gp, e, s) evaluates
// the constitutive law at Gauss point gp (index i + 2j + 4k).
template <typename T, int H, int NGP, typename MatF>
__device__ __forceinline__ void element_forces(const T (&u)[8][H], const Geom &g, const T *g0, MatF &&mat,
                                               T (&f)[8][H])
{
    constexpr int NS = H == 1 ? 3 : 6;
    T             Gx[2][2][H], Gy[2][2][H], Gz[2][2][H];
    grad_factors<T, H, NGP>(u, g, Gx, Gy, Gz);
    T Sx[2][2][H], Sy[2][2][H], Sz[2][2][H];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < H; ++c)
                Sx[a][b][c] = Sy[a][b][c] = Sz[a][b][c] = (T) 0;
#pragma unroll
    for (int gp = 0; gp < NGP; ++gp) {
        const int i = gp & 1, j = (gp >> 1) & 1, k = gp >> 2;
        T         gr[H][3];
#pragma unroll
        for (int c = 0; c < H; ++c) {
            gr[c][0] = Gx[j][k][c];
            gr[c][1] = Gy[i][k][c];
            gr[c][2] = Gz[i][j][c];
        }
        T e[NS], s[NS], Tt[H][3];
        grad_to_strain<T, H>(gr, g0, e);
        mat(gp, e, s);
        stress_to_T<T, H>(s, Tt);
#pragma unroll
        for (int c = 0; c < H; ++c) {
            Sx[j][k][c] += Tt[c][0];
            Sy[i][k][c] += Tt[c][1];
            Sz[i][j][c] += Tt[c][2];
        }
    }
    const T w = (T) g.w;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int c = 0; c < H; ++c) {
                Sx[a][b][c] *= w;
                Sy[a][b][c] *= w;
                Sz[a][b][c] *= w;
            }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < H; ++c)
            f[a][c] = (T) 0;
    grad_factors_T<T, H, NGP>(Sx, Sy, Sz, g, f);
}

// ---------------------------------------------------------------------------
// assembly kernel (x-marching tile)
// ---------------------------------------------------------------------------
enum { ASM_RESID = 0, ASM_CGLIN = 1 };
constexpr int TILE = 16;  // element tile edge (y and z)
constexpr int NOUT = TILE - 1;

template <typename T>
struct AsmArgs {
    int            nx, ny, nz;
    int            cx;          // x planes per block
    int            nty, ntz;    // number of tiles in y, z
    Geom           geom;
    const uint8_t *ms;          // phase ids [x][y][z]
    const Phase<T> *phases;
    J2State        js;
    // ASM_RESID: in = u, out = r; norm/dot partials
    // ASM_CGLIN: in = s, din = d_cur, dout = d_next (d = s + beta d), out = rnew
    const T       *in;
    const T       *din;
    T             *dout;
    T             *out;
    const T       *dotv;        // ASM_RESID optional: partial <out, dotv>
    const double  *scal;        // device scalars (beta at scal[0] for CGLIN)
    const T       *g0;          // macroscopic gradient (n_str), device
    double        *partial;     // [2][gridDim] : slot 0 = norm/dot, slot 1 = dot
    int            measure;     // 0 L1, 1 L2, 2 Linf  (ASM_RESID)
    const int     *done;
};

template <typename T, int H, int NGP, int MODE>
__global__ __launch_bounds__(256) void k_assemble(AsmArgs<T> a)
{
    if (a.done && *a.done)
        return;
    constexpr int NL = TILE + 1;  // node tile edge
    __shared__ T      Un[2][NL * NL * H];
    __shared__ T      Xz[TILE * TILE * 4 * H];
    __shared__ T      Xy[TILE * TILE * 2 * H];
    __shared__ double red[8];

    const int t = threadIdx.x, ey = t / TILE, ez = t - ey * TILE;
    int       bid  = blockIdx.x;
    const int tz   = bid % a.ntz;
    bid /= a.ntz;
    const int ty   = bid % a.nty;
    const int bx   = bid / a.nty;
    const int Y0   = ty * NOUT, Z0 = tz * NOUT, X0 = bx * a.cx;
    const int X1   = min(X0 + a.cx, a.nx);
    const int ny = a.ny, nz = a.nz, nx = a.nx;
    const int64_t plane = (int64_t) ny * nz;

    // node tile coords (ty_, tz_) in [0,17) map to global (Y0-1+ty_, Z0-1+tz_) mod n
    auto gy = [&](int v) { int y = Y0 - 1 + v; y += (y < 0) ? ny : 0; return y >= ny ? y - ny : y; };
    auto gz = [&](int v) { int z = Z0 - 1 + v; z += (z < 0) ? nz : 0; return z >= nz ? z - nz : z; };
    const int  my_y   = Y0 - 1 + ey, my_z = Z0 - 1 + ez;
    const bool own_yz = ey >= 1 && ez >= 1 && my_y < ny && my_z < nz;
    T          beta   = (T) 0;
    if constexpr (MODE == ASM_CGLIN)
        beta = (T) a.scal[0];

    auto load_plane = [&](int xp, int buf) {
        const int     xg = xp < 0 ? xp + nx : (xp >= nx ? xp - nx : xp);
        const bool    wr = (MODE == ASM_CGLIN) && xp >= X0 && xp < X1;
        for (int v = t; v < NL * NL; v += 256) {
            const int     a1 = v < 256 ? v / TILE : (v - 256 < NL ? TILE : (v - 256 - NL));
            const int     b1 = v < 256 ? v - (v / TILE) * TILE : (v - 256 < NL ? v - 256 : TILE);
            const int64_t n  = xg * plane + (int64_t) gy(a1) * nz + gz(b1);
#pragma unroll
            for (int c = 0; c < H; ++c) {
                T val;
                if constexpr (MODE == ASM_CGLIN) {
                    val = a.in[n * H + c] + beta * a.din[n * H + c];
                    if (wr && v < 256 && own_yz)
                        a.dout[n * H + c] = val;
                } else {
                    val = a.in[n * H + c];
                }
                Un[buf][(a1 * NL + b1) * H + c] = val;
            }
        }
    };

    T      carry[H];
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int c = 0; c < H; ++c)
        carry[c] = (T) 0;
    T g0[H == 1 ? 3 : 6];
#pragma unroll
    for (int i = 0; i < (H == 1 ? 3 : 6); ++i)
        g0[i] = (MODE == ASM_RESID) ? a.g0[i] : (T) 0;

    load_plane(X0 - 1, 0);
    for (int i = 0; X0 - 1 + i < X1; ++i) {
        const int xe = X0 - 1 + i;
        load_plane(xe + 1, (i + 1) & 1);
        __syncthreads();
        T u[8][H];
#pragma unroll
        for (int cr = 0; cr < 8; ++cr) {
            const int ax = cr & 1, ay = (cr >> 1) & 1, az = cr >> 2;
            const T  *src = &Un[(i + ax) & 1][((ey + ay) * NL + ez + az) * H];
#pragma unroll
            for (int c = 0; c < H; ++c)
                u[cr][c] = src[c];
        }
        T dv[H];  // d_next at node plane xe (CGLIN); read before the buffer is recycled
#pragma unroll
        for (int c = 0; c < H; ++c)
            dv[c] = Un[i & 1][(ey * NL + ez) * H + c];
        const int     xg  = xe < 0 ? xe + nx : xe;
        const int64_t el  = xg * plane + (int64_t) gy(ey) * nz + gz(ez);
        const Phase<T> &P = a.phases[a.ms[el]];
        T f[8][H];
        auto mat = [&](int gp, const T(&e)[H == 1 ? 3 : 6], T(&s)[H == 1 ? 3 : 6]) {
            material_point<T, H>(P, e, s, a.js, el, gp, NGP);
        };
        element_forces<T, H, NGP>(u, a.geom, g0, mat, f);
        // combine along z: node (ey, ez) gets az=0 corners of element ez and az=1 of element ez-1
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < H; ++c)
                Xz[(t * 4 + q) * H + c] = f[q + 4][c];
        __syncthreads();
        T P2[2][2][H];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < H; ++c)
                P2[q & 1][q >> 1][c] = f[q][c] + (ez >= 1 ? Xz[((t - 1) * 4 + q) * H + c] : (T) 0);
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
            for (int c = 0; c < H; ++c)
                Xy[(t * 2 + ax) * H + c] = P2[ax][1][c];
        __syncthreads();
        T Q[2][H];
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
            for (int c = 0; c < H; ++c)
                Q[ax][c] = P2[ax][0][c] + (ey >= 1 ? Xy[((t - TILE) * 2 + ax) * H + c] : (T) 0);
        if (i >= 1 && own_yz) {
            const int64_t n = xe * plane + (int64_t) my_y * nz + my_z;
#pragma unroll
            for (int c = 0; c < H; ++c) {
                const T res    = carry[c] + Q[0][c];
                a.out[n * H + c] = res;
                if constexpr (MODE == ASM_CGLIN) {
                    acc0 += (double) dv[c] * (double) res;
                } else {
                    const double ar = fabs((double) res);
                    if (a.measure == 2)
                        acc0 = fmax(acc0, ar);
                    else if (a.measure == 1)
                        acc0 += ar * ar;
                    else
                        acc0 += ar;
                    if (a.dotv)
                        acc1 += (double) res * (double) a.dotv[n * H + c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < H; ++c)
            carry[c] = Q[1][c];
    }
    if (a.partial) {
        double     v[2]  = {acc0, acc1};
        const bool mx[2] = {MODE == ASM_RESID && a.measure == 2, false};
        block_reduce<2>(v, mx, red);
        if (t == 0) {
            a.partial[blockIdx.x]             = v[0];
            a.partial[gridDim.x + blockIdx.x] = v[1];
        }
    }
}

// ---------------------------------------------------------------------------
// element-local pass (one thread per element): homogenized stress, fields,
// committing the J2 state.  Reads nodal u directly (no relative gather).
// ---------------------------------------------------------------------------
enum { ELEM_HOMOG = 0, ELEM_FIELDS = 1, ELEM_COMMIT = 2 };

template <typename T>
struct ElemArgs {
    int             nx, ny, nz;
    Geom            geom;
    const uint8_t  *ms;
    const Phase<T> *phases;
    J2State         js;
    const T        *u;
    const T        *g0;
    double         *partial;       // HOMOG: [NS][gridDim]
    double         *strain_elem;   // FIELDS (nullable)
    double         *stress_elem;
    double         *strain_gp;
    double         *stress_gp;
};

template <typename T, int H, int NGP, int MODE>
__global__ __launch_bounds__(256) void k_element_pass(ElemArgs<T> a)
{
    constexpr int NS = H == 1 ? 3 : 6;
    __shared__ double red[NS * 4];
    const int64_t     nel = (int64_t) a.nx * a.ny * a.nz;
    const int64_t     el  = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    double            sacc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k)
        sacc[k] = 0.0;
    if (el < nel) {
        const int z = (int) (el % a.nz), y = (int) ((el / a.nz) % a.ny), x = (int) (el / ((int64_t) a.nz * a.ny));
        const int x1 = x + 1 == a.nx ? 0 : x + 1, y1 = y + 1 == a.ny ? 0 : y + 1, z1 = z + 1 == a.nz ? 0 : z + 1;
        T         u[8][H];
#pragma unroll
        for (int cr = 0; cr < 8; ++cr) {
            const int64_t n = ((int64_t) ((cr & 1) ? x1 : x) * a.ny + ((cr & 2) ? y1 : y)) * a.nz + ((cr & 4) ? z1 : z);
#pragma unroll
            for (int c = 0; c < H; ++c)
                u[cr][c] = a.u[n * H + c];
        }
        const Phase<T> &P = a.phases[a.ms[el]];
        T               Gx[2][2][H], Gy[2][2][H], Gz[2][2][H];
        grad_factors<T, H, NGP>(u, a.geom, Gx, Gy, Gz);
        T g0[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g0[k] = a.g0[k];
        double eavg[NS], savg[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            eavg[k] = savg[k] = 0.0;
#pragma unroll
        for (int gp = 0; gp < NGP; ++gp) {
            const int i = gp & 1, j = (gp >> 1) & 1, k = gp >> 2;
            T         gr[H][3];
#pragma unroll
            for (int c = 0; c < H; ++c) {
                gr[c][0] = Gx[j][k][c];
                gr[c][1] = Gy[i][k][c];
                gr[c][2] = Gz[i][j][c];
            }
            T e[NS], s[NS];
            grad_to_strain<T, H>(gr, g0, e);
            if constexpr (MODE == ELEM_COMMIT) {
                if (P.kind == MODEL_J2_NEW) {
                    double ed[6], epin[6], sd[6], epo[6], qo, pd[4];
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        ed[c]   = (double) e[c];
                        epin[c] = a.js.st[((int64_t) c * NGP + gp) * a.js.nel + el];
                    }
                    const double qin = a.js.st[((int64_t) 6 * NGP + gp) * a.js.nel + el];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        pd[c] = (double) P.p[c];
                    j2_point(pd, ed, epin, qin, sd, epo, qo);
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        a.js.st[((int64_t) c * NGP + gp) * a.js.nel + el] = epo[c];
                    a.js.st[((int64_t) 6 * NGP + gp) * a.js.nel + el] = qo;
                }
            } else {
                material_point<T, H>(P, e, s, a.js, el, gp, NGP);
#pragma unroll
                for (int c = 0; c < NS; ++c) {
                    eavg[c] += (double) e[c] / NGP;
                    savg[c] += (double) s[c] / NGP;
                }
                if constexpr (MODE == ELEM_FIELDS) {
                    if (a.strain_gp)
#pragma unroll
                        for (int c = 0; c < NS; ++c)
                            a.strain_gp[(el * NGP + gp) * NS + c] = (double) e[c];
                    if (a.stress_gp)
#pragma unroll
                        for (int c = 0; c < NS; ++c)
                            a.stress_gp[(el * NGP + gp) * NS + c] = (double) s[c];
                }
            }
        }
        if constexpr (MODE == ELEM_FIELDS) {
#pragma unroll
            for (int c = 0; c < NS; ++c) {
                if (a.strain_elem)
                    a.strain_elem[el * NS + c] = eavg[c];
                if (a.stress_elem)
                    a.stress_elem[el * NS + c] = savg[c];
            }
        }
#pragma unroll
        for (int c = 0; c < NS; ++c)
            sacc[c] = savg[c];
    }
    if constexpr (MODE == ELEM_HOMOG) {
        bool mx[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            mx[k] = false;
        block_reduce<NS>(sacc, mx, red);
        if (threadIdx.x == 0)
#pragma unroll
            for (int k = 0; k < NS; ++k)
                a.partial[(int64_t) k * gridDim.x + blockIdx.x] = sacc[k];
    }
}

}  // namespace fans
