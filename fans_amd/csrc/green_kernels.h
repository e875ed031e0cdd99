// Green operator of the homogeneous reference medium, one thread per
// frequency; restates Solver::computeFundamentalSolution (solver.h:144-204).
//
// For frequency (ix, iy, iz) the eight element-corner phase factors
// A = [1, ex, ey, ex ey, ez, ex ez, ez ey, ex ey ez] (e_d = exp(2 pi i k_d / n_d))
// give AA = Re A Re A^T + Im A Im A^T; block_ij = sum(Ke_ref[8i:, 8j:] o AA);
// G = pinv(block, absolute cutoff 1e-14) / N, and G(0) = 0.
#pragma once
#include "common.h"

namespace fans {

// symmetric eigen-decomposition by cyclic Jacobi rotations (n <= 3)
template <int N>
__device__ __forceinline__ void jacobi_eig(double (&a)[N][N], double (&V)[N][N])
{
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
            V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, tot = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i)
#pragma unroll
            for (int j = 0; j < N; ++j) {
                tot += a[i][j] * a[i][j];
                if (i != j)
                    off += a[i][j] * a[i][j];
            }
        if (off <= 1e-32 * tot || off == 0.0)
            break;
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = a[p][q];
                if (apq == 0.0)
                    continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                const double t     = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p]          = c * akp - s * akq;
                    a[k][q]          = s * akp + c * akq;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k]          = c * apk - s * aqk;
                    a[q][k]          = s * apk + c * aqk;
                }
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p]          = c * vkp - s * vkq;
                    V[k][q]          = s * vkp + c * vkq;
                }
            }
    }
}

// keref: component-major (8H)x(8H) reference element stiffness (matmodel.h:237-253)
// G: [gxq][gyq][nz/2+1][H(H+1)/2] packed (00 01 02 11 12 22) for the stored
// frequencies kx < gxq, ky < gyq (see GreenView)
template <int H>
__global__ __launch_bounds__(256) void k_build_green(const double *__restrict__ keref, double *__restrict__ G, int nx,
                                                     int ny, int nz, int gxq, int gyq)
{
    constexpr int NP = H * (H + 1) / 2, NE = 8 * H;
    const int     Mc = nz / 2 + 1;
    const int64_t nf = (int64_t) gxq * gyq * Mc;
    const int64_t f  = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf)
        return;
    const int iz = (int) (f % Mc), iy = (int) ((f / Mc) % gyq), ix = (int) (f / ((int64_t) Mc * gyq));
    double   *g  = G + f * NP;
    if (ix == 0 && iy == 0 && iz == 0) {
#pragma unroll
        for (int k = 0; k < NP; ++k)
            g[k] = 0.0;
        return;
    }
    const double tpi = 6.283185307179586;  // 2 * acos(-1)
    double       sx, cx, sy, cy, sz, cz;
    sincos(tpi * ix / (double) nx, &sx, &cx);
    sincos(tpi * iy / (double) ny, &sy, &cy);
    sincos(tpi * iz / (double) nz, &sz, &cz);
    const cplx<double> ex{cx, sx}, ey{cy, sy}, ez{cz, sz};
    cplx<double>       A[8];
    A[0] = cplx<double>{1.0, 0.0};
    A[1] = ex;
    A[2] = ey;
    A[3] = cmul(ex, ey);
    A[4] = ez;
    A[5] = cmul(ex, ez);
    A[6] = cmul(ez, ey);
    A[7] = cmul(cmul(ex, ey), ez);
    double blk[H][H];
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = i; j < H; ++j) {
            double acc = 0.0;
            for (int a = 0; a < 8; ++a)
                for (int b = 0; b < 8; ++b)
                    acc += keref[(8 * i + a) * NE + 8 * j + b] * (A[a].x * A[b].x + A[a].y * A[b].y);
            blk[i][j] = acc;
            blk[j][i] = acc;
        }
    double V[H][H];
    jacobi_eig<H>(blk, V);
    const double Nd = (double) ((int64_t) nx * ny * nz);  // solver.h:198 divides by N
    int          k  = 0;
#pragma unroll
    for (int i = 0; i < H; ++i)
#pragma unroll
        for (int j = i; j < H; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int m = 0; m < H; ++m)
                if (fabs(blk[m][m]) > 1e-14)
                    acc += V[i][m] * V[j][m] / blk[m][m];
            g[k++] = acc / Nd;
        }
}

}  // namespace fans
