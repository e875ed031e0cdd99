/*
 * fans_oracle.c -- CPU FP64 restatement of the FANS per-iteration solver path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path in fans_amd/.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library; the product never routes through it.
 *
 * Each function cites the reference file:line (FANS 0.6.2) it restates.  The
 * restatement is single-rank: the ghost x-plane of u holds a copy of plane 0
 * and the ghost plane of r is folded back into plane 0 after the element loop,
 * exactly as the MPI_Sendrecv pair of solver.h:244-245,264-269 does at P=1.
 *
 * Parity status of each part (pins live in tests/test_oracle_*.py):
 *   FFT r2c/c2r            pinned: numpy.fft.rfftn/irfftn
 *   B, Ke, Ke_ref          pinned: partition of unity, exact linear gradients,
 *                                  symmetry/PSD/translation null space
 *   residual assembly      pinned: brute-force sparse assembly in numpy
 *   Green operator         pinned: G(xi) K(xi) = I from plane-wave probing
 *   CG / FP                pinned: scipy direct solve on tiny grids, homogeneous
 *                                  RVE, laminate closed form
 *   J2PlasticityNew        pinned: textbook radial return (Simo & Hughes Box 3.1)
 *   mixed BC + secant LS   pinned: golden macro strains of test_MixedBCs.json
 *   homogenized tangent    pinned: symmetry/SPD, Voigt/Reuss bounds, laminate
 *   compute_error          pinned: numpy L1 / L2 / L-inf norms of the first
 *                                  residual; relative-type stop rule with the
 *                                  iteration-0 constant 100
 *   triclinic models       pinned: homogeneous RVE Sigma = C g0 with C from the
 *                                  C_ij / K_ij keys; sparse assembly with full
 *                                  anisotropic matrices; mean reference stiffness
 *   PseudoPlastic (linear  pinned: deformation-plasticity closed form, elastic
 *   and power-law)                 3G and hardening H in series beyond sigma_y;
 *                                  power law sigma_y (eps_eq/eps_0)^n; golden
 *                                  strains of test_MixedBCs.json
 *   finite strain          pinned: dP/dF by differences of the stored energy,
 *                                  objectivity, golden F / P of
 *                                  test_MixedBCs_LargeStrain.json
 *   unpinned: exact CG iteration counts and error histories of the reference
 *   binary (FFTW / Eigen operation order), only restated
 */
#include "fans_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define MAXH 3
#define MAXSTR 9
#define MAXE 24 /* 8 * howmany */

/* ------------------------------------------------------------------------ */
/* FFT: iterative Stockham autosort, radix 4 / 2 / 3 / 5 (any other prime by */
/* a direct DFT); real transforms through a complex transform of half the    */
/* length (the unnormalised r2c / c2r of FFTW, which solver.h:221-222 uses)   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int   n;
    int   nf;
    int   f[40];
    cplx *w; /* w[k] = exp(-2 pi i k / n) */
} fft_plan;

static void plan_init(fft_plan *P, int n)
{
    P->n  = n;
    P->nf = 0;
    int m = n;
    while (m % 4 == 0) { P->f[P->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { P->f[P->nf++] = 2; m /= 2; }
    for (int p = 3; m > 1; p += 2) {
        while (m % p == 0) { P->f[P->nf++] = p; m /= p; }
    }
    P->w = (cplx *) malloc(sizeof(cplx) * (size_t) n);
    for (int k = 0; k < n; k++) {
        long double th = -2.0L * 3.14159265358979323846264338327950288L * (long double) k / (long double) n;
        P->w[k]        = (double) cosl(th) + I * (double) sinl(th);
    }
}

static void plan_free(fft_plan *P) { free(P->w); }

/* in-place DFT of length p (sign -1 forward, +1 inverse) */
static void dft_small(cplx *v, int p, int sign, const fft_plan *P)
{
    const double s3 = 0.86602540378443864676, c5a = 0.30901699437494742410, c5b = -0.80901699437494742410,
                 s5a = 0.95105651629515357212, s5b = 0.58778525229247312917;
    const cplx   mi = sign < 0 ? -I : I; /* exp(sign * i pi / 2) */
    if (p == 2) {
        cplx a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if (p == 4) {
        cplx s02 = v[0] + v[2], d02 = v[0] - v[2], s13 = v[1] + v[3], d13 = v[1] - v[3];
        v[0] = s02 + s13;
        v[1] = d02 + mi * d13;
        v[2] = s02 - s13;
        v[3] = d02 - mi * d13;
    } else if (p == 3) {
        cplx t1 = v[1] + v[2], t2 = v[0] - 0.5 * t1, t3 = mi * s3 * (v[1] - v[2]);
        v[0] = v[0] + t1;
        v[1] = t2 + t3;
        v[2] = t2 - t3;
    } else if (p == 5) {
        cplx a1 = v[1] + v[4], a2 = v[2] + v[3], b1 = v[1] - v[4], b2 = v[2] - v[3];
        cplx r1 = v[0] + c5a * a1 + c5b * a2, r2 = v[0] + c5b * a1 + c5a * a2;
        cplx i1 = mi * (s5a * b1 + s5b * b2), i2 = mi * (s5b * b1 - s5a * b2);
        v[0] = v[0] + a1 + a2;
        v[1] = r1 + i1;
        v[4] = r1 - i1;
        v[2] = r2 + i2;
        v[3] = r2 - i2;
    } else {
        cplx  stackbuf[16];
        cplx *tmp = p <= 16 ? stackbuf : (cplx *) malloc(sizeof(cplx) * (size_t) p);
        long  st  = P->n / p;
        for (int t = 0; t < p; t++) {
            cplx acc = 0;
            for (int q = 0; q < p; q++) {
                cplx w = P->w[((long) t * q % p) * st];
                acc += v[q] * (sign < 0 ? w : conj(w));
            }
            tmp[t] = acc;
        }
        memcpy(v, tmp, sizeof(cplx) * (size_t) p);
        if (tmp != stackbuf)
            free(tmp);
    }
}

/* one line of length P->n in a[] (work: scratch of the same length) */
static void fft_line(const fft_plan *P, cplx *a, cplx *work, int sign)
{
    const int n = P->n;
    cplx     *src = a, *dst = work;
    int       Ns  = 1;
    for (int s = 0; s < P->nf; s++) {
        const int  p  = P->f[s], nb = n / p;
        const long ts = n / (Ns * p);
        cplx       v[64];
        cplx      *vv = p <= 64 ? v : (cplx *) malloc(sizeof(cplx) * (size_t) p);
        for (int j = 0; j < nb; j++) {
            const int k = j % Ns;
            for (int t = 0; t < p; t++) {
                cplx x = src[j + t * nb];
                if (t && k) {
                    cplx w = P->w[(long) t * k * ts];
                    x *= sign < 0 ? w : conj(w);
                }
                vv[t] = x;
            }
            dft_small(vv, p, sign, P);
            const int o = (j / Ns) * Ns * p + k;
            for (int t = 0; t < p; t++)
                dst[o + t * Ns] = vv[t];
        }
        if (vv != v)
            free(vv);
        Ns *= p;
        cplx *tmp = src;
        src       = dst;
        dst       = tmp;
    }
    if (src != a)
        memcpy(a, src, sizeof(cplx) * (size_t) n);
}

/* complex transform of every line of one axis of a [nx][ny][nzc][h] array,
 * axis 0 (x) or 1 (y); lines of consecutive columns go through a contiguous
 * buffer in groups of CB */
#define CB 8
static void fft_axis(cplx *F, int nx, int ny, int nzc, int h, int axis, int sign)
{
    const int       n      = axis == 0 ? nx : ny;
    const ptrdiff_t stride = axis == 0 ? (ptrdiff_t) ny * nzc * h : (ptrdiff_t) nzc * h;
    const ptrdiff_t ncol   = (ptrdiff_t) nzc * h;                       /* contiguous columns per row */
    const ptrdiff_t nouter = axis == 0 ? ny : nx;                       /* rows of columns */
    const ptrdiff_t ostep  = axis == 0 ? (ptrdiff_t) nzc * h : (ptrdiff_t) ny * nzc * h;
    const ptrdiff_t ngrp   = (ncol + CB - 1) / CB;
    fft_plan P;
    plan_init(&P, n);
#pragma omp parallel
    {
        cplx *buf  = (cplx *) malloc(sizeof(cplx) * (size_t) n * CB);
        cplx *work = (cplx *) malloc(sizeof(cplx) * (size_t) n);
#pragma omp for schedule(static)
        for (ptrdiff_t L = 0; L < nouter * ngrp; L++) {
            const ptrdiff_t o = L / ngrp, c0 = (L % ngrp) * CB;
            const int       nc = (int) (ncol - c0 < CB ? ncol - c0 : CB);
            cplx           *base = F + o * ostep + c0;
            for (int i = 0; i < n; i++)
                for (int c = 0; c < nc; c++)
                    buf[(size_t) c * n + i] = base[i * stride + c];
            for (int c = 0; c < nc; c++)
                fft_line(&P, buf + (size_t) c * n, work, sign);
            for (int i = 0; i < n; i++)
                for (int c = 0; c < nc; c++)
                    base[i * stride + c] = buf[(size_t) c * n + i];
        }
        free(buf);
        free(work);
    }
    plan_free(&P);
}
#undef CB

/* unnormalised 3-D r2c like fftw_mpi_plan_many_dft_r2c (solver.h:221): z by a
 * half-length complex transform, X[k] = (Z[k] + conj Z[M-k]) / 2 + W^k (Z[k] -
 * conj Z[M-k]) / (2i), then y and x */
void or_fft_r2c(int nx, int ny, int nz, int h, const double *in, double *out_c)
{
    const int M = nz / 2, nc = M + 1;
    cplx     *o = (cplx *) out_c;
    fft_plan  Pz, Pw;
    plan_init(&Pz, M);
    plan_init(&Pw, nz);
#pragma omp parallel
    {
        cplx *z = (cplx *) malloc(sizeof(cplx) * (size_t) M), *work = (cplx *) malloc(sizeof(cplx) * (size_t) M);
#pragma omp for schedule(static)
        for (ptrdiff_t L = 0; L < (ptrdiff_t) nx * ny * h; L++) {
            const ptrdiff_t xy = L / h, c = L % h;
            const double   *x  = in + xy * nz * h + c;
            for (int m = 0; m < M; m++)
                z[m] = x[(2 * m) * h] + I * x[(2 * m + 1) * h];
            fft_line(&Pz, z, work, -1);
            for (int k = 0; k <= M; k++) {
                const cplx zk = z[k == M ? 0 : k], zc = conj(z[k == 0 ? 0 : M - k]);
                o[(xy * nc + k) * h + c] = 0.5 * (zk + zc) + Pw.w[k] * (zk - zc) / (2.0 * I);
            }
        }
        free(z);
        free(work);
    }
    plan_free(&Pz);
    plan_free(&Pw);
    fft_axis(o, nx, ny, nc, h, 1, -1);
    fft_axis(o, nx, ny, nc, h, 0, -1);
}

/* unnormalised 3-D c2r like fftw_mpi_plan_many_dft_c2r (solver.h:222): x, y,
 * then z through Z[k] = (X[k] + conj X[M-k]) + i W^-k (X[k] - conj X[M-k]) and a
 * half-length inverse transform (output = n_x n_y n_z x irfftn) */
void or_fft_c2r(int nx, int ny, int nz, int h, const double *in_c, double *out)
{
    const int       M = nz / 2, nc = M + 1;
    const ptrdiff_t NC = (ptrdiff_t) nx * ny * nc * h;
    cplx           *F  = (cplx *) malloc(sizeof(cplx) * (size_t) NC);
    memcpy(F, in_c, sizeof(cplx) * (size_t) NC);
    fft_axis(F, nx, ny, nc, h, 0, +1);
    fft_axis(F, nx, ny, nc, h, 1, +1);
    fft_plan Pz, Pw;
    plan_init(&Pz, M);
    plan_init(&Pw, nz);
#pragma omp parallel
    {
        cplx *z = (cplx *) malloc(sizeof(cplx) * (size_t) M), *work = (cplx *) malloc(sizeof(cplx) * (size_t) M);
#pragma omp for schedule(static)
        for (ptrdiff_t L = 0; L < (ptrdiff_t) nx * ny * h; L++) {
            const ptrdiff_t xy = L / h, c = L % h;
            const cplx     *X  = F + xy * nc * h + c;
            for (int k = 0; k < M; k++) {
                const cplx xk = X[k * h], xm = conj(X[(M - k) * h]);
                z[k]          = (xk + xm) + I * conj(Pw.w[k]) * (xk - xm);
            }
            fft_line(&Pz, z, work, +1);
            double *x = out + xy * nz * h + c;
            for (int m = 0; m < M; m++) {
                x[(2 * m) * h]     = creal(z[m]);
                x[(2 * m + 1) * h] = cimag(z[m]);
            }
        }
        free(z);
        free(work);
    }
    plan_free(&Pz);
    plan_free(&Pw);
    free(F);
}

/* ------------------------------------------------------------------------ */
/* small dense helpers                                                       */
/* ------------------------------------------------------------------------ */

/* eigen-decomposition of a symmetric n x n (n <= 9) matrix by cyclic Jacobi
 * rotations; used for the pseudo-inverse of solver.h:89-96 (JacobiSVD with an
 * absolute cutoff on singular values = |eigenvalues| for symmetric input) */
static void sym_eig(int n, const double *A, double *lam, double *V)
{
    double a[81];
    memcpy(a, A, sizeof(double) * n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            V[i * n + j] = (i == j);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += a[i * n + j] * a[i * n + j];
                if (i != j)
                    off += a[i * n + j] * a[i * n + j];
            }
        if (off <= 1e-32 * tot || off == 0.0)
            break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (apq == 0.0)
                    continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2 * apq);
                double t     = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++)
        lam[i] = a[i * n + i];
}

static void sym_pinv(int n, const double *A, double cutoff, double *P)
{
    double lam[9], V[81];
    sym_eig(n, A, lam, V);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double acc = 0;
            for (int k = 0; k < n; k++)
                if (fabs(lam[k]) > cutoff)
                    acc += V[i * n + k] * V[j * n + k] / lam[k];
            P[i * n + j] = acc;
        }
}

/* inverse of a general n x n (n <= 9) matrix by Gauss-Jordan elimination with
 * partial pivoting (the complete orthogonal decomposition pinv of
 * mixedBCs.h:43 for a nonsingular, non-symmetric block); nonzero if singular */
static int gauss_inverse(int n, const double *A, double *Ai)
{
    double a[9][18];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++)
            a[i][j] = j < n ? A[i * n + j] : (j - n == i);
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (fabs(a[r][c]) > fabs(a[piv][c]))
                piv = r;
        if (a[piv][c] == 0.0)
            return -1;
        for (int j = 0; j < 2 * n; j++) {
            double t    = a[c][j];
            a[c][j]     = a[piv][j];
            a[piv][j]   = t;
        }
        double d = a[c][c];
        for (int j = 0; j < 2 * n; j++)
            a[c][j] /= d;
        for (int r = 0; r < n; r++)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 2 * n; j++)
                    a[r][j] -= f * a[c][j];
            }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            Ai[i * n + j] = a[i][n + j];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* solver state                                                              */
/* ------------------------------------------------------------------------ */
typedef struct {
    int     kind;
    int     nphase;
    int     np;
    double *params; /* [nphase][np] */
} or_group;

struct or_solver {
    or_desc   cfg;
    int       nx, ny, nz, h, nstr, ngp;
    ptrdiff_t N, Nf;
    double    le[3], ve;
    double    Bint[8][MAXSTR * MAXE]; /* B_int[p]: nstr x 8h, row major */
    /* nonzero pattern of B_int (skipping exact zeros leaves every sum unchanged) */
    int       nzr[8][MAXSTR], nzr_col[8][MAXSTR][MAXE];
    int       nzc[8][MAXE], nzc_row[8][MAXE][MAXSTR];
    uint16_t *ms;
    int       ngroups;
    or_group *groups;
    int       nphases;
    int      *phase_group, *phase_local;
    int       all_linear;
    double    kref[MAXSTR * MAXSTR];
    double    keref[MAXE * MAXE];
    double   *Ke; /* [nphases][8h][8h], linear phases only */
    double   *G;  /* [Nf][h(h+1)/2]: (00 01 02 11 12 22) */
    double   *u, *u_prev, *r, *s, *d, *rnew, *rghost;
    double    g0[MAXSTR];
    double   *ep, *ep_t, *q, *q_t; /* per element per Gauss point state */
    double   *pb, *pb_t;           /* J2ViscoPlastic kinematic variable psi_bar (6 per point) */
    int       has_state, has_visco;
    double    alpha_warm;
    double   *err_all;
    int       iter;
    double    tol;
    int       err_type;
    /* mixed BC (mixedBCs.h:150-227) */
    int       mixed_active;
    int       nE, nF, nsteps, step;
    int       idxE[9], idxF[9];
    double   *FE, *PF;
    double    M[81];
    double    gvec[9], gvec_prev[9];
};

int or_model_nparams(int kind)
{
    switch (kind) {
    case OR_MODEL_THERMAL_ISO: return 1;
    case OR_MODEL_THERMAL_TRI: return 6;
    case OR_MODEL_ELASTIC_ISO: return 2;
    case OR_MODEL_ELASTIC_TRI: return 21;
    case OR_MODEL_J2_NEW: return 4;
    case OR_MODEL_PSEUDO_PLASTIC: return 4;
    case OR_MODEL_PSEUDO_NL: return 5;
    case OR_MODEL_SVK: return 2;
    case OR_MODEL_NEO_HOOKE: return 2;
    case OR_MODEL_J2_VISCO_LIN: return 7;
    case OR_MODEL_J2_VISCO_NL: return 9;
    default: return -1;
    }
}

static int model_is_linear(int kind)
{
    return kind == OR_MODEL_THERMAL_ISO || kind == OR_MODEL_THERMAL_TRI || kind == OR_MODEL_ELASTIC_ISO ||
           kind == OR_MODEL_ELASTIC_TRI;
}

void or_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void) n;
#endif
}

/* matmodel.h:157-188 */
static void basic_B(const or_solver *S, double x, double y, double z, double o[3][8])
{
    const double lx = S->le[0], ly = S->le[1], lz = S->le[2];
    o[0][0] = -(1. - y) * (1. - z) / lx;
    o[0][1] = (1. - y) * (1. - z) / lx;
    o[0][2] = -y * (1. - z) / lx;
    o[0][3] = y * (1. - z) / lx;
    o[0][4] = -(1. - y) * z / lx;
    o[0][5] = (1. - y) * z / lx;
    o[0][6] = -y * z / lx;
    o[0][7] = y * z / lx;

    o[1][0] = -(1. - x) * (1. - z) / ly;
    o[1][1] = -x * (1. - z) / ly;
    o[1][2] = (1. - x) * (1. - z) / ly;
    o[1][3] = x * (1. - z) / ly;
    o[1][4] = -(1. - x) * z / ly;
    o[1][5] = -x * z / ly;
    o[1][6] = (1. - x) * z / ly;
    o[1][7] = x * z / ly;

    o[2][0] = -(1. - x) * (1. - y) / lz;
    o[2][1] = -x * (1. - y) / lz;
    o[2][2] = -(1. - x) * y / lz;
    o[2][3] = -x * y / lz;
    o[2][4] = (1. - x) * (1. - y) / lz;
    o[2][5] = x * (1. - y) / lz;
    o[2][6] = (1. - x) * y / lz;
    o[2][7] = x * y / lz;
}

/* ThermalModel::Compute_B (matmodel.h:267-270) and SmallStrainMechModel::
 * Compute_B (matmodel.h:284-304): nstr x 8h row major */
static void compute_B(const or_solver *S, double x, double y, double z, double *out)
{
    double b[3][8];
    basic_B(S, x, y, z, b);
    int ncol = 8 * S->h;
    memset(out, 0, sizeof(double) * S->nstr * ncol);
    if (S->h == 1) {
        for (int i = 0; i < 3; i++)
            for (int q = 0; q < 8; q++)
                out[i * ncol + q] = b[i][q];
        return;
    }
    if (S->nstr == 9) {
        /* LargeStrainMechModel::Compute_B (LargeStrainMechModel.h:205-222):
         * row 3i+J = dN/dX_J acting on u_i */
        for (int i = 0; i < 3; i++)
            for (int J = 0; J < 3; J++)
                for (int q = 0; q < 8; q++)
                    out[(3 * i + J) * ncol + 3 * q + i] = b[J][q];
        return;
    }
    const double sh = 7.071067811865476e-01;
    for (int q = 0; q < 8; q++) {
        out[0 * ncol + 3 * q + 0] = b[0][q];
        out[1 * ncol + 3 * q + 1] = b[1][q];
        out[2 * ncol + 3 * q + 2] = b[2][q];
        out[3 * ncol + 3 * q + 0] = sh * b[1][q];
        out[4 * ncol + 3 * q + 0] = sh * b[2][q];
        out[5 * ncol + 3 * q + 1] = sh * b[2][q];
        out[3 * ncol + 3 * q + 1] = sh * b[0][q];
        out[4 * ncol + 3 * q + 2] = sh * b[0][q];
        out[5 * ncol + 3 * q + 2] = sh * b[1][q];
    }
}

/* Matmodel::Construct_B (matmodel.h:104-155) */
static void construct_B(or_solver *S)
{
    int          ncol = 8 * S->h, nstr = S->nstr;
    const double xp = 0.5 + sqrt(3.) / 6., xm = 0.5 - sqrt(3.) / 6.;
    const double xi[8][3] = {{xm, xm, xm}, {xp, xm, xm}, {xm, xp, xm}, {xp, xp, xm},
                             {xm, xm, xp}, {xp, xm, xp}, {xm, xp, xp}, {xp, xp, xp}};
    if (S->cfg.fe_type == OR_FE_HEX8R) {
        compute_B(S, 0.5, 0.5, 0.5, S->Bint[0]);
        return;
    }
    if (S->cfg.fe_type == OR_FE_BBAR && nstr > 3) {
        double Bvol[MAXSTR * MAXE], Bfull[MAXSTR * MAXE];
        compute_B(S, 0.5, 0.5, 0.5, Bvol);
        for (int p = 0; p < 8; p++) {
            compute_B(S, xi[p][0], xi[p][1], xi[p][2], Bfull);
            for (int col = 0; col < ncol; col++) {
                double vf = (Bfull[0 * ncol + col] + Bfull[1 * ncol + col] + Bfull[2 * ncol + col]) / 3.0;
                double vb = (Bvol[0 * ncol + col] + Bvol[1 * ncol + col] + Bvol[2 * ncol + col]) / 3.0;
                for (int row = 0; row < 3; row++)
                    S->Bint[p][row * ncol + col] = Bfull[row * ncol + col] - vf + vb;
                for (int row = 3; row < nstr; row++)
                    S->Bint[p][row * ncol + col] = Bfull[row * ncol + col];
            }
        }
        return;
    }
    for (int p = 0; p < 8; p++)
        compute_B(S, xi[p][0], xi[p][1], xi[p][2], S->Bint[p]);
}

or_solver *or_create(const or_desc *d)
{
    if (d->howmany != 1 && d->howmany != 3)
        return NULL;
    if (d->n_x < 1 || d->n_y < 1 || d->n_z < 2 || (d->n_z % 2) != 0)
        return NULL;
    if (d->large_strain && d->howmany != 3)
        return NULL;
    or_solver *S = (or_solver *) calloc(1, sizeof(or_solver));
    S->cfg       = *d;
    S->nx        = d->n_x;
    S->ny        = d->n_y;
    S->nz        = d->n_z;
    S->h         = d->howmany;
    S->nstr      = d->howmany == 1 ? 3 : (d->large_strain ? 9 : 6);
    S->ngp       = d->fe_type == OR_FE_HEX8R ? 1 : 8;
    S->N         = (ptrdiff_t) S->nx * S->ny * S->nz;
    S->Nf        = (ptrdiff_t) S->nx * S->ny * (S->nz / 2 + 1);
    /* reader.cpp:293-296 */
    S->le[0]     = d->L[0] / (double) S->nx;
    S->le[1]     = d->L[1] / (double) S->ny;
    S->le[2]     = d->L[2] / (double) S->nz;
    S->ve        = S->le[0] * S->le[1] * S->le[2];
    S->tol       = d->tol;
    S->err_type  = d->err_type;
    size_t nh    = (size_t) S->N * S->h;
    S->u         = (double *) calloc(nh, sizeof(double));
    S->u_prev    = (double *) calloc(nh, sizeof(double));
    S->r         = (double *) calloc(nh, sizeof(double));
    S->s         = (double *) calloc(nh, sizeof(double));
    S->d         = (double *) calloc(nh, sizeof(double));
    S->rnew      = (double *) calloc(nh, sizeof(double));
    S->rghost    = (double *) calloc((size_t) S->ny * S->nz * S->h, sizeof(double));
    S->ms        = (uint16_t *) calloc((size_t) S->N, sizeof(uint16_t));
    S->err_all   = (double *) calloc((size_t) d->n_it + 1, sizeof(double));
    construct_B(S);
    for (int p = 0; p < S->ngp; p++) {
        int ne = 8 * S->h;
        for (int i = 0; i < S->nstr; i++) {
            S->nzr[p][i] = 0;
            for (int j = 0; j < ne; j++)
                if (S->Bint[p][i * ne + j] != 0.0)
                    S->nzr_col[p][i][S->nzr[p][i]++] = j;
        }
        for (int j = 0; j < ne; j++) {
            S->nzc[p][j] = 0;
            for (int i = 0; i < S->nstr; i++)
                if (S->Bint[p][i * ne + j] != 0.0)
                    S->nzc_row[p][j][S->nzc[p][j]++] = i;
        }
    }
    return S;
}

void or_destroy(or_solver *S)
{
    if (!S)
        return;
    free(S->u);
    free(S->u_prev);
    free(S->r);
    free(S->s);
    free(S->d);
    free(S->rnew);
    free(S->rghost);
    free(S->ms);
    free(S->err_all);
    free(S->Ke);
    free(S->G);
    free(S->ep);
    free(S->ep_t);
    free(S->q);
    free(S->q_t);
    free(S->pb);
    free(S->pb_t);
    free(S->FE);
    free(S->PF);
    for (int g = 0; g < S->ngroups; g++)
        free(S->groups[g].params);
    free(S->groups);
    free(S->phase_group);
    free(S->phase_local);
    free(S);
}

int or_set_microstructure(or_solver *S, const uint16_t *ms)
{
    memcpy(S->ms, ms, sizeof(uint16_t) * (size_t) S->N);
    return 0;
}

int or_n_gp(or_solver *S) { return S->ngp; }

static const double *phase_params(const or_solver *S, int phase, int *kind)
{
    const or_group *g = &S->groups[S->phase_group[phase]];
    *kind             = g->kind;
    return g->params + (size_t) S->phase_local[phase] * g->np;
}

/* C matrix of a linear phase (LinearThermal.h:25, :89-92; LinearElastic.h:34-35, :118-126) */
static void linear_C(int n, int kind, const double *p, double *C)
{
    memset(C, 0, sizeof(double) * n * n);
    if (kind == OR_MODEL_THERMAL_ISO) {
        for (int i = 0; i < 3; i++)
            C[i * 3 + i] = p[0];
    } else if (kind == OR_MODEL_THERMAL_TRI) {
        double K[9] = {p[0], p[1], p[2], p[1], p[3], p[4], p[2], p[4], p[5]};
        memcpy(C, K, sizeof(K));
    } else if (kind == OR_MODEL_ELASTIC_ISO || kind == OR_MODEL_J2_NEW || kind == OR_MODEL_PSEUDO_PLASTIC ||
               kind == OR_MODEL_PSEUDO_NL) {
        double lam = p[0] - (2.0 / 3.0) * p[1], mu = p[1];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                C[i * 6 + j] = lam;
        for (int i = 0; i < 6; i++)
            C[i * 6 + i] += 2 * mu;
    } else if (kind == OR_MODEL_ELASTIC_TRI) {
        int k = 0;
        for (int row = 0; row < 6; row++)
            for (int col = row; col < 6; col++) {
                C[row * 6 + col] = p[k];
                C[col * 6 + row] = p[k];
                k++;
            }
    }
}

/* sum_p B_p^T C B_p v_e / n_gp (LinearElastic.h:37-39, :135-137; matmodel.h:243-245) */
static void element_stiffness(const or_solver *S, const double *C, double *Ke)
{
    int ne = 8 * S->h, ns = S->nstr;
    memset(Ke, 0, sizeof(double) * ne * ne);
    for (int p = 0; p < S->ngp; p++) {
        const double *B = S->Bint[p];
        double        CB[MAXSTR * MAXE];
        for (int i = 0; i < ns; i++)
            for (int j = 0; j < ne; j++) {
                double a = 0;
                for (int k = 0; k < ns; k++)
                    a += C[i * ns + k] * B[k * ne + j];
                CB[i * ne + j] = a;
            }
        for (int i = 0; i < ne; i++)
            for (int j = 0; j < ne; j++) {
                double a = 0;
                for (int k = 0; k < ns; k++)
                    a += B[k * ne + i] * CB[k * ne + j];
                Ke[i * ne + j] += a * S->ve / S->ngp;
            }
    }
}

/* model->get_reference_stiffness() averaged over groups (MaterialManager.h:198-204) */
static void large_reference_stiffness(const or_group *g, double *C);

static void group_reference_stiffness(const or_solver *S, const or_group *g, double *C)
{
    int n = S->nstr;
    memset(C, 0, sizeof(double) * n * n);
    if (g->kind == OR_MODEL_SVK || g->kind == OR_MODEL_NEO_HOOKE) {
        large_reference_stiffness(g, C);
        return;
    }
    if (g->kind == OR_MODEL_THERMAL_ISO || g->kind == OR_MODEL_THERMAL_TRI || g->kind == OR_MODEL_ELASTIC_TRI) {
        /* LinearThermal.h:42-45, :109-112 (kappa_average), LinearElastic.h:146-154 */
        double Ci[81];
        for (int i = 0; i < g->nphase; i++) {
            linear_C(S->nstr, g->kind, g->params + (size_t) i * g->np, Ci);
            for (int k = 0; k < n * n; k++)
                C[k] += Ci[k];
        }
        for (int k = 0; k < n * n; k++)
            C[k] /= g->nphase;
    } else if (g->kind == OR_MODEL_ELASTIC_ISO) {
        /* LinearElastic.h:55-69: (max+min)/2 of lambda and mu */
        double lmax = -INFINITY, lmin = INFINITY, mmax = -INFINITY, mmin = INFINITY;
        for (int i = 0; i < g->nphase; i++) {
            const double *p   = g->params + (size_t) i * g->np;
            double        lam = p[0] - (2.0 / 3.0) * p[1];
            lmax              = fmax(lmax, lam);
            lmin              = fmin(lmin, lam);
            mmax              = fmax(mmax, p[1]);
            mmin              = fmin(mmin, p[1]);
        }
        double lr = (lmax + lmin) / 2, mr = (mmax + mmin) / 2;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                C[i * 6 + j] = lr;
        for (int i = 0; i < 6; i++)
            C[i * 6 + i] += 2 * mr;
    } else {
        /* J2PlasticityNew.h:111-121, PseudoPlastic.h:43-54: mean K, mean G */
        double Kb = 0, Gb = 0;
        for (int i = 0; i < g->nphase; i++) {
            Kb += g->params[(size_t) i * g->np + 0];
            Gb += g->params[(size_t) i * g->np + 1];
        }
        Kb /= g->nphase;
        Gb /= g->nphase;
        double lb = Kb - 2.0 * Gb / 3.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                C[i * 6 + j] = lb;
        for (int i = 0; i < 6; i++)
            C[i * 6 + i] += 2.0 * Gb;
    }
}

/* Solver::computeFundamentalSolution (solver.h:144-204) */
static void build_green(or_solver *S)
{
    int          h = S->h, nx = S->nx, ny = S->ny, nz = S->nz, nc = nz / 2 + 1, np = h * (h + 1) / 2;
    const double tpi = 2 * acos(-1.0);
    const double *K0 = S->keref;
    int           ne = 8 * h;
    free(S->G);
    S->G = (double *) calloc((size_t) S->Nf * np, sizeof(double));
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S->Nf; f++) {
        int ix = (int) (f / ((ptrdiff_t) ny * nc));
        int iy = (int) ((f / nc) % ny);
        int iz = (int) (f % nc);
        if (ix == 0 && iy == 0 && iz == 0)
            continue;
        double thx = tpi * ix / (double) nx, thy = tpi * iy / (double) ny, thz = tpi * iz / (double) nz;
        cplx   ex = cos(thx) + I * sin(thx), ey = cos(thy) + I * sin(thy), ez = cos(thz) + I * sin(thz);
        cplx   A[8];
        A[0] = 1.0;
        A[1] = ex;
        A[2] = ey;
        A[3] = ex * ey;
        A[4] = ez;
        A[5] = ex * ez;
        A[6] = ez * ey;
        A[7] = ex * ey * ez;
        double AA[8][8];
        for (int a = 0; a < 8; a++)
            for (int b = 0; b < 8; b++)
                AA[a][b] = creal(A[a]) * creal(A[b]) + cimag(A[a]) * cimag(A[b]);
        double blk[9], P[9];
        for (int i = 0; i < h; i++)
            for (int j = i; j < h; j++) {
                double acc = 0;
                for (int a = 0; a < 8; a++)
                    for (int b = 0; b < 8; b++)
                        acc += K0[(8 * i + a) * ne + 8 * j + b] * AA[a][b];
                blk[i * h + j] = acc;
                blk[j * h + i] = acc;
            }
        sym_pinv(h, blk, 1e-14, P);
        double *g = S->G + f * np;
        int     k = 0;
        for (int i = 0; i < h; i++)
            for (int j = i; j < h; j++)
                g[k++] = P[i * h + j] / (double) ((ptrdiff_t) nx * ny * nz);
    }
}

int or_set_materials(or_solver *S, int n_groups, const int32_t *kind, const int32_t *nphase, const int32_t *phase_ids,
                     const double *params, const double *ref_material)
{
    int maxp = -1, tot = 0;
    for (int g = 0; g < n_groups; g++) {
        if (or_model_nparams(kind[g]) < 0)
            return -1;
        int thermal = kind[g] == OR_MODEL_THERMAL_ISO || kind[g] == OR_MODEL_THERMAL_TRI;
        if (thermal != (S->h == 1))
            return -2;
        /* createMatmodel<3,9> accepts only the finite strain models (setup.h:62-73) */
        int large = kind[g] == OR_MODEL_SVK || kind[g] == OR_MODEL_NEO_HOOKE;
        if (large != (S->nstr == 9))
            return -2;
        for (int i = 0; i < nphase[g]; i++)
            maxp = phase_ids[tot + i] > maxp ? phase_ids[tot + i] : maxp;
        tot += nphase[g];
    }
    if (maxp < 0)
        return -3;
    S->ngroups     = n_groups;
    S->groups      = (or_group *) calloc((size_t) n_groups, sizeof(or_group));
    S->nphases     = maxp + 1;
    S->phase_group = (int *) malloc(sizeof(int) * (size_t) S->nphases);
    S->phase_local = (int *) malloc(sizeof(int) * (size_t) S->nphases);
    for (int p = 0; p < S->nphases; p++)
        S->phase_group[p] = -1;
    S->all_linear = 1;
    int ofs = 0, pofs = 0;
    for (int g = 0; g < n_groups; g++) {
        or_group *G = &S->groups[g];
        G->kind     = kind[g];
        G->nphase   = nphase[g];
        G->np       = or_model_nparams(kind[g]);
        G->params   = (double *) malloc(sizeof(double) * (size_t) G->nphase * G->np);
        memcpy(G->params, params + pofs, sizeof(double) * (size_t) G->nphase * G->np);
        pofs += G->nphase * G->np;
        if (!model_is_linear(G->kind))
            S->all_linear = 0;
        for (int i = 0; i < nphase[g]; i++) {
            int p = phase_ids[ofs + i];
            if (p < 0 || S->phase_group[p] >= 0)
                return -4; /* MaterialManager.h:99-100 */
            S->phase_group[p] = g;
            S->phase_local[p] = i;
        }
        ofs += nphase[g];
    }
    for (int p = 0; p < S->nphases; p++)
        if (S->phase_group[p] < 0)
            return -5; /* MaterialManager.h:106-108 */
    /* reference stiffness (MaterialManager.h:177-205) */
    int n = S->nstr;
    if (ref_material) {
        memcpy(S->kref, ref_material, sizeof(double) * n * n);
    } else {
        memset(S->kref, 0, sizeof(S->kref));
        double C[81];
        for (int g = 0; g < n_groups; g++) {
            group_reference_stiffness(S, &S->groups[g], C);
            for (int k = 0; k < n * n; k++)
                S->kref[k] += C[k];
        }
        for (int k = 0; k < n * n; k++)
            S->kref[k] /= (double) n_groups;
    }
    /* per-phase element stiffness for linear phases (LinearModel::phase_stiffness) */
    int ne = 8 * S->h;
    S->Ke  = (double *) calloc((size_t) S->nphases * ne * ne, sizeof(double));
    for (int p = 0; p < S->nphases; p++) {
        int           k;
        const double *pp = phase_params(S, p, &k);
        if (model_is_linear(k)) {
            double C[81];
            linear_C(S->nstr, k, pp, C);
            element_stiffness(S, C, S->Ke + (size_t) p * ne * ne);
        }
    }
    /* Compute_Reference_ElementStiffness (matmodel.h:237-253): component-major */
    double tmp[MAXE * MAXE];
    element_stiffness(S, S->kref, tmp);
    for (int i = 0; i < ne; i++)
        for (int j = 0; j < ne; j++)
            S->keref[((i % S->h) * 8 + i / S->h) * ne + (j % S->h) * 8 + j / S->h] = tmp[i * ne + j];
    /* internal variables (Solver ctor solver.h:139) */
    S->has_state = S->has_visco = 0;
    for (int g = 0; g < n_groups; g++) {
        if (S->groups[g].kind == OR_MODEL_J2_NEW)
            S->has_state = 1;
        if (S->groups[g].kind == OR_MODEL_J2_VISCO_LIN || S->groups[g].kind == OR_MODEL_J2_VISCO_NL)
            S->has_state = S->has_visco = 1;
    }
    if (S->has_state) {
        /* initializeInternalVariables (J2PlasticityNew.h:29-35, J2Plasticity.h:46-54): zeros */
        size_t ng = (size_t) S->N * S->ngp;
        S->ep     = (double *) calloc(ng * 6, sizeof(double));
        S->ep_t   = (double *) calloc(ng * 6, sizeof(double));
        S->q      = (double *) calloc(ng, sizeof(double));
        S->q_t    = (double *) calloc(ng, sizeof(double));
        if (S->has_visco) {
            S->pb   = (double *) calloc(ng * 6, sizeof(double));
            S->pb_t = (double *) calloc(ng * 6, sizeof(double));
        }
    }
    build_green(S);
    return 0;
}

void or_get_green(or_solver *S, double *packed)
{
    memcpy(packed, S->G, sizeof(double) * (size_t) S->Nf * S->h * (S->h + 1) / 2);
}
void or_get_reference_stiffness(or_solver *S, double *k) { memcpy(k, S->kref, sizeof(double) * S->nstr * S->nstr); }
void or_get_ref_element_stiffness(or_solver *S, double *ke)
{
    memcpy(ke, S->keref, sizeof(double) * 64 * S->h * S->h);
}
void or_get_B(or_solver *S, double *B)
{
    int sz = S->nstr * 8 * S->h;
    for (int p = 0; p < S->ngp; p++)
        memcpy(B + (size_t) p * sz, S->Bint[p], sizeof(double) * sz);
}
int or_get_phase_stiffness(or_solver *S, int phase, double *ke)
{
    int ne = 8 * S->h;
    if (phase < 0 || phase >= S->nphases)
        return -1;
    memcpy(ke, S->Ke + (size_t) phase * ne * ne, sizeof(double) * ne * ne);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* constitutive updates (get_sigma)                                          */
/* ------------------------------------------------------------------------ */

/* J2PlasticityNew::get_sigma (J2PlasticityNew.h:43-109) */
static void j2new_point(const double *p, const double *e, const double *ep_in, double q_in, double *s, double *ep,
                        double *qo)
{
    const double c23 = sqrt(2.0 / 3.0);
    double       G   = p[1];
    for (int i = 0; i < 6; i++)
        s[i] = 2.0 * G * (e[i] - ep_in[i]);
    double lam = p[0] - 2.0 / 3.0 * G;
    double tr  = e[0] + e[1] + e[2];
    for (int i = 0; i < 3; i++)
        s[i] += lam * tr;
    double st[6];
    double mean = (s[0] + s[1] + s[2]) / 3.0;
    for (int i = 0; i < 6; i++)
        st[i] = s[i] - (i < 3 ? mean : 0.0);
    double s_t = 0;
    for (int i = 0; i < 6; i++)
        s_t += st[i] * st[i];
    s_t           = sqrt(s_t);
    double sy     = p[2] + q_in * p[3];
    double dsy    = p[3];
    double phi    = s_t - c23 * sy;
    double dgamma = fmax(0.0, phi / (2.0 * G + 2.0 / 3.0 * dsy));
    for (int i = 0; i < 6; i++)
        ep[i] = s_t > 1e-12 ? ep_in[i] + dgamma * st[i] / s_t : ep_in[i];
    *qo = q_in + c23 * dgamma;
    if (s_t > 1e-12)
        for (int i = 0; i < 6; i++)
            s[i] = s[i] - dgamma * 2.0 * G * st[i] / s_t;
}

/* J2Plasticity::get_sigma (J2Plasticity.h:65-108) with the hardening laws of
 * J2ViscoPlastic_LinearIsotropicHardening (:162-178) and
 * J2ViscoPlastic_NonLinearIsotropicHardening (:182-243).  p = [K G sigma_y
 * K_iso H_kin eta dt (sigma_inf delta)].  The trial state reads the committed
 * plastic strain ep_t, psi_t and psi_bar_t; ep is overwritten, while psi and
 * psi_bar (the current variables) accumulate gamma at every call, as the
 * reference does (:102-104).  The Newton iteration of the nonlinear law uses
 * the reference's derivative, whose (2 / 3) is integer division, i.e. 0 (:217). */
static void visco_point(int kind, const double *p, const double *e, const double *ep_t, double psi_t,
                        const double *pb_t, double *s, double *ep, double *psi, double *pb)
{
    const double c23 = sqrt(2.0 / 3.0);
    const double K = p[0], G = p[1], sy = p[2], Kh = p[3], Hk = p[4], eta = p[5], dt = p[6];
    double       ee[6], dev[6], dmq[6], n[6];
    for (int i = 0; i < 6; i++)
        ee[i] = e[i] - ep_t[i];
    const double treps = ee[0] + ee[1] + ee[2];
    for (int i = 0; i < 3; i++)
        s[i] = (K - 2.0 * G / 3.0) * treps;
    for (int i = 0; i < 3; i++)
        s[i] += 2 * G * ee[i];
    for (int i = 3; i < 6; i++)
        s[i] = 2 * G * ee[i];
    const double mean = (s[0] + s[1] + s[2]) / 3.0;
    for (int i = 0; i < 6; i++)
        dev[i] = s[i] - (i < 3 ? mean : 0.0);
    double q_tr;
    if (kind == OR_MODEL_J2_VISCO_LIN)
        q_tr = -Kh * psi_t;
    else
        q_tr = -Kh * psi_t - (p[7] - sy) * (1 - exp(-p[8] * psi_t));
    double nrm2 = 0;
    for (int i = 0; i < 6; i++) {
        dmq[i] = dev[i] - (-(2.0 / 3.0) * Hk * pb_t[i]);
        nrm2 += dmq[i] * dmq[i];
    }
    const double nrm = sqrt(nrm2);
    for (int i = 0; i < 6; i++)
        n[i] = nrm < 1e-12 ? 0.0 : dmq[i] / nrm;
    const double f   = nrm - c23 * (sy - q_tr);
    const double den = 2 * G + (2.0 / 3.0) * (Kh + Hk) + eta / dt;
    double       gamma = 0.0;
    if (f >= 0) {
        if (kind == OR_MODEL_J2_VISCO_LIN) {
            gamma = f / den;
        } else {
            const double sd = c23 * (p[7] - sy), dl = p[8];
            double       inc = 1;
            int          it  = 0;
            while (inc > 1e-10 && it < 10) {
                double g  = f - gamma * den - sd * (-exp(-dl * (psi_t + c23 * gamma)) + exp(-dl * psi_t));
                double dg = -den - (2 / 3) * (p[7] - sy) * dl * exp(-dl * (psi_t + c23 * gamma));
                inc       = -g / dg;
                gamma += inc;
                it++;
            }
        }
    }
    for (int i = 0; i < 6; i++) {
        s[i] -= gamma * 2 * G * n[i];
        ep[i] = ep_t[i] + gamma * n[i];
    }
    *psi += gamma * c23;
    for (int i = 0; i < 6; i++)
        pb[i] -= gamma * n[i];
}

void or_visco_point(int kind, const double *params, const double *eps, const double *ep_t, double psi_t,
                    const double *pb_t, double *sigma, double *ep, double *psi, double *pb)
{
    visco_point(kind, params, eps, ep_t, psi_t, pb_t, sigma, ep, psi, pb);
}

/* PseudoPlasticLinearHardening::get_sigma (PseudoPlastic.h:92-116) */
static void pseudo_point(const double *p, const double *e, double *s)
{
    const double a = 2. / 3, b = sqrt(a);
    double       K = p[0], G = p[1], sy = p[2], hp = p[3];
    double       eps_crit = sqrt(2. / 3.) * sy / (2. * G);
    double       E_s      = (3. * G) / (3. * G + hp);
    double       tr       = e[0] + e[1] + e[2];
    double       dev[6];
    for (int i = 0; i < 3; i++)
        dev[i] = e[i] - (1.0 / 3.0) * tr;
    for (int i = 3; i < 6; i++)
        dev[i] = e[i];
    double nd = 0;
    for (int i = 0; i < 6; i++)
        nd += dev[i] * dev[i];
    nd          = sqrt(nd);
    double buf1 = K * tr, buf2;
    if (nd <= eps_crit)
        buf2 = 2.0 * G;
    else
        buf2 = (b * sy + a * E_s * hp * (nd - eps_crit)) / nd;
    for (int i = 0; i < 3; i++)
        s[i] = buf1 + buf2 * dev[i];
    for (int i = 3; i < 6; i++)
        s[i] = buf2 * dev[i];
}

/* ------------------------------------------------------------------------ */
/* finite strain (LargeStrainMechModel.h, SaintVenantKirchhoff.h,             */
/* CompressibleNeoHookean.h): eps holds F row major, sigma holds P = F S      */
/* ------------------------------------------------------------------------ */
static int mandel_index(int a, int b)
{
    if (a == b)
        return a;
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 3 : 4) : 5;
}

static void inv3(const double A[3][3], double Ai[3][3], double *det)
{
    double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    double d   = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    Ai[0][0]   = c00 / d;
    Ai[1][0]   = c01 / d;
    Ai[2][0]   = c02 / d;
    Ai[0][1]   = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / d;
    Ai[1][1]   = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / d;
    Ai[2][1]   = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / d;
    Ai[0][2]   = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / d;
    Ai[1][2]   = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / d;
    Ai[2][2]   = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / d;
    *det       = d;
}

/* second Piola-Kirchhoff stress S(F) of the two hyperelastic models; returns
 * nonzero for det F <= 0 (CompressibleNeoHookean.h:40-42 throws there) */
static int large_S(int kind, const double *p, const double F[3][3], double S[3][3])
{
    const double mu = p[1], lam = p[0] - (2.0 / 3.0) * p[1];
    double       C[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            C[a][b] = F[0][a] * F[0][b] + F[1][a] * F[1][b] + F[2][a] * F[2][b];
    if (kind == OR_MODEL_SVK) {
        /* S = lam tr(E) I + 2 mu E, E = (C - I) / 2 (SaintVenantKirchhoff.h:34-39) */
        double trE = 0.5 * (C[0][0] + C[1][1] + C[2][2] - 3.0);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                S[a][b] = mu * (C[a][b] - (a == b)) + (a == b ? lam * trE : 0.0);
        return 0;
    }
    /* S = lam log(J) C^-1 + mu (I - C^-1) (CompressibleNeoHookean.h:35-49) */
    double Fa[3][3], Fi[3][3], J, Ci[3][3], dC;
    memcpy(Fa, F, sizeof(Fa));
    inv3(Fa, Fi, &J);
    if (!(J > 0.0))
        return -1;
    inv3(C, Ci, &dC);
    const double lj = log(J);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            S[a][b] = lam * lj * Ci[a][b] + mu * ((a == b) - Ci[a][b]);
    return 0;
}

/* LargeStrainMechModel::get_sigma (LargeStrainMechModel.h:193-199) */
static void large_point(int kind, const double *p, const double *e, double *P)
{
    double F[3][3], S[3][3];
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            F[a][b] = e[3 * a + b];
    if (large_S(kind, p, F, S)) {
        for (int k = 0; k < 9; k++)
            P[k] = NAN;
        return;
    }
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            P[3 * a + b] = F[a][0] * S[0][b] + F[a][1] * S[1][b] + F[a][2] * S[2][b];
}

/* material tangent dS/dE (Mandel, 6x6) of the two models at F
 * (SaintVenantKirchhoff.h:41-47, CompressibleNeoHookean.h:51-87) */
static void large_material_tangent(int kind, const double *p, const double F[3][3], double *Cm)
{
    const double mu = p[1], lam = p[0] - (2.0 / 3.0) * p[1];
    memset(Cm, 0, sizeof(double) * 36);
    if (kind == OR_MODEL_SVK) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                Cm[a * 6 + b] = lam;
        for (int a = 0; a < 6; a++)
            Cm[a * 6 + a] += 2.0 * mu;
        return;
    }
    double C[3][3], Ci[3][3], Fa[3][3], Fi[3][3], J, dC;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            C[a][b] = F[0][a] * F[0][b] + F[1][a] * F[1][b] + F[2][a] * F[2][b];
    memcpy(Fa, F, sizeof(Fa));
    inv3(Fa, Fi, &J);
    inv3(C, Ci, &dC);
    static const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    const double     r2       = sqrt(2.0);
    double           cv[6];
    for (int a = 0; a < 6; a++)
        cv[a] = Ci[ij[a][0]][ij[a][1]] * (a < 3 ? 1.0 : r2);
    const double c2 = mu - lam * log(J);
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) {
            int          ia = ij[a][0], ja = ij[a][1], kb = ij[b][0], lb = ij[b][1];
            const double pp2 =
                (Ci[ia][kb] * Ci[ja][lb] + Ci[ia][lb] * Ci[ja][kb]) * (a < 3 ? 1.0 : r2) * (b < 3 ? 1.0 : r2);
            Cm[a * 6 + b] = lam * cv[a] * cv[b] + c2 * pp2;
        }
}

/* dP/dF (9x9, rows iJ, columns kL) from S and dS/dE
 * (LargeStrainMechModel.h:101-176).  Like the reference, the sum over the
 * symmetric pair (P,Q) runs over P <= Q only, so the columns of the off-diagonal
 * F_kL carry half of the tensor's dE_PQ/dF_kL contribution. */
static void large_spatial_tangent(const double F[3][3], const double S[3][3], const double *Cm, double *A)
{
    memset(A, 0, sizeof(double) * 81);
    const double r2 = sqrt(2.0);
    for (int i = 0; i < 3; i++)
        for (int J = 0; J < 3; J++)
            for (int k = 0; k < 3; k++)
                for (int L = 0; L < 3; L++) {
                    double a = (i == k) ? S[L][J] : 0.0;
                    for (int M = 0; M < 3; M++) {
                        int mj = mandel_index(M, J);
                        for (int P = 0; P < 3; P++)
                            for (int Q = P; Q < 3; Q++) {
                                int    pq = mandel_index(P, Q);
                                double c  = Cm[mj * 6 + pq] / ((mj >= 3 ? r2 : 1.0) * (pq >= 3 ? r2 : 1.0));
                                double de = (Q == L ? 0.5 * F[k][P] : 0.0) + (P == L ? 0.5 * F[k][Q] : 0.0);
                                a += F[i][M] * c * de;
                            }
                    }
                    A[(3 * i + J) * 9 + 3 * k + L] = a;
                }
}

/* get_reference_stiffness of the two models: dP/dF at F = I averaged over the
 * group's phases (SaintVenantKirchhoff.h:49-66, CompressibleNeoHookean.h:89-104) */
static void large_reference_stiffness(const or_group *g, double *C)
{
    const double Fi[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    memset(C, 0, sizeof(double) * 81);
    for (int ph = 0; ph < g->nphase; ph++) {
        const double *p = g->params + (size_t) ph * g->np;
        double        S[3][3], Cm[36], A[81];
        large_S(g->kind, p, Fi, S);
        large_material_tangent(g->kind, p, Fi, Cm);
        large_spatial_tangent(Fi, S, Cm, A);
        for (int k = 0; k < 81; k++)
            C[k] += A[k];
    }
    for (int k = 0; k < 81; k++)
        C[k] /= g->nphase;
}

/* PseudoPlasticNonLinearHardening (PseudoPlastic.h:125-168): deformation
 * plasticity with the power law sigma_eq = sigma_y (eps_eq / eps_0)^n beyond
 * eps_crit = eps_0 (sigma_y / (3 G eps_0))^(1 / (1 - n)) (:137-140), where the
 * elastic branch 3 G eps_eq meets it; eps_eq = sqrt(2/3) |dev eps| (:149) */
static void pseudo_nl_point(const double *p, const double *e, double *s)
{
    const double K = p[0], G = p[1], sy = p[2], n = p[3], e0 = p[4];
    const double eps_crit = e0 * pow(sy / (3.0 * G * e0), 1.0 / (1.0 - n));
    const double tr       = e[0] + e[1] + e[2];
    double       dev[6], nd = 0;
    for (int i = 0; i < 3; i++)
        dev[i] = e[i] - (1.0 / 3.0) * tr;
    for (int i = 3; i < 6; i++)
        dev[i] = e[i];
    for (int i = 0; i < 6; i++)
        nd += dev[i] * dev[i];
    nd                = sqrt(nd);
    const double eq   = sqrt(2.0 / 3.0) * nd;
    const double buf1 = K * tr;
    const double buf2 = eq <= eps_crit ? 2.0 * G : sqrt(2.0 / 3.0) * sy * pow(eq / e0, n) / nd;
    for (int i = 0; i < 3; i++)
        s[i] = buf1 + buf2 * dev[i];
    for (int i = 3; i < 6; i++)
        s[i] = buf2 * dev[i];
}

void or_material_point(int kind, const double *p, const double *e, const double *ep_t, double q_t, double *sigma,
                       double *ep, double *q)
{
    if (kind == OR_MODEL_SVK || kind == OR_MODEL_NEO_HOOKE) {
        large_point(kind, p, e, sigma);
        return;
    }
    if (kind == OR_MODEL_J2_NEW) {
        j2new_point(p, e, ep_t, q_t, sigma, ep, q);
    } else if (kind == OR_MODEL_PSEUDO_PLASTIC) {
        pseudo_point(p, e, sigma);
    } else if (kind == OR_MODEL_PSEUDO_NL) {
        pseudo_nl_point(p, e, sigma);
    } else if (kind == OR_MODEL_ELASTIC_ISO) {
        /* LinearElastic.h:43-53 */
        double lam = p[0] - (2.0 / 3.0) * p[1];
        double b1 = lam * (e[0] + e[1] + e[2]), b2 = 2 * p[1];
        for (int i = 0; i < 3; i++)
            sigma[i] = b1 + b2 * e[i];
        for (int i = 3; i < 6; i++)
            sigma[i] = b2 * e[i];
    } else if (kind == OR_MODEL_THERMAL_ISO) {
        for (int i = 0; i < 3; i++)
            sigma[i] = p[0] * e[i];
    } else {
        int    n = kind == OR_MODEL_THERMAL_TRI ? 3 : 6;
        double C[36];
        linear_C(n, kind, p, C);
        for (int i = 0; i < n; i++) {
            double a = 0;
            for (int j = 0; j < n; j++)
                a += C[i * n + j] * e[j];
            sigma[i] = a;
        }
    }
}

/* eps = B ue + g0; sigma = get_sigma per GP (matmodel.h:190-201, 202-209) */
static void element_strain_stress(or_solver *S, const double *ue, int phase, ptrdiff_t el, double *eps, double *sig)
{
    int ne = 8 * S->h, ns = S->nstr;
    for (int p = 0; p < S->ngp; p++)
        for (int i = 0; i < ns; i++) {
            double a = 0;
            for (int k = 0; k < S->nzr[p][i]; k++) {
                int j = S->nzr_col[p][i][k];
                a += S->Bint[p][i * ne + j] * ue[j];
            }
            eps[p * ns + i] = a + S->g0[i];
        }
    int           kind;
    const double *pp = phase_params(S, phase, &kind);
    for (int p = 0; p < S->ngp; p++) {
        double *e = eps + p * ns, *s = sig + p * ns;
        if (kind == OR_MODEL_J2_NEW) {
            size_t gi = (size_t) el * S->ngp + p;
            j2new_point(pp, e, S->ep_t + gi * 6, S->q_t[gi], s, S->ep + gi * 6, S->q + gi);
        } else if (kind == OR_MODEL_J2_VISCO_LIN || kind == OR_MODEL_J2_VISCO_NL) {
            size_t gi = (size_t) el * S->ngp + p;
            visco_point(kind, pp, e, S->ep_t + gi * 6, S->q_t[gi], S->pb_t + gi * 6, s, S->ep + gi * 6, S->q + gi,
                        S->pb + gi * 6);
        } else {
            or_material_point(kind, pp, e, NULL, 0.0, s, NULL, NULL);
        }
    }
}

static void element_residual(or_solver *S, const double *ue, int phase, ptrdiff_t el, double *res)
{
    double eps[8 * MAXSTR], sig[8 * MAXSTR];
    int    ne = 8 * S->h, ns = S->nstr;
    element_strain_stress(S, ue, phase, el, eps, sig);
    for (int j = 0; j < ne; j++) {
        double a = 0;
        for (int p = 0; p < S->ngp; p++)
            for (int k = 0; k < S->nzc[p][j]; k++) {
                int i = S->nzc_row[p][j][k];
                a += S->Bint[p][i * ne + j] * sig[p * ns + i];
            }
        res[j] = a * S->ve / S->ngp;
    }
}

/* node offsets of element (ix,iy,iz) in the order of solver.h:318-382; x+1 of the
 * last plane lands in the ghost plane (index N..N+ny*nz-1) */
static inline void element_nodes(const or_solver *S, int ix, int iy, int iz, ptrdiff_t *idx)
{
    int       y1 = iy + 1 == S->ny ? 0 : iy + 1;
    int       z1 = iz + 1 == S->nz ? 0 : iz + 1;
    ptrdiff_t nz = S->nz, nyz = (ptrdiff_t) S->ny * S->nz;
    ptrdiff_t x0 = ix * nyz, x1 = (ptrdiff_t) (ix + 1) * nyz;
    idx[0]       = x0 + iy * nz + iz;
    idx[1]       = x1 + iy * nz + iz;
    idx[2]       = x0 + y1 * nz + iz;
    idx[3]       = x1 + y1 * nz + iz;
    idx[4]       = x0 + iy * nz + z1;
    idx[5]       = x1 + iy * nz + z1;
    idx[6]       = x0 + y1 * nz + z1;
    idx[7]       = x1 + y1 * nz + z1;
}

/* Solver::compute_residual_basic (solver.h:228-270) with f = element_residual
 * (compute_residual, :272-280) or f = Ke[phase] ue (solverCG.h:97-103).  x-planes
 * are processed in two colours so a threaded run stays race free and gives the
 * same bits for any thread count. */
void or_residual(or_solver *S, const double *u, double *r, int linear_kd)
{
    int       h = S->h, ne = 8 * h;
    ptrdiff_t N = S->N, nyz = (ptrdiff_t) S->ny * S->nz;
    memset(r, 0, sizeof(double) * (size_t) N * h);
    memset(S->rghost, 0, sizeof(double) * (size_t) nyz * h);
    for (int color = 0; color < 2; color++) {
#pragma omp parallel for schedule(static)
        for (int ix = color; ix < S->nx; ix += 2) {
            ptrdiff_t idx[8];
            double    ue[MAXE], res[MAXE];
            for (int iy = 0; iy < S->ny; iy++)
                for (int iz = 0; iz < S->nz; iz++) {
                    element_nodes(S, ix, iy, iz, idx);
                    for (int a = 0; a < 8; a++) {
                        ptrdiff_t n = idx[a] >= N ? idx[a] - N : idx[a]; /* ghost plane of u = plane 0 */
                        for (int c = 0; c < h; c++)
                            ue[h * a + c] = u[h * n + c] - u[h * idx[0] + c];
                    }
                    int phase = S->ms[idx[0]];
                    if (linear_kd) {
                        const double *K = S->Ke + (size_t) phase * ne * ne;
                        for (int i = 0; i < ne; i++) {
                            double a = 0;
                            for (int j = 0; j < ne; j++)
                                a += K[i * ne + j] * ue[j];
                            res[i] = a;
                        }
                    } else {
                        element_residual(S, ue, phase, idx[0], res);
                    }
                    for (int a = 0; a < 8; a++) {
                        double *dst = idx[a] >= N ? S->rghost + (idx[a] - N) * h : r + idx[a] * h;
                        for (int c = 0; c < h; c++)
                            dst[c] += res[h * a + c];
                    }
                }
        }
    }
    for (ptrdiff_t i = 0; i < nyz * h; i++)
        r[i] += S->rghost[i];
}

/* Solver::convolution (solver.h:387-412): r2c, per-frequency G r_hat, c2r */
void or_convolve(or_solver *S, const double *r, double *out)
{
    int   h = S->h, np = h * (h + 1) / 2;
    cplx *X = (cplx *) malloc(sizeof(cplx) * (size_t) S->Nf * h);
    or_fft_r2c(S->nx, S->ny, S->nz, h, r, (double *) X);
#pragma omp parallel for schedule(static)
    for (ptrdiff_t f = 0; f < S->Nf; f++) {
        const double *g = S->G + f * np;
        cplx         *x = X + f * h;
        if (h == 1) {
            x[0] = g[0] * x[0];
        } else {
            cplx a = x[0], b = x[1], c = x[2];
            x[0]   = g[0] * a + g[1] * b + g[2] * c;
            x[1]   = g[1] * a + g[3] * b + g[4] * c;
            x[2]   = g[2] * a + g[4] * b + g[5] * c;
        }
    }
    or_fft_c2r(S->nx, S->ny, S->nz, h, (double *) X, out);
    free(X);
}

/* deterministic dot: per x-plane sequential partials, summed in plane order */
static double dotp(const or_solver *S, const double *a, const double *b)
{
    ptrdiff_t pl = (ptrdiff_t) S->ny * S->nz * S->h;
    double   *part = (double *) malloc(sizeof(double) * (size_t) S->nx);
#pragma omp parallel for schedule(static)
    for (int x = 0; x < S->nx; x++) {
        double acc = 0;
        for (ptrdiff_t i = x * pl; i < (x + 1) * pl; i++)
            acc += a[i] * b[i];
        part[x] = acc;
    }
    double t = 0;
    for (int x = 0; x < S->nx; x++)
        t += part[x];
    free(part);
    return t;
}

/* Solver::compute_error (solver.h:414-452) */
static double compute_error(or_solver *S, const double *r)
{
    ptrdiff_t n = S->N * S->h;
    double    err;
    if (S->cfg.measure == OR_MEASURE_L1) {
        err = 0;
        for (ptrdiff_t i = 0; i < n; i++)
            err += fabs(r[i]);
    } else if (S->cfg.measure == OR_MEASURE_L2) {
        err = sqrt(dotp(S, r, r));
    } else {
        err = 0;
        for (ptrdiff_t i = 0; i < n; i++)
            err = fmax(err, fabs(r[i]));
    }
    S->err_all[S->iter] = err;
    double err0         = S->err_all[0];
    double err_rel      = S->iter == 0 ? 100 : err / err0;
    return S->err_type == OR_ERR_ABSOLUTE ? err : err_rel;
}

void or_set_gradient(or_solver *S, const double *g0)
{
    memcpy(S->g0, g0, sizeof(double) * S->nstr);
}

void or_get_macro_gradient(or_solver *S, double *g0) { memcpy(g0, S->g0, sizeof(double) * S->nstr); }

/* Solver::get_homogenized_stress (solver.h:707-737) */
void or_homogenized_stress(or_solver *S, double *out)
{
    int       h = S->h, ns = S->nstr;
    ptrdiff_t N = S->N;
    double    acc[MAXSTR] = {0};
    /* x planes in parallel, summed in plane order afterwards */
    double   *plane = (double *) calloc((size_t) S->nx * ns, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < S->nx; ix++)
        for (int iy = 0; iy < S->ny; iy++)
            for (int iz = 0; iz < S->nz; iz++) {
                ptrdiff_t idx[8];
                double    ue[MAXE], eps[8 * MAXSTR], sig[8 * MAXSTR];
                element_nodes(S, ix, iy, iz, idx);
                for (int a = 0; a < 8; a++) {
                    ptrdiff_t n = idx[a] >= N ? idx[a] - N : idx[a];
                    for (int c = 0; c < h; c++)
                        ue[h * a + c] = S->u[h * n + c];
                }
                element_strain_stress(S, ue, S->ms[idx[0]], idx[0], eps, sig);
                for (int j = 0; j < ns; j++) {
                    double av = 0;
                    for (int p = 0; p < S->ngp; p++)
                        av += sig[p * ns + j] / S->ngp;
                    plane[(size_t) ix * ns + j] += av;
                }
            }
    for (int ix = 0; ix < S->nx; ix++)
        for (int j = 0; j < ns; j++)
            acc[j] += plane[(size_t) ix * ns + j];
    free(plane);
    for (int j = 0; j < ns; j++)
        out[j] = acc[j] / (double) N;
}

/* MixedBCController::update (mixedBCs.h:161-178) */
static void mixed_update(or_solver *S)
{
    if (!S->mixed_active)
        return;
    double Pbar[MAXSTR];
    or_homogenized_stress(S, Pbar);
    if (S->nF) {
        double rhs[9], dE[9];
        for (int i = 0; i < S->nF; i++)
            rhs[i] = S->PF[(size_t) S->step * S->nF + i] - Pbar[S->idxF[i]];
        for (int i = 0; i < S->nF; i++) {
            double a = 0;
            for (int j = 0; j < S->nF; j++)
                a += S->M[i * S->nF + j] * rhs[j];
            dE[i] = a;
        }
        for (int i = 0; i < S->nF; i++)
            S->gvec[S->idxF[i]] += dE[i];
    }
    or_set_gradient(S, S->gvec);
}

/* MixedBCController::activate (mixedBCs.h:181-226) + MixedBC::finalize (:30-46) */
int or_enable_mixed_bc(or_solver *S, int nE, const int32_t *idxE, int nF, const int32_t *idxF, int nsteps,
                       const double *FE, const double *PF, int step)
{
    int n = S->nstr;
    if (nE + nF != n || step < 0 || step >= nsteps)
        return -1;
    S->mixed_active = 1;
    S->nE           = nE;
    S->nF           = nF;
    S->nsteps       = nsteps;
    S->step         = step;
    for (int i = 0; i < nE; i++)
        S->idxE[i] = idxE[i];
    for (int i = 0; i < nF; i++)
        S->idxF[i] = idxF[i];
    free(S->FE);
    free(S->PF);
    S->FE = (double *) malloc(sizeof(double) * (size_t) (nsteps * nE + 1));
    S->PF = (double *) malloc(sizeof(double) * (size_t) (nsteps * nF + 1));
    memcpy(S->FE, FE, sizeof(double) * (size_t) nsteps * nE);
    memcpy(S->PF, PF, sizeof(double) * (size_t) nsteps * nF);
    if (nF) {
        /* M = pinv(Q_F^T C0 Q_F): a principal submatrix of the SPD reference stiffness */
        double A[81];
        int    sym = 1;
        for (int i = 0; i < nF; i++)
            for (int j = 0; j < nF; j++) {
                A[i * nF + j] = S->kref[S->idxF[i] * n + S->idxF[j]];
                if (S->kref[S->idxF[i] * n + S->idxF[j]] != S->kref[S->idxF[j] * n + S->idxF[i]])
                    sym = 0;
            }
        if (sym || gauss_inverse(nF, A, S->M))
            sym_pinv(nF, A, 1e-300, S->M);
    }
    if (step == 0) {
        memset(S->gvec, 0, sizeof(S->gvec));
        if (n == 9) /* finite strain starts from F = I (mixedBCs.h:191-197) */
            S->gvec[0] = S->gvec[4] = S->gvec[8] = 1.0;
        memcpy(S->gvec_prev, S->gvec, sizeof(S->gvec));
    } else {
        double delta[9];
        for (int i = 0; i < n; i++)
            delta[i] = S->gvec[i] - S->gvec_prev[i];
        memcpy(S->gvec_prev, S->gvec, sizeof(S->gvec));
        for (int i = 0; i < nF; i++)
            S->gvec[S->idxF[i]] += delta[S->idxF[i]];
    }
    for (int i = 0; i < nE; i++)
        S->gvec[S->idxE[i]] = S->FE[(size_t) step * nE + i];
    or_set_gradient(S, S->gvec);
    mixed_update(S);
    return 0;
}

void or_disable_mixed_bc(or_solver *S) { S->mixed_active = 0; }
int  or_mixed_active(or_solver *S) { return S->mixed_active; }

/* SolverCG::LineSearchSecant (solverCG.h:119-160) */
static void line_search(or_solver *S)
{
    ptrdiff_t    n     = S->N * S->h;
    double       err   = 10.0;
    const int    MaxIt = S->cfg.ls_max_iter;
    const double tol   = S->cfg.ls_tol;
    int          it    = 0;
    double       ap = 0.0, ac = S->alpha_warm;
    double       rpd = dotp(S, S->r, S->d);
    for (ptrdiff_t i = 0; i < n; i++)
        S->u[i] += S->d[i] * ac;
    mixed_update(S);
    or_residual(S, S->u, S->rnew, 0);
    double r1pd = dotp(S, S->rnew, S->d);
    while (it < MaxIt && err > tol) {
        double denom = r1pd - rpd;
        if (fabs(denom) < 1e-14 * (fabs(r1pd) + fabs(rpd)))
            break;
        double an = ac - r1pd * (ac - ap) / denom;
        if (an <= 0.0)
            an = 0.5 * (ap + ac);
        err = fabs(an - ac);
        for (ptrdiff_t i = 0; i < n; i++)
            S->u[i] += S->d[i] * (an - ac);
        ap  = ac;
        rpd = r1pd;
        ac  = an;
        it++;
        mixed_update(S);
        or_residual(S, S->u, S->rnew, 0);
        r1pd = dotp(S, S->rnew, S->d);
    }
    S->alpha_warm = (it == MaxIt && err > tol) ? 0.1 : ac;
    memcpy(S->r, S->rnew, sizeof(double) * (size_t) n);
}

/* SolverCG::internalSolve (solverCG.h:61-117) */
static void solve_cg(or_solver *S)
{
    ptrdiff_t n = S->N * S->h;
    S->alpha_warm = 0.1;
    memset(S->s, 0, sizeof(double) * (size_t) n);
    memset(S->d, 0, sizeof(double) * (size_t) n);
    or_residual(S, S->u, S->r, 0);
    S->iter        = 0;
    double err_rel = compute_error(S, S->r);
    double delta = 1.0, delta0, deltamid;
    while (S->iter < S->cfg.n_it && err_rel > S->tol) {
        deltamid = dotp(S, S->r, S->s);
        or_convolve(S, S->r, S->s);
        for (ptrdiff_t i = 0; i < n; i++)
            S->s[i] = -S->s[i];
        delta0      = delta;
        delta       = dotp(S, S->r, S->s);
        double beta = fmax(0, (delta - deltamid) / delta0);
        for (ptrdiff_t i = 0; i < n; i++)
            S->d[i] = S->s[i] + beta * S->d[i];
        if (S->all_linear && !S->mixed_active) {
            or_residual(S, S->d, S->rnew, 1);
            double alpha = delta / dotp(S, S->d, S->rnew);
            for (ptrdiff_t i = 0; i < n; i++)
                S->r[i] -= alpha * S->rnew[i];
            for (ptrdiff_t i = 0; i < n; i++)
                S->u[i] -= alpha * S->d[i];
        } else {
            line_search(S);
        }
        S->iter++;
        err_rel = compute_error(S, S->r);
    }
}

/* SolverFP::internalSolve (solverFP.h:33-55) */
static void solve_fp(or_solver *S)
{
    ptrdiff_t n = S->N * S->h;
    or_residual(S, S->u, S->r, 0);
    S->iter        = 0;
    double err_rel = compute_error(S, S->r);
    while (S->iter < S->cfg.n_it && err_rel > S->tol) {
        or_convolve(S, S->r, S->r);
        for (ptrdiff_t i = 0; i < n; i++)
            S->u[i] -= S->r[i];
        mixed_update(S);
        or_residual(S, S->u, S->r, 0);
        S->iter++;
        err_rel = compute_error(S, S->r);
    }
}

/* Solver::solve (solver.h:282-300) */
int or_solve(or_solver *S)
{
    memset(S->err_all, 0, sizeof(double) * (size_t) (S->cfg.n_it + 1));
    if (S->cfg.method == OR_METHOD_FP)
        solve_fp(S);
    else
        solve_cg(S);
    if (S->has_state) { /* updateInternalVariables (J2PlasticityNew.h:37-41, J2Plasticity.h:56-61) */
        size_t ng = (size_t) S->N * S->ngp;
        memcpy(S->ep_t, S->ep, sizeof(double) * ng * 6);
        memcpy(S->q_t, S->q, sizeof(double) * ng);
        if (S->has_visco)
            memcpy(S->pb_t, S->pb, sizeof(double) * ng * 6);
    }
    return S->iter;
}

int or_error_history(or_solver *S, double *out, int cap)
{
    int n = S->iter + 1 < cap ? S->iter + 1 : cap;
    memcpy(out, S->err_all, sizeof(double) * (size_t) n);
    return S->iter + 1;
}

/* Solver::get_homogenized_tangent (solver.h:739-778); like the reference it
 * permanently switches the error type to relative with TOL >= 1e-6 */
int or_homogenized_tangent(or_solver *S, double pert, double *out)
{
    int    n = S->nstr;
    double unpert[MAXSTR], pstress[MAXSTR], g0[MAXSTR], pe[MAXSTR];
    or_homogenized_stress(S, unpert);
    memcpy(g0, S->g0, sizeof(g0));
    S->err_type = OR_ERR_RELATIVE;
    S->tol      = fmax(1e-6, S->tol);
    double C[81];
    for (int i = 0; i < n; i++) {
        if (S->all_linear) {
            memset(pe, 0, sizeof(pe));
            pe[i] = 1.0;
        } else {
            memcpy(pe, g0, sizeof(pe));
            pe[i] += pert;
        }
        or_set_gradient(S, pe);
        S->mixed_active = 0;
        or_solve(S);
        or_homogenized_stress(S, pstress);
        for (int j = 0; j < n; j++)
            C[j * n + i] = S->all_linear ? pstress[j] : (pstress[j] - unpert[j]) / pert;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            out[i * n + j] = 0.5 * (C[i * n + j] + C[j * n + i]);
    return 0;
}

/* Solver::extrapolateDisplacement (solver.h:302-311) */
void or_extrapolate_displacement(or_solver *S)
{
    ptrdiff_t n = S->N * S->h;
    for (ptrdiff_t i = 0; i < n; i++) {
        double delta = S->u[i] - S->u_prev[i];
        S->u_prev[i] = S->u[i];
        S->u[i] += delta;
    }
}

void or_get_u(or_solver *S, double *out) { memcpy(out, S->u, sizeof(double) * (size_t) S->N * S->h); }
void or_set_u(or_solver *S, const double *in) { memcpy(S->u, in, sizeof(double) * (size_t) S->N * S->h); }
void or_get_r(or_solver *S, double *out) { memcpy(out, S->r, sizeof(double) * (size_t) S->N * S->h); }

void or_get_kinematic_state(or_solver *S, double *pb, int committed)
{
    size_t ng = (size_t) S->N * S->ngp;
    if (!S->has_visco)
        memset(pb, 0, sizeof(double) * ng * 6);
    else
        memcpy(pb, committed ? S->pb_t : S->pb, sizeof(double) * ng * 6);
}

void or_get_state(or_solver *S, double *pe, double *q, int committed)
{
    size_t ng = (size_t) S->N * S->ngp;
    if (!S->has_state) {
        if (pe)
            memset(pe, 0, sizeof(double) * ng * 6);
        if (q)
            memset(q, 0, sizeof(double) * ng);
        return;
    }
    if (pe)
        memcpy(pe, committed ? S->ep_t : S->ep, sizeof(double) * ng * 6);
    if (q)
        memcpy(q, committed ? S->q_t : S->q, sizeof(double) * ng);
}

/* element pass of Solver::postprocess (solver.h:503-574) */
/* PseudoPlastic::postprocess (PseudoPlastic.h:56-64) with the flags that
 * get_sigma sets at every Gauss point (PseudoPlasticLinearHardening :103-110 on
 * |dev eps| <= eps_crit, PseudoPlasticNonLinearHardening :153-167 on
 * sqrt(2/3) |dev eps| <= eps_crit): the phase's index within its material group
 * where elastic, that plus the group's phase count where plastic; elements of
 * other models keep 0.  flag[e] is the mean over the element's Gauss points. */
void or_plastic_flag(or_solver *S, double *flag)
{
    int       h = S->h, ns = S->nstr, ng = S->ngp;
    ptrdiff_t N = S->N;
    for (int ix = 0; ix < S->nx; ix++)
        for (int iy = 0; iy < S->ny; iy++)
            for (int iz = 0; iz < S->nz; iz++) {
                ptrdiff_t idx[8];
                double    ue[MAXE], eps[8 * MAXSTR], sig[8 * MAXSTR];
                element_nodes(S, ix, iy, iz, idx);
                int             phase = S->ms[idx[0]];
                const or_group *G     = &S->groups[S->phase_group[phase]];
                flag[idx[0]]          = 0.0;
                if (h != 3 || (G->kind != OR_MODEL_PSEUDO_PLASTIC && G->kind != OR_MODEL_PSEUDO_NL))
                    continue;
                for (int a = 0; a < 8; a++) {
                    ptrdiff_t n = idx[a] >= N ? idx[a] - N : idx[a];
                    for (int c = 0; c < h; c++)
                        ue[h * a + c] = S->u[h * n + c];
                }
                element_strain_stress(S, ue, phase, idx[0], eps, sig);
                const int     loc = S->phase_local[phase];
                const double *p   = G->params + (size_t) loc * G->np;
                double        acc = 0;
                for (int gp = 0; gp < ng; gp++) {
                    const double *e  = eps + gp * ns;
                    double        tr = e[0] + e[1] + e[2], nd = 0;
                    for (int i = 0; i < 6; i++) {
                        double d = i < 3 ? e[i] - tr / 3.0 : e[i];
                        nd += d * d;
                    }
                    nd = sqrt(nd);
                    int plastic;
                    if (G->kind == OR_MODEL_PSEUDO_PLASTIC)
                        plastic = !(nd <= sqrt(2. / 3.) * p[2] / (2. * p[1]));
                    else
                        plastic = !(sqrt(2.0 / 3.0) * nd <= p[4] * pow(p[2] / (3.0 * p[1] * p[4]), 1.0 / (1.0 - p[3])));
                    acc += loc + (plastic ? G->nphase : 0);
                }
                flag[idx[0]] = acc / ng;
            }
}

void or_element_fields(or_solver *S, double *strain_elem, double *stress_elem, double *strain_gp, double *stress_gp,
                       double *stress_avg, double *strain_avg, double *phase_stress_avg, double *phase_strain_avg,
                       int32_t *phase_counts, int n_slots)
{
    int       h = S->h, ns = S->nstr, ng = S->ngp;
    ptrdiff_t N = S->N;
    double    sa[MAXSTR] = {0}, ea[MAXSTR] = {0};
    if (phase_stress_avg)
        memset(phase_stress_avg, 0, sizeof(double) * n_slots * ns);
    if (phase_strain_avg)
        memset(phase_strain_avg, 0, sizeof(double) * n_slots * ns);
    if (phase_counts)
        memset(phase_counts, 0, sizeof(int32_t) * n_slots);
    for (int ix = 0; ix < S->nx; ix++)
        for (int iy = 0; iy < S->ny; iy++)
            for (int iz = 0; iz < S->nz; iz++) {
                ptrdiff_t idx[8];
                double    ue[MAXE], eps[8 * MAXSTR], sig[8 * MAXSTR];
                element_nodes(S, ix, iy, iz, idx);
                for (int a = 0; a < 8; a++) {
                    ptrdiff_t n = idx[a] >= N ? idx[a] - N : idx[a];
                    for (int c = 0; c < h; c++)
                        ue[h * a + c] = S->u[h * n + c];
                }
                int phase = S->ms[idx[0]];
                element_strain_stress(S, ue, phase, idx[0], eps, sig);
                double es[MAXSTR] = {0}, ss[MAXSTR] = {0};
                for (int p = 0; p < ng; p++)
                    for (int j = 0; j < ns; j++) {
                        es[j] += eps[p * ns + j] / ng;
                        ss[j] += sig[p * ns + j] / ng;
                    }
                ptrdiff_t e = idx[0];
                for (int j = 0; j < ns; j++) {
                    if (strain_elem)
                        strain_elem[e * ns + j] = es[j];
                    if (stress_elem)
                        stress_elem[e * ns + j] = ss[j];
                    sa[j] += ss[j];
                    ea[j] += es[j];
                    if (phase < n_slots) {
                        if (phase_stress_avg)
                            phase_stress_avg[phase * ns + j] += ss[j];
                        if (phase_strain_avg)
                            phase_strain_avg[phase * ns + j] += es[j];
                    }
                }
                if (phase_counts && phase < n_slots)
                    phase_counts[phase]++;
                if (strain_gp)
                    memcpy(strain_gp + e * ng * ns, eps, sizeof(double) * ng * ns);
                if (stress_gp)
                    memcpy(stress_gp + e * ng * ns, sig, sizeof(double) * ng * ns);
            }
    for (int j = 0; j < ns; j++) {
        if (stress_avg)
            stress_avg[j] = sa[j] / (double) N;
        if (strain_avg)
            strain_avg[j] = ea[j] / (double) N;
    }
    if (phase_counts)
        for (int p = 0; p < n_slots; p++)
            if (phase_counts[p] > 0)
                for (int j = 0; j < ns; j++) {
                    if (phase_stress_avg)
                        phase_stress_avg[p * ns + j] /= phase_counts[p];
                    if (phase_strain_avg)
                        phase_strain_avg[p * ns + j] /= phase_counts[p];
                }
}
