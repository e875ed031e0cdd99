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
// This header holds the per-element arithmetic and the element-local passes;
// the assembly (gather/scatter) kernel built on it is in assemble_kernels.h.
#pragma once
#include "common.h"

namespace fans {

enum {
    MODEL_THERMAL_ISO    = 1,
    MODEL_THERMAL_TRI    = 2,
    MODEL_ELASTIC_ISO    = 3,
    MODEL_ELASTIC_TRI    = 4,
    MODEL_J2_NEW         = 5,
    MODEL_PSEUDO_PLASTIC = 6,
    MODEL_SVK            = 7,  // SaintVenantKirchhoff (finite strain)
    MODEL_NEO_HOOKE      = 8,  // CompressibleNeoHookean (finite strain)
    MODEL_PSEUDO_NL      = 9,  // PseudoPlasticNonLinearHardening
    MODEL_J2_VISCO_LIN   = 10, // J2ViscoPlastic_LinearIsotropicHardening
    MODEL_J2_VISCO_NL    = 11  // J2ViscoPlastic_NonLinearIsotropicHardening
};

// Strain / stress vector length: thermal gradient (3), small strain Mandel (6)
// or, with finite strain, F and P row major (9).
template <int H, bool LARGE>
constexpr int nstr_of()
{
    return H == 1 ? 3 : (LARGE ? 9 : 6);
}

// Per-phase material table.  p[] holds, by kind:
//   THERMAL_ISO  k                      THERMAL_TRI  K (3x3 row major)
//   ELASTIC_ISO  lambda mu              ELASTIC_TRI  C (6x6 row major)
//   J2_NEW       K G sigma_y H_iso      PSEUDO       K G sigma_y h eps_crit E_s
//   SVK          lambda mu              NEO_HOOKE    lambda mu
//   PSEUDO_NL    K G sigma_y n eps_crit eps_0
//   J2_VISCO_*   K G sigma_y K_iso H_kin eta dt (sigma_inf delta)
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
    int    bbar;           // B-bar element (mechanics, 8 Gauss points)
    int    pad;
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
    // one reciprocal instead of twelve divisions by s_t; below the 1e-12 guard
    // the updates add exact zeros, which leaves ep_in and s untouched as the
    // reference's branch does
    const double inv = s_t > 1e-12 ? 1.0 / s_t : 0.0;
    const double ge = dgamma * inv, gs = dgamma * 2.0 * G * inv;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        ep[i] = ep_in[i] + ge * st[i];
        s[i]  = s[i] - gs * st[i];
    }
    qo = q_in + c23 * dgamma;
}

// J2Plasticity::get_sigma (J2Plasticity.h:65-108) with the isotropic hardening
// of J2ViscoPlastic_LinearIsotropicHardening (:166-174) or
// J2ViscoPlastic_NonLinearIsotropicHardening (:197-221), FP64.  p = K G sigma_y
// K_iso H_kin eta dt (sigma_inf delta); st = committed plastic strain (6),
// psi_t, psi_bar_t (6).  Returns sigma, the plastic strain eps_t + gamma n and
// gamma and n, with which the caller advances the current psi (+ gamma
// sqrt(2/3)) and psi_bar (- gamma n) once per call, as the reference does at
// every get_sigma (:102-104).  The nonlinear law's Newton loop is the
// reference's: slope -den (its second term is (2 / 3) * ... in integer
// arithmetic, i.e. zero, :217), stopping at the first increment not above
// 1e-10 or after 10 steps.
__device__ __forceinline__ void visco_point(int kind, const double *p, const double (&e)[6], const double (&st)[13],
                                            double (&s)[6], double (&ep)[6], double &gamma, double (&n)[6])
{
    const double c23 = 0.81649658092772603273;  // sqrt(2/3)
    const double K = p[0], G = p[1], sy = p[2], Kh = p[3], Hk = p[4];
    const double psi_t = st[6];
    double       ee[6];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        ee[i] = e[i] - st[i];
    const double treps = ee[0] + ee[1] + ee[2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s[i] = (K - 2.0 * G / 3.0) * treps + 2 * G * ee[i];
#pragma unroll
    for (int i = 3; i < 6; ++i)
        s[i] = 2 * G * ee[i];
    const double mean = (s[0] + s[1] + s[2]) / 3.0;
    double       dmq[6], nrm2 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        dmq[i] = (s[i] - (i < 3 ? mean : 0.0)) - (-(2.0 / 3.0) * Hk * st[7 + i]);
        nrm2 += dmq[i] * dmq[i];
    }
    const double nrm  = sqrt(nrm2);
    const bool   nlaw = kind == MODEL_J2_VISCO_NL;
    const double q_tr = nlaw ? -Kh * psi_t - (p[7] - sy) * (1 - exp(-p[8] * psi_t)) : -Kh * psi_t;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        n[i] = nrm < 1e-12 ? 0.0 : dmq[i] / nrm;
    const double f   = nrm - c23 * (sy - q_tr);
    const double den = 2 * G + (2.0 / 3.0) * (Kh + Hk) + p[5] / p[6];
    gamma            = 0.0;
    if (!(f < 0)) {
        if (!nlaw) {
            gamma = f / den;
        } else {
            const double sd = c23 * (p[7] - sy), dl = p[8], e0 = exp(-dl * psi_t);
            double       inc = 1.0;
            for (int it = 0; inc > 1e-10 && it < 10; ++it) {
                const double g = f - gamma * den - sd * (-exp(-dl * (psi_t + c23 * gamma)) + e0);
                inc            = g / den;  // -g / dg with dg = -den
                gamma += inc;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        s[i] -= gamma * 2 * G * n[i];
        ep[i] = st[i] + gamma * n[i];
    }
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

// PseudoPlasticNonLinearHardening::get_sigma (PseudoPlastic.h:143-168):
// sigma = K tr(eps) 1 + 2 G dev(eps) up to eps_eq = sqrt(2/3) |dev eps| =
// eps_crit, beyond it sigma_y (eps_eq / eps_0)^n along dev / |dev| (scaled by
// sqrt(2/3), Mandel)
template <typename T>
__device__ __forceinline__ void pseudo_nl_point(const T *p, const T (&e)[6], T (&s)[6])
{
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
    nd            = sqrt(nd);
    const T c23   = (T) 0.81649658092772603273;  // sqrt(2/3)
    const T eq    = c23 * nd;
    const T buf1  = p[0] * tr;
    const T buf2  = eq <= p[4] ? (T) 2.0 * p[1] : c23 * p[2] * pow(eq / p[5], p[3]) / nd;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        s[i] = buf1 + buf2 * dev[i];
#pragma unroll
    for (int i = 3; i < 6; ++i)
        s[i] = buf2 * dev[i];
}

// PseudoPlastic flag of one Gauss point (PseudoPlastic.h:103-110, :153-167):
// the phase's index in its material group (p[6]), plus the group's phase count
// (p[7]) where the point is plastic, decided as in pseudo_point / pseudo_nl_point
template <typename T>
__device__ __forceinline__ double pseudo_flag(const Phase<T> &P, const T (&e)[6])
{
    const T tr = e[0] + e[1] + e[2];
    T       nd = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const T d = i < 3 ? e[i] - (T) (1.0 / 3.0) * tr : e[i];
        nd += d * d;
    }
    nd                 = sqrt(nd);
    const bool elastic = P.kind == MODEL_PSEUDO_PLASTIC ? nd <= P.p[4] : (T) 0.81649658092772603273 * nd <= P.p[4];
    return (double) P.p[6] + (elastic ? 0.0 : (double) P.p[7]);
}

// First Piola-Kirchhoff stress P = F S(F) of the finite strain models
// (LargeStrainMechModel.h:193-199); F and P row major.
//   SaintVenantKirchhoff   S = lam tr(E) I + 2 mu E, E = (F^T F - I) / 2
//                          (SaintVenantKirchhoff.h:34-39)
//   CompressibleNeoHookean S = lam ln(J) C^-1 + mu (I - C^-1), C = F^T F
//                          (CompressibleNeoHookean.h:35-49); J <= 0 gives NaN
//                          (the reference throws there)
template <typename T>
__device__ __forceinline__ void large_point(int kind, const T *p, const T (&e)[9], T (&s)[9])
{
    const T lam = p[0], mu = p[1];
    T       C[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            C[a][b] = e[a] * e[b] + e[3 + a] * e[3 + b] + e[6 + a] * e[6 + b];
    T S[3][3];
    if (kind == MODEL_SVK) {
        const T htr = (T) 0.5 * (C[0][0] + C[1][1] + C[2][2] - (T) 3);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                S[a][b] = mu * (C[a][b] - (a == b ? (T) 1 : (T) 0)) + (a == b ? lam * htr : (T) 0);
    } else {
        // C^-1 from the cofactors of the symmetric C; det C = J^2
        const T c00 = C[1][1] * C[2][2] - C[1][2] * C[1][2];
        const T c01 = C[0][2] * C[1][2] - C[0][1] * C[2][2];
        const T c02 = C[0][1] * C[1][2] - C[0][2] * C[1][1];
        const T c11 = C[0][0] * C[2][2] - C[0][2] * C[0][2];
        const T c12 = C[0][1] * C[0][2] - C[0][0] * C[1][2];
        const T c22 = C[0][0] * C[1][1] - C[0][1] * C[0][1];
        const T detC = C[0][0] * c00 + C[0][1] * c01 + C[0][2] * c02;
        const T J    = e[0] * (e[4] * e[8] - e[5] * e[7]) - e[1] * (e[3] * e[8] - e[5] * e[6]) +
                    e[2] * (e[3] * e[7] - e[4] * e[6]);
        const T ic   = (T) 1 / detC;
        const T Ci[3][3] = {{c00 * ic, c01 * ic, c02 * ic}, {c01 * ic, c11 * ic, c12 * ic},
                            {c02 * ic, c12 * ic, c22 * ic}};
        const T lj = J > (T) 0 ? log(J) : (T) NAN;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                S[a][b] = lam * lj * Ci[a][b] + mu * ((a == b ? (T) 1 : (T) 0) - Ci[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            s[3 * a + b] = e[3 * a] * S[0][b] + e[3 * a + 1] * S[1][b] + e[3 * a + 2] * S[2][b];
}

// State of J2 points, structure of arrays over slots laid out like the tiles
// of the stencil pass: z is cut into tiles of kStateTile elements, and the 16
// slots of tile tz hold elements z = kStateTile * tz - 1 + lz (lz = 0..15,
// periodic in z), so that one row of the stencil pass reads whole 128-byte
// lines.  Elements at a tile seam (and elements wrapped past the last tile)
// own two slots, which always hold the same values.  Component c (0..5 plastic
// strain, 6 q) of Gauss point g of slot s = ((x * ny + y) * ntz + tz) * 16 + lz
// sits at st[(c * ngp + g) * nslot + s].
constexpr int kStateTile = 15;
// J2PlasticityNew: components 0..5 plastic strain, 6 q.  J2ViscoPlastic: 0..5
// plastic strain, 6 psi, 7..12 psi_bar (committed, all slots of an element),
// 13 psi and 14..19 psi_bar of the step in progress (canonical slot only).
constexpr int kJ2Comps = 7, kViscoComps = 20;
struct J2State {
    double *st;     // committed (_t) state
    int64_t nslot;  // slots per component and Gauss point: nx * ny * ntz * 16
};

// the element of slot s; canonical: the one slot of the element that every
// pass but the state commit (which writes all of them) works on
__device__ __forceinline__ void state_slot_element(int64_t s, int ny, int nz, int ntz, int &x, int &y, int &z,
                                                   bool &canonical)
{
    const int     lz = (int) (s & 15);
    const int64_t r  = s >> 4;
    const int     tz = (int) (r % ntz);
    const int64_t xy = r / ntz;
    y                = (int) (xy % ny);
    x                = (int) (xy / ny);
    const int zz     = kStateTile * tz - 1 + lz;
    canonical        = lz >= 1 && zz < nz;
    z                = ((zz % nz) + nz) % nz;
}

// sigma = model(eps) for one Gauss point; returns in s
template <typename T, int H>
__device__ __forceinline__ void material_point(const Phase<T> &P, const T (&e)[H == 1 ? 3 : 6], T (&s)[H == 1 ? 3 : 6])
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
        case MODEL_J2_NEW:
        case MODEL_PSEUDO_NL:
            // J2 points need their state, and both these models are nonlinear:
            // PointModel evaluates them; the linear-operator paths that call
            // this never see such a phase
#pragma unroll
            for (int i = 0; i < 6; ++i)
                s[i] = (T) NAN;
            break;
        default: pseudo_point<T>(P.p, e, s); break;
        }
    }
}

// One element's constitutive model on the small-strain per-Gauss-point paths.
// The phase kind and its first six parameters are read once per element (every
// kind but the triclinic ones needs no more), and the J2 state of Gauss point
// gp + 1 is requested while point gp is evaluated, so the state loads overlap
// the previous point's arithmetic instead of each waiting on its own.  A caller
// that walks a row of elements names the next element's state (sn) so that the
// last Gauss point requests that element's point 0 (begin() then skips it).
// PNL: the power-law PseudoPlastic model may occur (its pow() pushes the
// residual kernel past 256 registers, so that kernel has a variant without it).
// VISCO: the J2ViscoPlastic models may occur; their state is read at the
// element's slot without prefetch, and with `acc` every evaluation advances the
// current psi / psi_bar of the slot (an element's canonical slot, evaluated
// once per pass) like one get_sigma call of the reference.
template <typename T, int H, int NGP, bool PNL = true, bool VISCO = false>
struct PointModel {
    const Phase<T> *P;
    int             kind;
    T               pr[6];
    const double   *sb;     // J2 element: its state slot at Gauss point 0, else null
    const double   *sn;     // next J2 element of the row (point 0), else null
    int64_t         nel;    // state stride between Gauss points (slots)
    double          nx[7];  // state of the next Gauss point
    double         *vs;     // VISCO: the element's state slot (Gauss point 0), else null
    bool            vacc;   // VISCO: advance the current psi / psi_bar

    __device__ __forceinline__ static const double *state_of(int kind, const J2State &js, int64_t slot)
    {
        return (H == 3 && kind == MODEL_J2_NEW) ? js.st + slot : nullptr;
    }
    // slot: a state slot of the element (J2State); sn: the next element's state
    // (from state_of), or null; prefetched: point 0 of this element was
    // requested by the previous element's last point
    __device__ __forceinline__ void begin(const Phase<T> &ph, int knd, const J2State &js, int64_t slot,
                                          const double *next, bool prefetched, bool accumulate = false)
    {
        P    = &ph;
        kind = knd;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            pr[i] = ph.p[i];
        nel = js.nslot;
        sb  = state_of(kind, js, slot);
        sn  = next;
        if constexpr (VISCO) {
            vs   = (kind == MODEL_J2_VISCO_LIN || kind == MODEL_J2_VISCO_NL) ? js.st + slot : nullptr;
            vacc = accumulate;
        }
        if (!prefetched)
            fetch(0);
    }
    __device__ __forceinline__ void init(const Phase<T> &ph, const J2State &js, int64_t slot, bool accumulate = false)
    {
        begin(ph, ph.kind, js, slot, nullptr, false, accumulate);
    }
    __device__ __forceinline__ static void load(const double *b, int gp, int64_t nel, double (&v)[7])
    {
        if (b)
#pragma unroll
            for (int c = 0; c < 7; ++c)
                v[c] = b[((int64_t) c * NGP + gp) * nel];
    }
    __device__ __forceinline__ void fetch(int gp) { load(sb, gp, nel, nx); }
    __device__ __forceinline__ void visco_eval(int gp, const T (&e)[H == 1 ? 3 : 6], T (&s)[H == 1 ? 3 : 6]);
    __device__ __forceinline__ void eval(int gp, const T (&e)[H == 1 ? 3 : 6], T (&s)[H == 1 ? 3 : 6])
    {
        if constexpr (H == 1) {
            if (kind == MODEL_THERMAL_ISO) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    s[i] = pr[0] * e[i];
            } else {
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    s[i] = P->p[3 * i] * e[0] + P->p[3 * i + 1] * e[1] + P->p[3 * i + 2] * e[2];
            }
        } else {
            double cur[7];
#pragma unroll
            for (int c = 0; c < 7; ++c)
                cur[c] = nx[c];
            if (gp + 1 < NGP)
                fetch(gp + 1);
            else
                load(sn, 0, nel, nx);
            switch (kind) {
            case MODEL_ELASTIC_ISO: {
                const T b1 = pr[0] * (e[0] + e[1] + e[2]), b2 = (T) 2 * pr[1];
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
                        a += P->p[6 * i + j] * e[j];
                    s[i] = a;
                }
                break;
            case MODEL_J2_NEW: {
                double ed[6], epin[6], sd[6], epo[6], qo, pd[4];
#pragma unroll
                for (int i = 0; i < 6; ++i) {
                    ed[i]   = (double) e[i];
                    epin[i] = cur[i];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    pd[i] = (double) pr[i];
                j2_point(pd, ed, epin, cur[6], sd, epo, qo);
#pragma unroll
                for (int i = 0; i < 6; ++i)
                    s[i] = (T) sd[i];
            } break;
            default:
                if constexpr (VISCO && sizeof(T) == 8)
                    if (vs) {
                        visco_eval(gp, e, s);
                        break;
                    }
                if constexpr (PNL)
                    if (kind == MODEL_PSEUDO_NL) {
                        pseudo_nl_point<T>(pr, e, s);
                        break;
                    }
                pseudo_point<T>(pr, e, s);
                break;
            }
        }
    }
};

// one J2ViscoPlastic get_sigma at Gauss point gp (PointModel with VISCO)
template <typename T, int H, int NGP, bool PNL, bool VISCO>
__device__ __forceinline__ void visco_gp(PointModel<T, H, NGP, PNL, VISCO> &m, int gp, const T (&e)[6], T (&s)[6])
{
    const int64_t nel = m.nel;
    double        st[13], ed[6], sd[6], epo[6], gamma, n[6], pd[9];
#pragma unroll
    for (int c = 0; c < 13; ++c)
        st[c] = m.vs[((int64_t) c * NGP + gp) * nel];
#pragma unroll
    for (int i = 0; i < 6; ++i)
        ed[i] = (double) e[i];
#pragma unroll
    for (int i = 0; i < 9; ++i)
        pd[i] = (double) m.P->p[i];
    visco_point(m.kind, pd, ed, st, sd, epo, gamma, n);
#pragma unroll
    for (int i = 0; i < 6; ++i)
        s[i] = (T) sd[i];
    if (m.vacc && gamma != 0.0) {
        // psi += gamma sqrt(2/3), psi_bar -= gamma n (J2Plasticity.h:103-104)
        m.vs[((int64_t) 13 * NGP + gp) * nel] += gamma * 0.81649658092772603273;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            m.vs[((int64_t) (14 + i) * NGP + gp) * nel] -= gamma * n[i];
    }
}

template <typename T, int H, int NGP, bool PNL, bool VISCO>
__device__ __forceinline__ void PointModel<T, H, NGP, PNL, VISCO>::visco_eval(int gp, const T (&e)[H == 1 ? 3 : 6],
                                                                           T (&s)[H == 1 ? 3 : 6])
{
    if constexpr (H == 3)
        visco_gp(*this, gp, e, s);
}

// sigma of the linear models only (the linear matvec K d never sees another
// kind; anything else gives NaN)
template <typename T, int H>
__device__ __forceinline__ void linear_point(const Phase<T> &P, const T (&e)[H == 1 ? 3 : 6],
                                             T (&s)[H == 1 ? 3 : 6])
{
    constexpr int n = H == 1 ? 3 : 6;
    if (P.kind == MODEL_THERMAL_ISO || P.kind == MODEL_ELASTIC_ISO) {
        if constexpr (H == 1) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[i] = P.p[0] * e[i];
        } else {
            const T b1 = P.p[0] * (e[0] + e[1] + e[2]), b2 = (T) 2 * P.p[1];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                s[i] = b1 + b2 * e[i];
#pragma unroll
            for (int i = 3; i < 6; ++i)
                s[i] = b2 * e[i];
        }
    } else if (P.kind == MODEL_THERMAL_TRI || P.kind == MODEL_ELASTIC_TRI) {
#pragma unroll
        for (int i = 0; i < n; ++i) {
            T a = 0;
#pragma unroll
            for (int j = 0; j < n; ++j)
                a += P.p[n * i + j] * e[j];
            s[i] = a;
        }
    } else {
#pragma unroll
        for (int i = 0; i < n; ++i)
            s[i] = (T) NAN;
    }
}

// material_point for any strain vector length (finite strain: NS = 9)
template <typename T, int H, int NS>
__device__ __forceinline__ void material_point_ns(const Phase<T> &P, const T (&e)[NS], T (&s)[NS])
{
    if constexpr (NS == 9)
        large_point<T>(P.kind, P.p, e, s);
    else
        material_point<T, H>(P, e, s);
}

// ---------------------------------------------------------------------------
// element kernel core: u[8][H] (corner values) -> f[8][H] = B^T sigma(B u + g0) w
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T lerp1(T a, T b, T t)
{
    return a + t * (b - a);
}

// strain (Mandel) from displacement gradient g[c][d] = d u_c / d x_d; with
// finite strain (NS = 9) F = g + g0 row major (LargeStrainMechModel.h:205-222)
template <typename T, int H, int NS = (H == 1 ? 3 : 6)>
__device__ __forceinline__ void grad_to_strain(const T (&g)[H][3], const T *g0, T (&e)[NS])
{
    if constexpr (NS == 9) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                e[3 * c + d] = g[c][d] + g0[3 * c + d];
    } else if constexpr (H == 1) {
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
// (finite strain: T = P)
template <typename T, int H, int NS = (H == 1 ? 3 : 6)>
__device__ __forceinline__ void stress_to_T(const T (&s)[NS], T (&Tt)[H][3])
{
    if constexpr (NS == 9) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                Tt[c][d] = s[3 * c + d];
    } else if constexpr (H == 1) {
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

// B-bar (include/matmodel.h:113-140): rows 0..2 of B at every Gauss point are
// shifted by the difference between their centre sum and their own sum over
// 3; the centre value of the trilinear element equals the mean over the 2x2x2
// Gauss points.  Rows 0..2 are the normal strains (gradient entries (r, r));
// with finite strain the reference applies the same rows, F_11 F_12 F_13
// (entries (0, r)).  Applied to the gradient at Gauss point (i,j,k) those
// entries shift by (t_c - t_gp) / 3 with t the row sum; the transpose maps a
// stress s_gp to B^T (s_gp + (mean_gp(t(s)) - t(s_gp)) / 3 on rows 0..2).
template <int NS>
__device__ constexpr int bbar_comp(int r)
{
    return NS == 9 ? 0 : r;
}

template <typename T, int NS>
__device__ __forceinline__ T bbar_mean_trace(const T (&Gx)[2][2][3], const T (&Gy)[2][2][3], const T (&Gz)[2][2][3])
{
    constexpr int c0 = bbar_comp<NS>(0), c1 = bbar_comp<NS>(1), c2 = bbar_comp<NS>(2);
    return (T) 0.25 * (((Gx[0][0][c0] + Gx[1][1][c0]) + (Gx[0][1][c0] + Gx[1][0][c0])) +
                       ((Gy[0][0][c1] + Gy[1][1][c1]) + (Gy[0][1][c1] + Gy[1][0][c1])) +
                       ((Gz[0][0][c2] + Gz[1][1][c2]) + (Gz[0][1][c2] + Gz[1][0][c2])));
}

template <typename T, int NS>
__device__ __forceinline__ void bbar_project_grad(T (&gr)[3][3], T trc)
{
    constexpr int c0 = bbar_comp<NS>(0), c1 = bbar_comp<NS>(1), c2 = bbar_comp<NS>(2);
    const T       d  = (trc - (gr[c0][0] + gr[c1][1] + gr[c2][2])) * (T) (1.0 / 3.0);
    gr[c0][0] += d;
    gr[c1][1] += d;
    gr[c2][2] += d;
}

// Full element: f = sum_gp B^T sigma(B u + g0) * w.  MatF(gp, e, s) evaluates
// the constitutive law at Gauss point gp (index i + 2j + 4k).
template <typename T, int H, int NGP, int NS = (H == 1 ? 3 : 6), typename MatF>
__device__ __forceinline__ void element_forces(const T (&u)[8][H], const Geom &g, const T *g0, MatF &&mat,
                                               T (&f)[8][H])
{
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
    constexpr bool CAN_BBAR = H == 3 && NGP == 8;
    const bool     bbar     = CAN_BBAR && g.bbar;
    T              trc = (T) 0, trs = (T) 0;
    if constexpr (CAN_BBAR)
        if (bbar)
            trc = bbar_mean_trace<T, NS>(Gx, Gy, Gz);
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
        if constexpr (CAN_BBAR)
            if (bbar)
                bbar_project_grad<T, NS>(gr, trc);
        T e[NS], s[NS], Tt[H][3];
        grad_to_strain<T, H, NS>(gr, g0, e);
        mat(gp, e, s);
        if constexpr (CAN_BBAR)
            if (bbar) {
                const T t = s[0] + s[1] + s[2];
                trs += t;
                const T t3 = t * (T) (1.0 / 3.0);
                s[0] -= t3;
                s[1] -= t3;
                s[2] -= t3;
            }
        stress_to_T<T, H, NS>(s, Tt);
#pragma unroll
        for (int c = 0; c < H; ++c) {
            Sx[j][k][c] += Tt[c][0];
            Sy[i][k][c] += Tt[c][1];
            Sz[i][j][c] += Tt[c][2];
        }
    }
    if constexpr (CAN_BBAR)
        if (bbar) {
            // m mean(tr sigma) / 3 at each of the 8 points: every factor entry
            // collects two points
            const T       v  = trs * (T) (2.0 / 24.0);
            constexpr int c0 = bbar_comp<NS>(0), c1 = bbar_comp<NS>(1), c2 = bbar_comp<NS>(2);
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    Sx[a][b][c0] += v;
                    Sy[a][b][c1] += v;
                    Sz[a][b][c2] += v;
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

// Isotropic fast path: with sigma = lambda tr(eps) I + 2 mu eps (Mandel), the
// Gauss-point loop collapses.  T[c][d] = lambda tr(grad u) delta_cd +
// mu (du_c/dx_d + du_d/dx_c) is linear in the gradient factors, and the sums
// over the Gauss index that a factor does not depend on reduce to pairwise
// sums.  lw = lambda * w, mw = mu * w (w = v_e / n_gp); T0 is the traction
// tensor of the macroscopic gradient (already multiplied by w), added once
// per Gauss point.
template <typename T, int NGP>
__device__ __forceinline__ void element_forces_iso3(const T (&u)[8][3], const Geom &g, T lw, T mw, const T (&T0)[3][3],
                                                    T (&f)[8][3])
{
    T Gx[2][2][3], Gy[2][2][3], Gz[2][2][3];
    grad_factors<T, 3, NGP>(u, g, Gx, Gy, Gz);
    T Sx[2][2][3], Sy[2][2][3], Sz[2][2][3];
    if constexpr (NGP == 8) {
        const T l2m2 = (T) 2 * (lw + (T) 2 * mw), m2 = (T) 2 * mw, g2 = (T) 2;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // sums over the Gauss index the factor does not depend on
                const T sGy_i0 = Gy[0][b][0] + Gy[1][b][0], sGy_i1 = Gy[0][b][1] + Gy[1][b][1];  // [k=b]
                const T sGz_i0 = Gz[0][a][0] + Gz[1][a][0], sGz_i2 = Gz[0][a][2] + Gz[1][a][2];  // [j=a]
                Sx[a][b][0] = l2m2 * Gx[a][b][0] + lw * (sGy_i1 + sGz_i2) + g2 * T0[0][0];
                Sx[a][b][1] = m2 * Gx[a][b][1] + mw * sGy_i0 + g2 * T0[1][0];
                Sx[a][b][2] = m2 * Gx[a][b][2] + mw * sGz_i0 + g2 * T0[2][0];
                const T sGx_j0 = Gx[0][b][0] + Gx[1][b][0], sGx_j1 = Gx[0][b][1] + Gx[1][b][1];  // [k=b]
                const T sGz_j1 = Gz[a][0][1] + Gz[a][1][1], sGz_j2 = Gz[a][0][2] + Gz[a][1][2];  // [i=a]
                Sy[a][b][0] = m2 * Gy[a][b][0] + mw * sGx_j1 + g2 * T0[0][1];
                Sy[a][b][1] = l2m2 * Gy[a][b][1] + lw * (sGx_j0 + sGz_j2) + g2 * T0[1][1];
                Sy[a][b][2] = m2 * Gy[a][b][2] + mw * sGz_j1 + g2 * T0[2][1];
                const T sGx_k0 = Gx[b][0][0] + Gx[b][1][0], sGx_k2 = Gx[b][0][2] + Gx[b][1][2];  // [j=b]
                const T sGy_k1 = Gy[a][0][1] + Gy[a][1][1], sGy_k2 = Gy[a][0][2] + Gy[a][1][2];  // [i=a]
                Sz[a][b][0] = m2 * Gz[a][b][0] + mw * sGx_k2 + g2 * T0[0][2];
                Sz[a][b][1] = m2 * Gz[a][b][1] + mw * sGy_k2 + g2 * T0[1][2];
                Sz[a][b][2] = l2m2 * Gz[a][b][2] + lw * (sGx_k0 + sGy_k1) + g2 * T0[2][2];
            }
    } else {
        const T gr[3][3] = {{Gx[0][0][0], Gy[0][0][0], Gz[0][0][0]},
                            {Gx[0][0][1], Gy[0][0][1], Gz[0][0][1]},
                            {Gx[0][0][2], Gy[0][0][2], Gz[0][0][2]}};
        const T tr       = gr[0][0] + gr[1][1] + gr[2][2];
        T       Tt[3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d)
                Tt[c][d] = mw * (gr[c][d] + gr[d][c]) + (c == d ? lw * tr : (T) 0) + T0[c][d];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Sx[0][0][c] = Tt[c][0];
            Sy[0][0][c] = Tt[c][1];
            Sz[0][0][c] = Tt[c][2];
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            f[a][c] = (T) 0;
    grad_factors_T<T, 3, NGP>(Sx, Sy, Sz, g, f);
}

// isotropic conduction: sigma = k grad + k g0, the Gauss sums are factors of 2
template <typename T, int NGP>
__device__ __forceinline__ void element_forces_iso1(const T (&u)[8][1], const Geom &g, T kw, const T (&T0)[3],
                                                    T (&f)[8][1])
{
    T Gx[2][2][1], Gy[2][2][1], Gz[2][2][1];
    grad_factors<T, 1, NGP>(u, g, Gx, Gy, Gz);
    const T s = NGP == 8 ? (T) 2 * kw : kw, s0 = NGP == 8 ? (T) 2 : (T) 1;
    T       Sx[2][2][1], Sy[2][2][1], Sz[2][2][1];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            Sx[a][b][0] = s * Gx[a][b][0] + s0 * T0[0];
            Sy[a][b][0] = s * Gy[a][b][0] + s0 * T0[1];
            Sz[a][b][0] = s * Gz[a][b][0] + s0 * T0[2];
        }
    f[0][0] = f[1][0] = f[2][0] = f[3][0] = f[4][0] = f[5][0] = f[6][0] = f[7][0] = (T) 0;
    grad_factors_T<T, 1, NGP>(Sx, Sy, Sz, g, f);
}

// ---------------------------------------------------------------------------
// Modal form of the isotropic element operators.
//
// With reference coordinates (xi, eta, zeta) in [-1,1]^3, the derivative of the
// trilinear field along x is bilinear in (eta, zeta):
//   du/dx = kx/4 (A0 + A1 eta + A2 zeta + A3 eta zeta),  kx = 1/hx,
// where A = H Dx is the 4-point Hadamard transform of the four x-edge
// differences (likewise for y over (xi, zeta) and z over (xi, eta)).  The
// products of two such fields have degree <= 3 in every coordinate, so the
// 2x2x2 Gauss rule integrates the energy density exactly, and the modes
// {1, xi, eta, zeta, xi eta, ...} are orthogonal under it (norms 8 / 3^deg).
// The element energy lambda/2 tr(eps)^2 + mu eps:eps therefore splits into
// independent modal blocks:
//   mode 1        : all nine gradient entries
//   mode xi       : d/dy, d/dz   (eta: d/dx, d/dz; zeta: d/dx, d/dy)
//   mode eta zeta : d/dx only    (xi zeta: d/dy; xi eta: d/dz)
// and the forces are the transposed Hadamard / edge-difference maps applied to
// the modal stresses.  With 8 Gauss points both maps are one 3-D Walsh
// transform of the corner values per component (walsh8: the Hadamard modes of
// the edge differences are Walsh coefficients, and the mixed modes that two
// derivative directions share are the same coefficient).  This is the same
// operator as the Gauss-point loop (the quadrature is exact for it) at about
// 250 instead of ~600 FP64 operations per element.  For one Gauss point
// (HEX8R) only mode 1 remains.
// ---------------------------------------------------------------------------

// forward: D[p][q] (p, q in {0,1}) -> A0..A3
template <typename T>
__device__ __forceinline__ void hadamard4(T &d00, T &d10, T &d01, T &d11)
{
    const T s0 = d00 + d01, s1 = d10 + d11, e0 = d01 - d00, e1 = d11 - d10;
    d00 = s0 + s1;  // A0
    d10 = s1 - s0;  // A1 (slope in p)
    d01 = e0 + e1;  // A2 (slope in q)
    d11 = e1 - e0;  // A3 (p q)
}
// transpose: A0..A3 -> D[p][q]
template <typename T>
__device__ __forceinline__ void hadamard4_T(T &a0, T &a1, T &a2, T &a3)
{
    const T p0 = a0 - a2, p1 = a1 - a3, q0 = a0 + a2, q1 = a1 + a3;
    a0 = p0 - p1;  // D00
    a1 = p0 + p1;  // D10
    a2 = q0 - q1;  // D01
    a3 = q0 + q1;  // D11
}

// unnormalised 3-D Walsh-Hadamard transform of the 8 corner values (corner /
// coefficient index i + 2j + 4k): W[ijk] = sum_xyz (-1)^(ix+jy+kz) v[xyz].
// The Hadamard modes of the edge differences are Walsh coefficients:
//   d/dx: A0..A3 = -W100, W110, W101, -W111   (d/dy: -W010, W110, W011, -W111;
//   d/dz: -W001, W101, W011, -W111)
// so one transform of 24 additions replaces the 3 x (4 differences + 8-addition
// Hadamard) per component, and the transposed maps are one transform of the
// mode sums.  The transform is symmetric (its own transpose).
template <typename T>
__device__ __forceinline__ void walsh8(T (&v)[8])
{
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int a = 0; a < 8; ++a)
            if ((a & s) == 0) {
                const T x = v[a], y = v[a + s];
                v[a]      = x + y;
                v[a + s]  = x - y;
            }
}

// modal block with two derivative directions dA, dB present (third missing):
// a[c] = coefficient of d u_c / d x_dA, b[c] of d u_c / d x_dB
template <typename T, int dA, int dB>
__device__ __forceinline__ void modal_pair(T (&a)[3], T (&b)[3], T L, T M)
{
    constexpr int c0 = 3 - dA - dB;
    const T       t  = L * (a[dA] + b[dB]);
    const T       m2 = M + M;
    const T       off = M * (a[dB] + b[dA]);
    a[dA]            = t + m2 * a[dA];
    b[dB]            = t + m2 * b[dB];
    a[dB]            = off;
    b[dA]            = off;
    a[c0] *= M;
    b[c0] *= M;
}

template <typename T, int d>
__device__ __forceinline__ void modal_single(T (&a)[3], T L, T M)
{
#pragma unroll
    for (int c = 0; c < 3; ++c)
        a[c] *= (c == d) ? (L + M + M) : M;
}

// f = Ke(lambda, mu) u (+ the forces of the traction T0 = w sigma(g0)); same
// arguments as element_forces_iso3.  CUBIC (hx = hy = hz): the 1/h of the edge
// differences and of their transpose fold into the modal constants (k^2) and
// the traction term (k), which saves 72 multiplications per element.
template <typename T, int NGP, bool CUBIC = false>
__device__ __forceinline__ void element_forces_iso3_modal(const T (&u)[8][3], const Geom &g, T lw, T mw,
                                                          const T (&T0)[3][3], T (&f)[8][3])
{
    const T kc   = (T) g.ilx;
    const T k[3] = {CUBIC ? (T) 1 : (T) g.ilx, CUBIC ? (T) 1 : (T) g.ily, CUBIC ? (T) 1 : (T) g.ilz};
    // A[d][mode][c]: Hadamard modes of the edge differences along d, scaled by 1/h_d
    T A[3][4][3];
    if constexpr (NGP == 8) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T w[8];
#pragma unroll
            for (int n = 0; n < 8; ++n)
                w[n] = u[n][c];
            walsh8(w);
            A[0][0][c] = -w[1] * k[0], A[0][1][c] = w[3] * k[0], A[0][2][c] = w[5] * k[0], A[0][3][c] = -w[7] * k[0];
            A[1][0][c] = -w[2] * k[1], A[1][1][c] = w[3] * k[1], A[1][2][c] = w[6] * k[1], A[1][3][c] = -w[7] * k[1];
            A[2][0][c] = -w[4] * k[2], A[2][1][c] = w[5] * k[2], A[2][2][c] = w[6] * k[2], A[2][3][c] = -w[7] * k[2];
        }
    } else {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    A[0][p + 2 * q][c] = (u[1 + 2 * p + 4 * q][c] - u[2 * p + 4 * q][c]) * k[0];
                    A[1][p + 2 * q][c] = (u[p + 2 + 4 * q][c] - u[p + 4 * q][c]) * k[1];
                    A[2][p + 2 * q][c] = (u[p + 2 * q + 4][c] - u[p + 2 * q][c]) * k[2];
                }
    }
    // Gauss weight w = V / NGP; mode weights NGP / 3^deg; 1/16 from the two
    // factors 1/4 of the bilinear expansion
    const T s  = CUBIC ? (T) NGP / (T) 16 * kc * kc : (T) NGP / (T) 16;
    const T L0 = s * lw, M0 = s * mw;
    if constexpr (NGP == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                A[d][0][c] = (A[d][0][c] + A[d][2][c]) + (A[d][1][c] + A[d][3][c]);
    }
    // mode 1: all nine entries
    {
        const T t  = L0 * (A[0][0][0] + A[1][0][1] + A[2][0][2]);
        const T m2 = M0 + M0;
        const T t0 = CUBIC ? (T) NGP * (T) 0.25 * kc : (T) NGP * (T) 0.25;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            A[d][0][d] = t + m2 * A[d][0][d] + t0 * T0[d][d];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = d + 1; c < 3; ++c) {
                // entry (c, d) holds du_c/dx_d, entry (d, c) du_d/dx_c
                const T v  = M0 * (A[d][0][c] + A[c][0][d]);
                A[d][0][c] = v + t0 * T0[c][d];
                A[c][0][d] = v + t0 * T0[d][c];
            }
    }
    if constexpr (NGP == 8) {
        // B-bar keeps only the element-mean trace: the trace energy of the
        // higher modes drops by K, which leaves lambda - K = -2 mu / 3
        const T third = (T) (1.0 / 3.0);
        const T Lh = g.bbar ? (T) (-2.0 / 3.0) * M0 : L0;
        const T L1 = Lh * third, M1 = M0 * third, L2 = L1 * third, M2 = M1 * third;
        // xi: d/dy mode 1, d/dz mode 1; eta: d/dx mode 1, d/dz mode 2; zeta: d/dx mode 2, d/dy mode 2
        modal_pair<T, 1, 2>(A[1][1], A[2][1], L1, M1);
        modal_pair<T, 0, 2>(A[0][1], A[2][2], L1, M1);
        modal_pair<T, 0, 1>(A[0][2], A[1][2], L1, M1);
        modal_single<T, 0>(A[0][3], L2, M2);  // eta zeta
        modal_single<T, 1>(A[1][3], L2, M2);  // xi zeta
        modal_single<T, 2>(A[2][3], L2, M2);  // xi eta
        // transposed maps: the Walsh transform of the mode sums
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            T b[8];
            b[0] = (T) 0;
            b[1] = -A[0][0][c] * k[0];
            b[2] = -A[1][0][c] * k[1];
            b[4] = -A[2][0][c] * k[2];
            b[3] = A[0][1][c] * k[0] + A[1][1][c] * k[1];
            b[5] = A[0][2][c] * k[0] + A[2][1][c] * k[2];
            b[6] = A[1][2][c] * k[1] + A[2][2][c] * k[2];
            b[7] = -(A[0][3][c] * k[0] + A[1][3][c] * k[1] + A[2][3][c] * k[2]);
            walsh8(b);
#pragma unroll
            for (int n = 0; n < 8; ++n)
                f[n][c] = b[n];
        }
        return;
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                A[d][1][c] = A[d][2][c] = A[d][3][c] = A[d][0][c];
    }
    // transposed edge differences (one Gauss point)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int n = 0; n < 8; ++n)
            f[n][c] = (T) 0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const T dx = A[0][p + 2 * q][c] * k[0], dy = A[1][p + 2 * q][c] * k[1],
                        dz = A[2][p + 2 * q][c] * k[2];
                f[1 + 2 * p + 4 * q][c] += dx;
                f[2 * p + 4 * q][c] -= dx;
                f[p + 2 + 4 * q][c] += dy;
                f[p + 4 * q][c] -= dy;
                f[p + 2 * q + 4][c] += dz;
                f[p + 2 * q][c] -= dz;
            }
    }
}

// isotropic conduction in modal form: every mode of every derivative is its
// own block, weighted by NGP / 3^deg
template <typename T, int NGP, bool CUBIC = false>
__device__ __forceinline__ void element_forces_iso1_modal(const T (&u)[8][1], const Geom &g, T kw, const T (&T0)[3],
                                                          T (&f)[8][1])
{
    const T kc   = (T) g.ilx;
    const T k[3] = {CUBIC ? (T) 1 : (T) g.ilx, CUBIC ? (T) 1 : (T) g.ily, CUBIC ? (T) 1 : (T) g.ilz};
    const T c0   = CUBIC ? (T) NGP / (T) 16 * kw * kc * kc : (T) NGP / (T) 16 * kw;
    const T t0   = CUBIC ? (T) NGP * (T) 0.25 * kc : (T) NGP * (T) 0.25;
    if constexpr (NGP == 8) {
        // Walsh coefficients -> modes (see walsh8), modal scaling, mode sums -> Walsh
        T w[8];
#pragma unroll
        for (int n = 0; n < 8; ++n)
            w[n] = u[n][0];
        walsh8(w);
        const T c1 = c0 * (T) (1.0 / 3.0), c2 = c1 * (T) (1.0 / 3.0);
        T       b[8];
        b[0]        = (T) 0;
        b[1]        = -(c0 * (-w[1] * k[0]) + t0 * T0[0]) * k[0];
        b[2]        = -(c0 * (-w[2] * k[1]) + t0 * T0[1]) * k[1];
        b[4]        = -(c0 * (-w[4] * k[2]) + t0 * T0[2]) * k[2];
        b[3]        = c1 * w[3] * (k[0] * k[0] + k[1] * k[1]);
        b[5]        = c1 * w[5] * (k[0] * k[0] + k[2] * k[2]);
        b[6]        = c1 * w[6] * (k[1] * k[1] + k[2] * k[2]);
        b[7]        = c2 * w[7] * (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        walsh8(b);
#pragma unroll
        for (int n = 0; n < 8; ++n)
            f[n][0] = b[n];
        return;
    }
    T A[3][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            A[0][p + 2 * q] = (u[1 + 2 * p + 4 * q][0] - u[2 * p + 4 * q][0]) * k[0];
            A[1][p + 2 * q] = (u[p + 2 + 4 * q][0] - u[p + 4 * q][0]) * k[1];
            A[2][p + 2 * q] = (u[p + 2 * q + 4][0] - u[p + 2 * q][0]) * k[2];
        }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (NGP == 8) {
            const T c1 = c0 * (T) (1.0 / 3.0), c2 = c1 * (T) (1.0 / 3.0);
            hadamard4(A[d][0], A[d][1], A[d][2], A[d][3]);
            A[d][0] = c0 * A[d][0] + t0 * T0[d];
            A[d][1] *= c1;
            A[d][2] *= c1;
            A[d][3] *= c2;
            hadamard4_T(A[d][0], A[d][1], A[d][2], A[d][3]);
        } else {
            const T a = c0 * ((A[d][0] + A[d][2]) + (A[d][1] + A[d][3])) + t0 * T0[d];
            A[d][0] = A[d][1] = A[d][2] = A[d][3] = a;
        }
    }
#pragma unroll
    for (int n = 0; n < 8; ++n)
        f[n][0] = (T) 0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const T dx = A[0][p + 2 * q] * k[0], dy = A[1][p + 2 * q] * k[1], dz = A[2][p + 2 * q] * k[2];
            f[1 + 2 * p + 4 * q][0] += dx;
            f[2 * p + 4 * q][0] -= dx;
            f[p + 2 + 4 * q][0] += dy;
            f[p + 4 * q][0] -= dy;
            f[p + 2 * q + 4][0] += dz;
            f[p + 2 * q][0] -= dz;
        }
}

// lane l <- lane l-1 inside each row of 16 lanes (DPP row_shr:1); row starts get 0
// (bound_ctrl set: the first lane of each row reads 0 from the DPP source
// itself, so no zeroed destination has to be prepared)
__device__ __forceinline__ double dpp_row_shr1(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x111, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x111, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float dpp_row_shr1(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x111, 0xf, 0xf, true));
}

// ---------------------------------------------------------------------------
// element-local pass (one thread per element, or per J2 state slot for the
// state commit): homogenized stress, fields,
// committing the J2 state.  Reads nodal u directly (no relative gather).
// ---------------------------------------------------------------------------
// ELEM_SENS: per-element integrals of eps_a^T dC/dp eps_b for the isotropic
// moduli p (bulk K, shear G; thermal: conductivity), eps_a from (u, g0) and
// eps_b from (ub, g0b) -- the sensitivity of v . Sigma_bar (times the volume)
// when (u, g0) solves for g0 and (ub, g0b) for the macro strain v.
enum { ELEM_HOMOG = 0, ELEM_FIELDS = 1, ELEM_COMMIT = 2, ELEM_SENS = 3 };

template <typename T>
struct ElemArgs {
    int             nx, ny, nz;
    int             ntz;           // z tiles of the state layout (J2State)
    Geom            geom;
    const uint8_t  *ms;
    int             ms16;          // phase ids are 16-bit (phase_id)
    const Phase<T> *phases;
    J2State         js;
    const T        *u;
    const T        *g0;
    double         *partial;       // HOMOG: [NS][gridDim]
    double         *strain_elem;   // FIELDS (nullable)
    double         *stress_elem;
    double         *strain_gp;
    double         *stress_gp;
    double         *plastic_flag;  // FIELDS (nullable): element mean of the PseudoPlastic flags, else 0
    const T        *vk;            // per-voxel moduli (nullable): K (thermal: k) ...
    const T        *vg;            // ... and G, replacing the phase table's
    const T        *ub;            // SENS: second displacement field and gradient
    const T        *g0b;
    double         *sens;          // SENS: [N][H == 1 ? 1 : 2]
    Slots           sl;            // RVE slots (u, ms, phases, g0, partial at the slot's offset)
};

// isotropic stress with per-voxel moduli: K, G (mechanics, Mandel) or k (thermal)
template <typename T, int H>
__device__ __forceinline__ void voxel_iso_stress(T K, T G, const T (&e)[H == 1 ? 3 : 6], T (&s)[H == 1 ? 3 : 6])
{
    if constexpr (H == 1) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            s[i] = K * e[i];
    } else {
        const T lam = K - (T) (2.0 / 3.0) * G, b1 = lam * (e[0] + e[1] + e[2]), b2 = (T) 2 * G;
#pragma unroll
        for (int i = 0; i < 3; ++i)
            s[i] = b1 + b2 * e[i];
#pragma unroll
        for (int i = 3; i < 6; ++i)
            s[i] = b2 * e[i];
    }
}

// The state commit walks the slots of the J2 state layout (J2State) and
// evaluates every slot (an element with two slots twice, with the same
// result); every other mode walks the elements and reads their canonical slots.
template <typename T, int H, int NGP, int MODE, int NS = (H == 1 ? 3 : 6)>
__global__ __launch_bounds__(256) void k_element_pass(ElemArgs<T> a)
{
    __shared__ double red[NS * 4];
    {
        const int sb = slot_of(a.sl, false);
        if (sb < 0)
            return;
        if (a.sl.map) {  // batched launch: this slot's buffers
            a.u       = slot_ptr(a.u, sb * a.sl.field);
            a.g0      = slot_ptr(a.g0, (int64_t) sb * 9);
            a.ms      = a.ms + sb * a.sl.ms;
            a.phases  = a.phases + sb * a.sl.ph;
            a.partial = slot_ptr(a.partial, sb * a.sl.part);
        }
    }
    const int64_t     tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    double            sacc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k)
        sacc[k] = 0.0;
    int     x = 0, y = 0, z = 0;
    int64_t slot   = tid;
    int64_t canon  = tid;  // ELEM_COMMIT: the canonical slot of the slot's element
    bool    active = false;
    if constexpr (MODE == ELEM_COMMIT) {
        bool canonical;
        active = tid < a.js.nslot;
        if (active) {
            state_slot_element(tid, a.ny, a.nz, a.ntz, x, y, z, canonical);
            canon = (((int64_t) x * a.ny + y) * a.ntz + z / kStateTile) * 16 + z % kStateTile + 1;
        }
    } else {
        active = tid < (int64_t) a.nx * a.ny * a.nz;
        if (active) {
            z                = (int) (tid % a.nz);
            const int64_t xy = tid / a.nz;
            y                = (int) (xy % a.ny);
            x                = (int) (xy / a.ny);
            slot             = (xy * a.ntz + z / kStateTile) * 16 + z % kStateTile + 1;
        }
    }
    if (active) {
        const int64_t el = ((int64_t) x * a.ny + y) * a.nz + z;
        const int x1 = x + 1 == a.nx ? 0 : x + 1, y1 = y + 1 == a.ny ? 0 : y + 1, z1 = z + 1 == a.nz ? 0 : z + 1;
        T         u[8][H];
#pragma unroll
        for (int cr = 0; cr < 8; ++cr) {
            const int64_t n = ((int64_t) ((cr & 1) ? x1 : x) * a.ny + ((cr & 2) ? y1 : y)) * a.nz + ((cr & 4) ? z1 : z);
#pragma unroll
            for (int c = 0; c < H; ++c)
                u[cr][c] = a.u[n * H + c];
        }
        const Phase<T> &P = a.phases[phase_id(a.ms, a.ms16, el)];
        T               Gx[2][2][H], Gy[2][2][H], Gz[2][2][H];
        grad_factors<T, H, NGP>(u, a.geom, Gx, Gy, Gz);
        constexpr bool CAN_BBAR = H == 3 && NGP == 8;
        const bool     bbar     = CAN_BBAR && a.geom.bbar;
        T              trc      = (T) 0;
        if constexpr (CAN_BBAR)
            trc = bbar_mean_trace<T, NS>(Gx, Gy, Gz);
        T g0[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k)
            g0[k] = a.g0[k];
        const T vK = a.vk ? a.vk[el] : (T) 0, vG = (a.vk && H == 3) ? a.vg[el] : (T) 0;
        if constexpr (MODE == ELEM_SENS) {
            // second field at the same corners
            T ub[8][H];
#pragma unroll
            for (int cr = 0; cr < 8; ++cr) {
                const int64_t n =
                    ((int64_t) ((cr & 1) ? x1 : x) * a.ny + ((cr & 2) ? y1 : y)) * a.nz + ((cr & 4) ? z1 : z);
#pragma unroll
                for (int c = 0; c < H; ++c)
                    ub[cr][c] = a.ub[n * H + c];
            }
            T Hx[2][2][H], Hy[2][2][H], Hz[2][2][H];
            grad_factors<T, H, NGP>(ub, a.geom, Hx, Hy, Hz);
            T trcb = (T) 0;
            if constexpr (CAN_BBAR)
                trcb = bbar_mean_trace<T, NS>(Hx, Hy, Hz);
            T g0b[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k)
                g0b[k] = a.g0b[k];
            double sk = 0.0, sg = 0.0;
#pragma unroll
            for (int gp = 0; gp < NGP; ++gp) {
                const int i = gp & 1, j = (gp >> 1) & 1, k = gp >> 2;
                T         ga[H][3], gb[H][3];
#pragma unroll
                for (int c = 0; c < H; ++c) {
                    ga[c][0] = Gx[j][k][c], ga[c][1] = Gy[i][k][c], ga[c][2] = Gz[i][j][c];
                    gb[c][0] = Hx[j][k][c], gb[c][1] = Hy[i][k][c], gb[c][2] = Hz[i][j][c];
                }
                if constexpr (CAN_BBAR)
                    if (bbar) {
                        bbar_project_grad<T, NS>(ga, trc);
                        bbar_project_grad<T, NS>(gb, trcb);
                    }
                T ea[NS], eb[NS];
                grad_to_strain<T, H>(ga, g0, ea);
                grad_to_strain<T, H>(gb, g0b, eb);
                double dot = 0.0;
#pragma unroll
                for (int c = 0; c < NS; ++c)
                    dot += (double) ea[c] * (double) eb[c];
                if constexpr (H == 1) {
                    sk += dot;  // dC/dk = I
                } else {
                    const double tra = (double) ea[0] + ea[1] + ea[2], trb = (double) eb[0] + eb[1] + eb[2];
                    sk += tra * trb;                             // dC/dK = m m^T
                    sg += 2.0 * dot - (2.0 / 3.0) * tra * trb;  // dC/dG = 2 I - 2/3 m m^T
                }
            }
            const double w = a.geom.w;
            if constexpr (H == 1) {
                a.sens[el] = w * sk;
            } else {
                a.sens[2 * el]     = w * sk;
                a.sens[2 * el + 1] = w * sg;
            }
            return;
        }
        // every evaluation of a homogenized-stress or fields pass is one
        // get_sigma call of the reference (canonical slot: once per element)
        PointModel<T, H, NGP, true, H == 3 && NS == 6> pm;
        if constexpr (MODE != ELEM_COMMIT && NS != 9)
            pm.init(P, a.js, slot, true);
        double eavg[NS], savg[NS], flag = 0.0;
        bool   pseudo = false;
        if constexpr (MODE == ELEM_FIELDS && H == 3 && NS == 6)
            pseudo = a.plastic_flag && !a.vk && (P.kind == MODEL_PSEUDO_PLASTIC || P.kind == MODEL_PSEUDO_NL);
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
            if constexpr (CAN_BBAR)
                if (bbar)
                    bbar_project_grad<T, NS>(gr, trc);
            T e[NS], s[NS];
            grad_to_strain<T, H, NS>(gr, g0, e);
            if constexpr (MODE == ELEM_COMMIT) {
                if constexpr (H == 3 && NS == 6 && sizeof(T) == 8)
                    if (P.kind == MODEL_J2_VISCO_LIN || P.kind == MODEL_J2_VISCO_NL) {
                        // updateInternalVariables (J2Plasticity.h:56-61): the plastic
                        // strain of the last evaluation (at this u, g0), psi and
                        // psi_bar of the step from the element's canonical slot
                        const int64_t ns = a.js.nslot, cs = canon;
                        double        st[13], ed[6], sd[6], epo[6], gamma, n[6], pd[9];
#pragma unroll
                        for (int c = 0; c < 13; ++c)
                            st[c] = a.js.st[((int64_t) c * NGP + gp) * ns + slot];
#pragma unroll
                        for (int c = 0; c < 6; ++c)
                            ed[c] = (double) e[c];
#pragma unroll
                        for (int c = 0; c < 9; ++c)
                            pd[c] = (double) P.p[c];
                        visco_point(P.kind, pd, ed, st, sd, epo, gamma, n);
#pragma unroll
                        for (int c = 0; c < 6; ++c)
                            a.js.st[((int64_t) c * NGP + gp) * ns + slot] = epo[c];
#pragma unroll
                        for (int c = 0; c < 7; ++c)
                            a.js.st[((int64_t) (6 + c) * NGP + gp) * ns + slot] =
                                a.js.st[((int64_t) (13 + c) * NGP + gp) * ns + cs];
                    }
                if (P.kind == MODEL_J2_NEW) {
                    double ed[6], epin[6], sd[6], epo[6], qo, pd[4];
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        ed[c]   = (double) e[c];
                        epin[c] = a.js.st[((int64_t) c * NGP + gp) * a.js.nslot + slot];
                    }
                    const double qin = a.js.st[((int64_t) 6 * NGP + gp) * a.js.nslot + slot];
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        pd[c] = (double) P.p[c];
                    j2_point(pd, ed, epin, qin, sd, epo, qo);
#pragma unroll
                    for (int c = 0; c < 6; ++c)
                        a.js.st[((int64_t) c * NGP + gp) * a.js.nslot + slot] = epo[c];
                    a.js.st[((int64_t) 6 * NGP + gp) * a.js.nslot + slot] = qo;
                }
            } else {
                if constexpr (NS == 9)
                    material_point_ns<T, H, NS>(P, e, s);
                else if (a.vk)
                    voxel_iso_stress<T, H>(vK, vG, e, s);
                else
                    pm.eval(gp, e, s);
#pragma unroll
                for (int c = 0; c < NS; ++c) {
                    eavg[c] += (double) e[c] / NGP;
                    savg[c] += (double) s[c] / NGP;
                }
                if constexpr (MODE == ELEM_FIELDS) {
                    if constexpr (H == 3 && NS == 6)
                        if (pseudo)
                            flag += pseudo_flag<T>(P, e);
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
            if (a.plastic_flag)
                a.plastic_flag[el] = pseudo ? flag / NGP : 0.0;
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
