// Residual assembly kernel (Solver::compute_residual_basic, solver.h:228-270,
// and the linear CG matvec of solverCG.h:94-105) for the periodic voxel grid.
//
// The kernel marches along x.  A workgroup of 256 threads owns a 16x16 (y,z)
// tile of elements (one element per thread, lanes ordered z-fastest so one
// wave holds 4 rows of 16) and emits the 15x15 nodes whose four surrounding
// elements of the current plane lie in the tile.  Per x plane:
//   1. the 17x17 node plane x+1 is staged in LDS;
//   2. each thread gathers its 8 corners from LDS and computes its element;
//   3. corner forces are combined along z with DPP row_shr (in registers),
//      along y with a 16-lane shuffle (LDS only for the wave-boundary row),
//      and along x by a register carry into the next plane.
// Two barriers per plane; every node is produced by exactly one thread in a
// fixed summation order, so results are deterministic and atomic-free.
#pragma once
#include "element_kernels.h"

namespace fans {

enum { ASM_RESID = 0, ASM_CGLIN = 1 };
#ifndef FANS_ASM_MATVEC_WAVES
#define FANS_ASM_MATVEC_WAVES 3
#endif
constexpr int kAsmMatvecWaves = FANS_ASM_MATVEC_WAVES;
constexpr int TILE = 16;  // element tile edge (y and z)
constexpr int NOUT = TILE - 1;

template <typename T>
struct AsmArgs {
    int             nx, ny, nz;
    int             cx;        // x planes per block
    int             nty, ntz;  // number of tiles in y, z
    Geom            geom;
    const uint8_t  *ms;        // phase ids [x][y][z]
    int             ms16;      // 16-bit phase ids (phase_id)
    const Phase<T> *phases;
    int             nphases;
    J2State         js;
    // ASM_RESID: in = u, out = r; norm/dot partials
    // ASM_CGLIN: in = d, out = rnew = K d; partial <d, rnew>
    const T        *in;
    T              *out;
    const T        *dotv;      // ASM_RESID optional: partial <out, dotv>
    const T        *g0;        // macroscopic gradient (n_str), device
    double         *partial;   // [2][gridDim] : slot 0 = norm/dot, slot 1 = dot
    int             measure;   // 0 L1, 1 L2, 2 Linf  (ASM_RESID)
    Slots           sl;        // RVE slots (done flags: finished slots return at entry)
    const T        *vk;        // per-voxel isotropic moduli (nullable, ISO only):
    const T        *vg;        // bulk K and shear G, or conductivity k in vk
};

template <typename T, int H>
__device__ __forceinline__ T shfl_up16(T v)
{
    return __shfl_up(v, 16, 64);
}

// VAR: ASM_GENERAL = per-Gauss-point constitutive path (any model);
// ASM_ISO = every phase LinearElasticIsotropic (H=3) or LinearThermalIsotropic
// (H=1), modal element arithmetic; ASM_ISO_CUBIC = the same on cubic voxels;
// ASM_LARGE = the per-Gauss-point path with finite strain (F and P, n_str 9);
// ASM_GENERAL_PNL = ASM_GENERAL with power-law PseudoPlastic phases present.
// ASM_GENERAL_VISCO = ASM_GENERAL with J2ViscoPlastic phases present (their
// per-call psi / psi_bar update, see PointModel).
enum { ASM_GENERAL = 0, ASM_ISO = 1, ASM_ISO_CUBIC = 2, ASM_LARGE = 3, ASM_GENERAL_PNL = 4, ASM_GENERAL_VISCO = 5 };

// The per-Gauss-point FP64 residual lands just above 256 registers (one wave
// per SIMD) unless asked for two waves, which it reaches without spilling; the
// finite-strain, power-law and visco-plastic residuals and the FP64
// per-Gauss-point linear matvec would spill there, so they keep one.  The FP64
// modal matvec takes 132 registers (three waves per SIMD); squeezed into 128
// for four it was 2 % slower at 256^3.
template <typename T, int MODE, int VAR>
constexpr int kAsmWaves()
{
    if (VAR == ASM_LARGE || VAR == ASM_GENERAL_PNL || VAR == ASM_GENERAL_VISCO ||
        (VAR == ASM_GENERAL && MODE == ASM_CGLIN && sizeof(T) == 8))
        return 1;
    if ((VAR == ASM_ISO || VAR == ASM_ISO_CUBIC) && MODE == ASM_CGLIN)
        return sizeof(T) == 8 ? kAsmMatvecWaves : 6;
    return 2;
}
template <typename T, int H, int NGP, int MODE, int VAR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kAsmWaves<T, MODE, VAR>()))) void k_assemble(AsmArgs<T> a)
{
    constexpr bool ISO = VAR == ASM_ISO || VAR == ASM_ISO_CUBIC, CUBIC = VAR == ASM_ISO_CUBIC;
    const int sb = slot_of(a.sl);
    if (sb < 0)
        return;
    if (a.sl.map) {  // batched launch: this slot's buffers
        a.in      = a.in + sb * a.sl.field;
        a.out     = a.out + sb * a.sl.field;
        a.dotv    = slot_ptr(a.dotv, sb * a.sl.field);
        a.g0      = slot_ptr(a.g0, (int64_t) sb * 9);
        a.ms      = a.ms + sb * a.sl.ms;
        a.phases  = a.phases + sb * a.sl.ph;
        a.partial = slot_ptr(a.partial, sb * a.sl.part);
        a.vk      = slot_ptr(a.vk, sb * (a.sl.field / H));
        a.vg      = slot_ptr(a.vg, sb * (a.sl.field / H));
    }
    constexpr int NL = TILE + 1;  // node tile edge
    constexpr int NS = nstr_of<H, VAR == ASM_LARGE>();
    __shared__ T      Un[2][NL * NL * H];
    __shared__ T      Yb[4][TILE * 2 * H];  // last row of each wave for the y exchange
    __shared__ double red[8];

    const int t = threadIdx.x, ey = t / TILE, ez = t - ey * TILE;
    const int wave = t >> 6, lane = t & 63;
    int       bid  = xcd_tile_index(blockIdx.x, gridDim.x);
    const int tz   = bid % a.ntz;
    bid /= a.ntz;
    const int     ty = bid % a.nty;
    const int     bx = bid / a.nty;
    const int     Y0 = ty * NOUT, Z0 = tz * NOUT, X0 = bx * a.cx;
    const int     X1 = min(X0 + a.cx, a.nx);
    const int     ny = a.ny, nz = a.nz, nx = a.nx;
    const int64_t plane = (int64_t) ny * nz;

    // periodic node index of tile position v (several wraps on grids below 16)
    auto wrapy = [&](int v) { return ((Y0 - 1 + v) % ny + ny) % ny; };
    auto wrapz = [&](int v) { return ((Z0 - 1 + v) % nz + nz) % nz; };
    const int  my_y   = Y0 - 1 + ey, my_z = Z0 - 1 + ez;
    const bool own_yz = ey >= 1 && ez >= 1 && my_y < ny && my_z < nz;
    // node-plane slots of this thread: its own node (ey, ez) and, for 33 threads
    // spread over the waves, one of the extra row/column 16 nodes
    const int  ext    = (t % 7 == 0 && t / 7 < 2 * NL - 1) ? t / 7 : -1;
    const int  e_a    = ext < 0 ? 0 : (ext < NL ? TILE : ext - NL);
    const int  e_b    = ext < 0 ? 0 : (ext < NL ? ext : TILE);
    const int64_t off_own = (int64_t) wrapy(ey) * nz + wrapz(ez);
    const int64_t off_ext = (int64_t) wrapy(e_a) * nz + wrapz(e_b);
    // J2 state slot of this thread's element within an x plane of the state
    // layout: the tile's 16 slots are its 16 element columns in z
    static_assert(NOUT == kStateTile && TILE == 16, "the J2 state layout follows the stencil tiles");
    const int64_t splane = (int64_t) ny * a.ntz * 16;
    const int64_t sl_own = ((int64_t) wrapy(ey) * a.ntz + tz) * 16 + ez;
    // elements next to an owned node; the others (rows of the last tile past
    // the grid edge when 15 does not divide it) only feed unowned sums
    const bool el_needed = Y0 - 1 + max(ey, 1) < ny && Z0 - 1 + max(ez, 1) < nz;

    T      carry[H];
    double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
    for (int c = 0; c < H; ++c)
        carry[c] = (T) 0;
    T g0[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
        g0[i] = (MODE == ASM_RESID) ? a.g0[i] : (T) 0;
    const T w = (T) a.geom.w;
    // isotropic moduli of the first 256 phases, staged once per block (ids
    // beyond read the phase table)
    __shared__ T PL[256][2];
    if constexpr (ISO) {
        for (int p = t; p < min(a.nphases, 256); p += blockDim.x) {
            PL[p][0] = a.phases[p].p[0];
            PL[p][1] = H == 3 ? a.phases[p].p[1] : (T) 0;
        }
    }
    // per-Gauss-point residual: model kind of the first 256 phases, staged once
    // per block
    constexpr bool PM = !ISO && VAR != ASM_LARGE && MODE == ASM_RESID;
    __shared__ int PK[PM ? 256 : 1];
    if constexpr (PM)
        for (int p = t; p < min(a.nphases, 256); p += blockDim.x)
            PK[p] = a.phases[p].kind;
    auto kind_of = [&](int ph) { return ph < 256 ? PK[ph] : a.phases[ph].kind; };

    // node plane x+1: fetch (global) then commit (LDS); the phase id of the
    // next element plane is loaded one plane ahead.  Addresses are a
    // per-thread base (loop invariant) plus a per-plane offset that is the
    // same for the whole workgroup (scalar arithmetic).
    const int64_t planeH = plane * H;
    const T *const in_own = a.in + off_own * H, *const in_ext = a.in + off_ext * H;
    T    pf_own[H], pf_ext[H];
    auto fetch_plane = [&](int xp) {
        const int     xg = xp < 0 ? xp + nx : (xp >= nx ? xp - nx : xp);
        const int64_t po = xg * planeH;
#pragma unroll
        for (int c = 0; c < H; ++c)
            pf_own[c] = in_own[po + c];
        if (ext >= 0)
#pragma unroll
            for (int c = 0; c < H; ++c)
                pf_ext[c] = in_ext[po + c];
    };
    auto commit_plane = [&](int buf) {
#pragma unroll
        for (int c = 0; c < H; ++c)
            Un[buf][(ey * NL + ez) * H + c] = pf_own[c];
        if (ext >= 0)
#pragma unroll
            for (int c = 0; c < H; ++c)
                Un[buf][(e_a * NL + e_b) * H + c] = pf_ext[c];
    };
    const uint8_t *const ms_own = a.ms + (a.ms16 ? 2 : 1) * off_own;
    auto phase_of = [&](int xe) {
        const int xg = xe < 0 ? xe + nx : xe;
        return phase_id(ms_own, a.ms16, xg * plane);
    };

    fetch_plane(X0 - 1);
    commit_plane(0);
    fetch_plane(X0);
    int                   ph_next = phase_of(X0 - 1);
    PointModel<T, H, NGP, VAR == ASM_GENERAL_PNL || VAR == ASM_GENERAL_VISCO, VAR == ASM_GENERAL_VISCO> pm;
    bool pm_pre = false;  // point 0 of the coming element requested
    for (int i = 0; X0 - 1 + i < X1; ++i) {
        const int xe = X0 - 1 + i;
        commit_plane((i + 1) & 1);
        const int ph = ph_next;
        if (xe + 1 < X1)
            ph_next = phase_of(xe + 1);
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
        T dv[H];  // d at node plane xe (CG mode)
#pragma unroll
        for (int c = 0; c < H; ++c)
            dv[c] = u[0][c];
        const int     xg = xe < 0 ? xe + nx : xe;
        const int64_t el = xg * plane + off_own;
        T             f[8][H];
        if constexpr (ISO) {
            if constexpr (H == 3) {
                T lam = ph < 256 ? PL[ph][0] : a.phases[ph].p[0], mu = ph < 256 ? PL[ph][1] : a.phases[ph].p[1];
                if (a.vk) {
                    mu  = a.vg[el];
                    lam = a.vk[el] - (T) (2.0 / 3.0) * mu;
                }
                const T lw = lam * w, mw = mu * w;
                T       T0[3][3];
                if constexpr (MODE == ASM_RESID) {
                    const T r = (T) kSqrtHalf, tr = g0[0] + g0[1] + g0[2];
                    const T E[3][3] = {{g0[0], r * g0[3], r * g0[4]}, {r * g0[3], g0[1], r * g0[5]},
                                       {r * g0[4], r * g0[5], g0[2]}};
#pragma unroll
                    for (int p = 0; p < 3; ++p)
#pragma unroll
                        for (int q = 0; q < 3; ++q)
                            T0[p][q] = (T) 2 * mw * E[p][q] + (p == q ? lw * tr : (T) 0);
                } else {
#pragma unroll
                    for (int p = 0; p < 3; ++p)
#pragma unroll
                        for (int q = 0; q < 3; ++q)
                            T0[p][q] = (T) 0;
                }
                element_forces_iso3_modal<T, NGP, CUBIC>(reinterpret_cast<const T(&)[8][3]>(u), a.geom, lw, mw, T0,
                                            reinterpret_cast<T(&)[8][3]>(f));
            } else {
                const T kw    = (a.vk ? a.vk[el] : (ph < 256 ? PL[ph][0] : a.phases[ph].p[0])) * w;
                const T T0[3] = {kw * g0[0], kw * g0[1], kw * g0[2]};
                element_forces_iso1_modal<T, NGP, CUBIC>(reinterpret_cast<const T(&)[8][1]>(u), a.geom, kw, T0,
                                            reinterpret_cast<T(&)[8][1]>(f));
            }
        } else {
            const Phase<T> &P = a.phases[ph];
            // the linear matvec sees linear models only: no state to prefetch
            if constexpr (NS == 9) {
                auto mat = [&](int gp, const T(&e)[NS], T(&s)[NS]) {
                    material_point_ns<T, H, NS>(P, e, s);
                };
                element_forces<T, H, NGP, NS>(u, a.geom, g0, mat, f);
            } else if constexpr (MODE == ASM_CGLIN) {
                auto mat = [&](int gp, const T(&e)[NS], T(&s)[NS]) { linear_point<T, H>(P, e, s); };
                element_forces<T, H, NGP, NS>(u, a.geom, g0, mat, f);
            } else {
                // the next element of this thread's row is xe + 1 (phase ph_next),
                // 0 <= xe + 1 < X1 <= nx: no wrap (xe = -1 itself wraps to nx - 1)
                // an element feeding no owned node reads no state (elastic stand-in)
                const double *sn = nullptr;
                if (xe + 1 < X1 && el_needed)
                    sn = decltype(pm)::state_of(kind_of(ph_next), a.js, (int64_t) (xe + 1) * splane + sl_own);
                // owned elements of this block's x range are evaluated here and
                // nowhere else in the pass: they carry the per-call state updates
                pm.begin(P, el_needed ? kind_of(ph) : (int) MODEL_ELASTIC_ISO, a.js, (int64_t) xg * splane + sl_own, sn,
                         pm_pre, own_yz && xe >= X0);
                pm_pre = true;
                auto mat = [&](int gp, const T(&e)[NS], T(&s)[NS]) { pm.eval(gp, e, s); };
                element_forces<T, H, NGP, NS>(u, a.geom, g0, mat, f);
            }
        }
        // the loads of the plane after next fly during the exchange and the
        // stores (the element's registers are free again here)
        if (xe + 2 <= X1)
            fetch_plane(xe + 2);
        // z: node (ey,ez) <- az=0 corners of element ez + az=1 corners of element ez-1
        T P2[2][2][H];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < H; ++c)
                P2[q & 1][q >> 1][c] = f[q][c] + dpp_row_shr1(f[q + 4][c]);
        // y: node (ey,ez) <- ay=0 of own row + ay=1 of row ey-1 (lane-16, or previous wave)
        if (lane >= 48)
#pragma unroll
            for (int ax = 0; ax < 2; ++ax)
#pragma unroll
                for (int c = 0; c < H; ++c)
                    Yb[wave][(ez * 2 + ax) * H + c] = P2[ax][1][c];
        T up[2][H];
#pragma unroll
        for (int ax = 0; ax < 2; ++ax)
#pragma unroll
            for (int c = 0; c < H; ++c)
                up[ax][c] = shfl_up16<T, H>(P2[ax][1][c]);
        __syncthreads();
        if (lane < 16 && wave > 0)
#pragma unroll
            for (int ax = 0; ax < 2; ++ax)
#pragma unroll
                for (int c = 0; c < H; ++c)
                    up[ax][c] = Yb[wave - 1][(ez * 2 + ax) * H + c];
        if (i >= 1 && own_yz) {
            const int64_t po = xe * planeH;
#pragma unroll
            for (int c = 0; c < H; ++c) {
                const T res            = carry[c] + P2[0][0][c] + up[0][c];
                a.out[off_own * H + po + c] = res;
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
                        acc1 += (double) res * (double) a.dotv[off_own * H + po + c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < H; ++c)
            carry[c] = P2[1][0][c] + up[1][c];
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

}  // namespace fans
