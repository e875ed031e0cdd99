// libfans_gpu.so: MI355X (gfx950) implementation of the FANS per-iteration
// solver path behind the C ABI of include/fans_gpu.h.
//
// Host side of one load case (the reference's MaterialManager + Solver pair):
// device buffers, FFT plans, Green operator, and the solver loops of
// include/solverCG.h (linear fast path with device-resident scalars; nonlinear
// path with the secant line search), include/solverFP.h and the mixed
// strain/stress controller of include/mixedBCs.h.
#include "fans_gpu.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "assemble_kernels.h"
#include "common.h"
#include "element_kernels.h"
#include "fft_kernels.h"
#include "fft_reg_kernels.h"
#include "green_kernels.h"
#include "vec_kernels.h"

using namespace fans;

namespace {

constexpr int kThreads = 256;

// kernel families for the timing report
enum {
    KF_FFT_Z_FWD = 0,
    KF_FFT_Y_FWD,
    KF_FFT_X_GREEN,
    KF_FFT_Y_INV,
    KF_FFT_Z_INV,
    KF_ASSEMBLE_LIN,
    KF_ASSEMBLE_RES,
    KF_CG_UPDATE,
    KF_REDUCE,
    KF_VECTOR,
    KF_ELEMENT,
    KF_CG_FFT_Z_FWD,  // CG update fused into the forward z pass
    KF_COUNT
};
const char *kFamilyNames[KF_COUNT] = {"fft_z_r2c",    "fft_y_fwd",    "fft_x_green",  "fft_y_inv",
                                      "fft_z_c2r",    "assemble_lin", "assemble_res", "cg_update",
                                      "reduce",       "vector",       "element_pass", "cg_update_fft_z_r2c"};

int nparams(int kind)
{
    switch (kind) {
    case FANS_MODEL_THERMAL_ISO: return 1;
    case FANS_MODEL_THERMAL_TRI: return 6;
    case FANS_MODEL_ELASTIC_ISO: return 2;
    case FANS_MODEL_ELASTIC_TRI: return 21;
    case FANS_MODEL_J2_NEW: return 4;
    case FANS_MODEL_PSEUDO_PLASTIC: return 4;
    case FANS_MODEL_PSEUDO_PLASTIC_NL: return 5;
    case FANS_MODEL_SVK: return 2;
    case FANS_MODEL_NEO_HOOKE: return 2;
    case FANS_MODEL_J2_VISCO_LIN: return 7;
    case FANS_MODEL_J2_VISCO_NL: return 9;
    default: return -1;
    }
}
bool model_visco(int kind) { return kind == FANS_MODEL_J2_VISCO_LIN || kind == FANS_MODEL_J2_VISCO_NL; }
bool model_linear(int kind) { return kind >= 1 && kind <= 4; }
bool model_large(int kind) { return kind == FANS_MODEL_SVK || kind == FANS_MODEL_NEO_HOOKE; }

// ---- host twiddles (long double, then rounded) ------------------------------
template <typename T>
std::vector<cplx<T>> twiddles(int n)
{
    std::vector<cplx<T>> w(n);
    const long double    pi = 3.14159265358979323846264338327950288L;
    for (int k = 0; k < n; ++k) {
        long double th = -2.0L * pi * (long double) k / (long double) n;
        w[k]           = cplx<T>{(T) cosl(th), (T) sinl(th)};
    }
    return w;
}

bool make_len(int n, int tws, FftLen &L)
{
    L.n   = n;
    L.tws = tws;
    L.nst = 0;
    int m = n;
    for (int r : {5, 3})
        while (m % r == 0) {
            if (L.nst >= 14)
                return false;
            L.rad[L.nst++] = r;
            m /= r;
        }
    while (m % 8 == 0) {
        L.rad[L.nst++] = 8;
        m /= 8;
    }
    while (m % 4 == 0) {
        L.rad[L.nst++] = 4;
        m /= 4;
    }
    while (m % 2 == 0) {
        L.rad[L.nst++] = 2;
        m /= 2;
    }
    return m == 1;
}

// lines per block of the compile-time z kernels (0: generic kernels)
int zct_lines(int h, int M)
{
    switch (M) {
    case 128: return h == 1 ? zct_lpb<1, 128>() : zct_lpb<3, 128>();
    case 64: return h == 1 ? zct_lpb<1, 64>() : zct_lpb<3, 64>();
    case 48: return h == 1 ? zct_lpb<1, 48>() : zct_lpb<3, 48>();
    case 40: return h == 1 ? zct_lpb<1, 40>() : zct_lpb<3, 40>();
    case 32: return h == 1 ? zct_lpb<1, 32>() : zct_lpb<3, 32>();
    case 24: return h == 1 ? zct_lpb<1, 24>() : zct_lpb<3, 24>();
    case 16: return h == 1 ? zct_lpb<1, 16>() : zct_lpb<3, 16>();
    default: return 0;
    }
}

template <typename T>
struct Fft3D {
    int           nx = 0, ny = 0, nz = 0, h = 0, M = 0, Mc = 0;
    FftLen        Lz{}, Ly{}, Lx{};
    cplx<T>      *twz = nullptr, *twy = nullptr, *twx = nullptr;
    int           lpb = 1, bc = 8, bk = 2, bcx = 8;
    size_t        smz = 0, smy = 0, smx = 0, smx_plain = 0;
    // LDS budget of one workgroup of the generic (LDS-tiled) passes; it sets
    // their lines / columns per workgroup (FANS_OPT_FFT_LDS_KIB)
    int           lds_kib = 40;

    // largest power of two <= min(limit, cap), at least 1
    static int pow2_cols(int limit, int cap)
    {
        int v = 1;
        while (v * 2 <= limit && v * 2 <= cap)
            v *= 2;
        return v;
    }
    int64_t       nblk_z = 0;
    bool          zct    = false;  // compile-time z kernels in use
    bool          ready  = false;

    hipError_t init(int nx_, int ny_, int nz_, int h_)
    {
        hipError_t e = plan(nx_, ny_, nz_, h_);
        if (e != hipSuccess)
            return e;
        ready   = false;
        auto up = [](cplx<T> *&dst, const std::vector<cplx<T>> &v) {
            hipError_t e = hipMalloc(&dst, sizeof(cplx<T>) * v.size());
            if (e != hipSuccess)
                return e;
            return hipMemcpy(dst, v.data(), sizeof(cplx<T>) * v.size(), hipMemcpyHostToDevice);
        };
        if ((e = up(twz, twiddles<T>(nz))) != hipSuccess)
            return e;
        if ((e = up(twy, twiddles<T>(ny))) != hipSuccess)
            return e;
        if ((e = up(twx, twiddles<T>(nx))) != hipSuccess)
            return e;
        ready = true;
        return hipSuccess;
    }
    // launch shapes and length factorisations only (host arithmetic, no device
    // memory): the twiddle tables twz / twy / twx are the caller's to provide
    hipError_t plan(int nx_, int ny_, int nz_, int h_)
    {
        nx = nx_, ny = ny_, nz = nz_, h = h_;
        M  = nz / 2;
        Mc = M + 1;
        if (!make_len(M, 2, Lz) || !make_len(ny, 1, Ly) || !make_len(nx, 1, Lx))
            return hipErrorInvalidValue;
        // One padded LDS tile per workgroup, n * columns <= FFT_MAX_ELEMS so the
        // in-place stages keep every butterfly input in registers.
        const size_t cb = sizeof(cplx<T>);
        if ((int64_t) M * h > FFT_MAX_ELEMS || ny > FFT_MAX_ELEMS || (int64_t) nx * h > FFT_MAX_ELEMS)
            return hipErrorInvalidValue;
        // the default 40 KiB of LDS per workgroup keeps four workgroups
        // resident per CU
        const int lds_elems = (int) ((int64_t) lds_kib * 1024 / cb);
        const int lim       = std::min(FFT_MAX_ELEMS, lds_elems);
        lpb                 = std::max(1, std::min(16, lim / (M * h)));
        smz                 = cb * M * (lpb * h + 1);
        const int Kc        = Mc * h;
        bc                  = pow2_cols(lim / ny, std::min(Kc, 16));
        smy                 = cb * ny * (bc + 1);
        bk                  = std::max(1, std::min(Mc, lim / (nx * h)));
        smx                 = cb * nx * (bk * h + 1);
        bcx                 = pow2_cols(lim / nx, 16);
        smx_plain    = cb * nx * (bcx + 1);
        nblk_z       = ((int64_t) nx * ny + lpb - 1) / lpb;
        const int lct = zct_lines(h, M);
        zct           = lct > 0 && ((int64_t) nx * ny) % lct == 0;
        if (zct)
            nblk_z = (int64_t) nx * ny / lct;
        ready = true;
        return hipSuccess;
    }
    void release()
    {
        (void) hipFree(twz);
        (void) hipFree(twy);
        (void) hipFree(twx);
        twz = twy = twx = nullptr;
        ready           = false;
    }
};

// RVE slots of one batched context at most
constexpr int kMaxSlots = 4096;

// per-RVE device buffers of a context and the stride kind of each (Slots);
// g0 holds 9 entries per slot, the device scalars SC_COUNT / IS_COUNT
#define FANS_SLOT_BUFFERS(X)                                                                                   \
    X(double, u, field) X(double, u_prev, field) X(double, r, field) X(double, s, field) X(double, dA, field)   \
    X(double, dB, field) X(double, rnew, field) X(cplx<double>, spec64, spec) X(double, G64, green)            \
    X(float, G32, green) X(uint8_t, ms, ms) X(Phase<double>, dph64, ph) X(Phase<float>, dph32, ph)            \
    X(double, g0d, g0) X(float, g0f, g0) X(double, partial, part) X(double, partial_e, part) X(double, sc, sc) \
    X(int, is, is) X(double, hist, hist) X(float, r32, field) X(float, s32, field) X(float, dA32, field)       \
    X(float, dB32, field) X(float, rnew32, field) X(float, x32, field) X(cplx<float>, spec32, spec)
enum SlotKind { SK_field, SK_spec, SK_green, SK_ms, SK_ph, SK_g0, SK_part, SK_sc, SK_is, SK_hist };

// slot 0's address of every per-RVE buffer of a batched context
struct SlotBase {
#define X(T, name, kind) T *name = nullptr;
    FANS_SLOT_BUFFERS(X)
#undef X
    double *ring64[kDirRing]{};
    float  *ring32[kDirRing]{};
};

// host state of one RVE of a batched context while another slot is selected
#define FANS_SLOT_HOST(X)                                                                                      \
    X(have_ms) X(have_mat) X(ms16) X(ms_max) X(nphases) X(phase_kind) X(all_linear) X(has_j2) X(all_iso)        \
    X(has_pnl) X(has_visco) X(user_ref) X(voxel_moduli) X(ph64) X(ph32) X(gxq) X(gyq) X(iter) X(inner_iters)    \
    X(outer_iters) X(err_hist) X(solve_ms) X(kref) X(keref) X(g0) X(gvec) X(gvec_prev)
struct SlotHost {
    bool                       have_ms = false, have_mat = false, ms16 = false;
    int                        ms_max = -1, nphases = 0;
    std::vector<int>           phase_kind;
    bool                       all_linear = true, has_j2 = false, all_iso = false, has_pnl = false;
    bool                       has_visco = false, user_ref = false, voxel_moduli = false;
    std::vector<Phase<double>> ph64;
    std::vector<Phase<float>>  ph32;
    int                        gxq = 0, gyq = 0, iter = 0, inner_iters = 0, outer_iters = 0;
    std::vector<double>        err_hist;
    double                     solve_ms = 0;
    double                     kref[81]{}, keref[576]{}, g0[9]{}, gvec[9]{}, gvec_prev[9]{};
};

}  // namespace

struct fans_gpu_ctx {
    fans_gpu_desc d{};
    std::string   err;
    int           nx = 0, ny = 0, nz = 0, h = 0, nstr = 0, ngp = 0;
    int64_t       N = 0, Nh = 0, Nf = 0;
    hipStream_t   stream = nullptr;
    int           tol_type = 0;  // current error type (tangent switches it, solver.h:749)
    double        tol      = 0;

    // materials
    bool                        have_ms = false, have_mat = false;
    int                         ms_max  = -1;  // largest phase id of the microstructure
    int                         nphases = 0;
    std::vector<int>            phase_kind;
    bool                        all_linear = true, has_j2 = false, all_iso = false;
    bool                        has_pnl = false;  // a PseudoPlasticNonLinearHardening phase
    bool                        has_visco = false;  // a J2ViscoPlastic phase
    bool                        cubic = false;  // hx == hy == hz
    double                      kref[81]{};
    double                      keref[576]{};
    std::vector<Phase<double>>  ph64;
    std::vector<Phase<float>>   ph32;

    // device buffers
    uint8_t       *ms = nullptr;  // phase ids: bytes, or 16-bit ids when ms16
    bool           ms16 = false;
    int64_t        ms_bytes = 0;  // allocated bytes of ms
    int            nph_cap  = 0;  // allocated entries of dph64 / dph32
    Phase<double> *dph64 = nullptr;
    Phase<float>  *dph32 = nullptr;
    double        *u = nullptr, *u_prev = nullptr, *r = nullptr, *s = nullptr, *dA = nullptr, *dB = nullptr,
           *rnew = nullptr;
    cplx<double> *spec64 = nullptr;
    double       *G64 = nullptr, *keref_d = nullptr;
    float        *G32 = nullptr;
    int           gxq = 0, gyq = 0;  // stored kx / ky counts of the Green operator
    // per-voxel isotropic moduli (K, G; thermal: k in vk), replacing the phase table
    bool          voxel_moduli = false, user_ref = false;
    double       *vk64 = nullptr, *vg64 = nullptr;
    float        *vk32 = nullptr, *vg32 = nullptr;
    double       *ub_d = nullptr, *sens_d = nullptr;
    double       *js = nullptr;
    double       *g0d = nullptr;
    float        *g0f = nullptr;
    double       *partial = nullptr;
    int64_t       partial_len = 0;
    double       *partial_e   = nullptr;  // [nblk_z] of the linear CG's inverse z pass
    int64_t       partial_e_len = 0;
    double       *sc = nullptr;
    int          *is = nullptr;
    double       *hist = nullptr;
    // FP32 inner solve (mixed form)
    float        *r32 = nullptr, *s32 = nullptr, *dA32 = nullptr, *dB32 = nullptr, *rnew32 = nullptr, *x32 = nullptr;
    cplx<float>  *spec32 = nullptr;
    // linear CG direction rings: slots 0 and 1 are dA / dB (dA32 / dB32), the
    // others are allocated by the first linear CG solve
    double *ring64[kDirRing]{};
    float  *ring32[kDirRing]{};
    Fft3D<double> fft64;
    Fft3D<float>  fft32;

    // RVE slots (fans_gpu_create_batch): nb same-shape RVEs whose per-RVE
    // device buffers lie nb times at the strides of `strides`.  The buffer
    // fields above address the selected slot `cur`, `base` keeps slot 0's
    // addresses and `parked` the host state of the slots not selected.  A
    // batched solve points the fields at slot 0 and launches gz slots per
    // kernel (grid z) through the device map `slot_map`.
    int                   nb = 1, cur = 0, gz = 1;
    bool                  use_map  = false;
    int                  *slot_map = nullptr;
    int                   map_cap  = 0;
    std::vector<int>      map_h, map_slots;  // uploaded map and the slots it holds
    Slots                 strides{};
    SlotBase              base{};
    std::vector<SlotHost> parked;
    std::array<int, 7>    batch_sig{};

    // launch geometry
    int     cx = 16, nty = 0, ntz = 0, nblk_asm = 0, nblk_vec = 0, nblk_el = 0;
    int64_t nslot = 0;      // J2 state slots per component and Gauss point
    int     js_comps = 0;   // components of the allocated J2 state
    int     nblk_slot = 0;  // workgroups of the state commit (one thread per slot)
    double  g0[9]{};
    Geom    geom{};

    // mixed BC (mixedBCs.h)
    bool                mixed_active = false;
    std::vector<int>    idxE, idxF;
    std::vector<double> FE, PF, M;
    int                 nsteps = 0, step = 0;
    double              gvec[9]{}, gvec_prev[9]{};
    double              alpha_warm = 0.1;

    // results
    int                 iter = 0, inner_iters = 0, outer_iters = 0;
    std::vector<double> err_hist;
    double              solve_ms = 0;

    // timing
    bool                                                timing = false;
    int                                                 fft_chunks = 1;
    // the single-operator entry points run the mixed form's FP32 kernels
    bool                                                op_fp32 = false;
    // single-RVE mixed form with restarted inner solves (as batched contexts)
    bool                                                mixed_restart = false;
    // captured CG iterations (hipGraph), keyed by precision and the direction
    // buffer current at the first captured iteration
    bool                                                use_graphs = true;
    // key: the direction buffer and the slots per launch (gz)
    std::vector<std::pair<std::pair<const void *, int>, hipGraphExec_t>> graphs;
    std::array<int, 7>                                  graph_sig{};
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
    std::vector<hipEvent_t>                             ev_pool;
    double                                              fam_ms[KF_COUNT]{};
    int                                                 fam_n[KF_COUNT]{};
};

namespace {

// message of the last failed context-free call of this thread
thread_local char g_noalloc_err[256] = "";

int fail(fans_gpu_ctx *c, int code, const char *fmt, ...)
{
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                                      \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return fail(ctx, FANS_GPU_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                             \
    } while (0)

hipEvent_t get_event(fans_gpu_ctx *c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void) hipEventCreate(&e);
    return e;
}

// optional per-launch event bracketing (timing mode)
struct Timed {
    fans_gpu_ctx *c;
    int           fam;
    hipEvent_t    a = nullptr;
    Timed(fans_gpu_ctx *c_, int f) : c(c_), fam(f)
    {
        if (c->timing) {
            a = get_event(c);
            (void) hipEventRecord(a, c->stream);
        }
    }
    ~Timed()
    {
        if (c->timing) {
            hipEvent_t b = get_event(c);
            (void) hipEventRecord(b, c->stream);
            c->ev_used.push_back({fam, {a, b}});
        }
    }
};

void harvest_timing(fans_gpu_ctx *c)
{
    for (auto &p : c->ev_used) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            c->fam_ms[p.first] += ms;
            c->fam_n[p.first] += 1;
        }
        c->ev_pool.push_back(p.second.first);
        c->ev_pool.push_back(p.second.second);
    }
    c->ev_used.clear();
}

// Is the component-major reference element stiffness symmetric under the
// mirror of axis `ax` (0 = x, 1 = y)?  The mirror maps corner a to a ^ (1 << ax)
// and flips the sign of displacement component ax (mechanics); then
// G(-k_ax) = R G(k_ax) R and only half of that axis' frequencies need storing.
bool keref_mirror_symmetric(const double *ke, int h, int ax)
{
    const int ne = 8 * h;
    double    mx = 0.0;
    for (int k = 0; k < ne * ne; ++k)
        mx = std::max(mx, std::fabs(ke[k]));
    for (int i = 0; i < h; ++i)
        for (int a = 0; a < 8; ++a)
            for (int j = 0; j < h; ++j)
                for (int b = 0; b < 8; ++b) {
                    const double ri = (h == 3 && i == ax) ? -1.0 : 1.0, rj = (h == 3 && j == ax) ? -1.0 : 1.0;
                    const double v  = ke[(8 * i + a) * ne + 8 * j + b];
                    const double m  = ke[(8 * i + (a ^ (1 << ax))) * ne + 8 * j + (b ^ (1 << ax))];
                    if (std::fabs(v - ri * rj * m) > 1e-13 * mx)
                        return false;
                }
    return true;
}

// captured iterations hold kernel arguments (buffers, material flags, Green
// storage); forget them when any of those change
void drop_graphs(fans_gpu_ctx *c)
{
    for (auto &g : c->graphs)
        (void) hipGraphExecDestroy(g.second);
    c->graphs.clear();
}

// after new materials: keep the captured iterations when every by-value
// kernel argument they hold is unchanged (tables and the Green operator are
// read through fixed pointers), e.g. for a batch of RVEs on one context
void refresh_graphs(fans_gpu_ctx *c)
{
    const std::array<int, 7> sig = {c->all_iso, c->cubic, c->voxel_moduli, c->gxq, c->gyq, c->nphases, c->has_pnl};
    if (sig != c->graph_sig) {
        drop_graphs(c);
        c->graph_sig = sig;
    }
}

// Copies and clears ordered on the context's stream (non-blocking, so other
// host threads' legacy-stream calls neither wait for it nor disturb a capture
// in progress on it), then waited for.
hipError_t copy_on(fans_gpu_ctx *c, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}
hipError_t fill_on(fans_gpu_ctx *c, void *dst, size_t bytes)
{
    hipError_t e = hipMemsetAsync(dst, 0, bytes, c->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

// forget the timing records from index `from` on (launches that returned at entry)
void drop_timing(fans_gpu_ctx *c, size_t from)
{
    for (size_t k = from; k < c->ev_used.size(); ++k) {
        c->ev_pool.push_back(c->ev_used[k].second.first);
        c->ev_pool.push_back(c->ev_used[k].second.second);
    }
    if (from < c->ev_used.size())
        c->ev_used.resize(from);
}

int grid_vec(int64_t n) { return (int) std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 2048)); }

// ---------------------------------------------------------------------------
// RVE slots
// ---------------------------------------------------------------------------
int64_t slot_stride(const fans_gpu_ctx *c, SlotKind k)
{
    switch (k) {
    case SK_field: return c->strides.field;
    case SK_spec: return c->strides.spec;
    case SK_green: return c->strides.green;
    case SK_ms: return c->strides.ms;
    case SK_ph: return c->strides.ph;
    case SK_g0: return 9;
    case SK_part: return c->strides.part;
    case SK_sc: return SC_COUNT;
    case SK_is: return IS_COUNT;
    case SK_hist: return c->strides.hist;
    }
    return 0;
}

// point the buffer fields at slot b (batched contexts)
void set_views(fans_gpu_ctx *c, int b)
{
#define X(T, name, kind) c->name = c->base.name ? c->base.name + b * slot_stride(c, SK_##kind) : nullptr;
    FANS_SLOT_BUFFERS(X)
#undef X
    for (int m = 2; m < kDirRing; ++m) {
        c->ring64[m] = c->base.ring64[m] ? c->base.ring64[m] + b * c->strides.field : nullptr;
        c->ring32[m] = c->base.ring32[m] ? c->base.ring32[m] + b * c->strides.field : nullptr;
    }
    c->cur = b;
}

void swap_host(fans_gpu_ctx *c, SlotHost &h)
{
#define X(f) std::swap(c->f, h.f);
    FANS_SLOT_HOST(X)
#undef X
}

// make slot b the one that the single-RVE entry points address
void select_slot(fans_gpu_ctx *c, int b)
{
    if (b == c->cur)
        return;
    swap_host(c, c->parked[c->cur]);
    swap_host(c, c->parked[b]);
    set_views(c, b);
}

// slots of a launch: the selected slot alone (its buffers are the fields), or
// during a batched solve the slots of the device map; `gated`: finished slots
// return at entry
Slots launch_slots(const fans_gpu_ctx *c, bool gated)
{
    Slots s = c->strides;
    s.map   = c->use_map ? c->slot_map : nullptr;
    s.is    = gated ? c->is : nullptr;
    return s;
}

// launch the given slots from now on, one grid z plane each
int set_map(fans_gpu_ctx *c, const std::vector<int> &slots)
{
    const int gz = std::max<int>(1, (int) slots.size());
    c->map_h.assign(gz, -1);
    std::copy(slots.begin(), slots.end(), c->map_h.begin());
    c->map_slots = slots;
    HIPCHK(c, copy_on(c, c->slot_map, c->map_h.data(), sizeof(int) * gz, hipMemcpyHostToDevice));
    c->gz      = gz;
    c->use_map = true;
    return 0;
}

void clear_map(fans_gpu_ctx *c)
{
    c->gz      = 1;
    c->use_map = false;
    c->map_slots.clear();
}

// ---------------------------------------------------------------------------
// power-of-two lengths 16..256 use the register four-step kernels
// ---------------------------------------------------------------------------
bool reg_len(int n) { return n == 16 || n == 32 || n == 64 || n == 128 || n == 256; }
// mixed-radix lengths of the register y pass (24 x 4, 20 x 4, 12 x 4); the
// FP32 x+Green pass takes them too
bool reg_len_y(int n) { return n == 96 || n == 80 || n == 48; }

// columns per block: a power of two (>= 4) within ~72 KiB of LDS (two
// workgroups per CU) and at most MAXT threads; at 256^3 FP64 this gives 16
// columns (256-byte row segments) for the y pass and 4 frequencies for the x
// pass.  The FP32 y pass takes 512 threads (32 columns, 256-byte segments:
// 0.095 -> 0.085 ms at 256^3); the FP32 x pass stays at 256 (wider was slower).
constexpr int kRegLdsKiB = 72;
template <typename T, int N1, int N2, int MAXT = 256>
constexpr int reg_cols()
{
    constexpr int n   = N1 * N2;
    constexpr int lim = (int) (kRegLdsKiB * 1024 / (n * sizeof(cplx<T>)));
    int           c   = 4;
    while (c * 2 <= lim && c * 2 * N2 <= MAXT)
        c *= 2;
    return c;
}
template <typename T>
constexpr int strided_max_threads()
{
    return sizeof(T) == 4 ? 512 : 256;
}

template <typename T, bool INV, int N1, int N2>
void launch_strided_reg(fans_gpu_ctx *c, cplx<T> *spec, const cplx<T> *tw, int64_t sn, int c0, int c1,
                        int64_t ostride, int nouter, const int *done)
{
    constexpr int C   = reg_cols<T, N1, N2, strided_max_threads<T>()>();
    const int     nbx = (c1 - c0 + C - 1) / C;
    hipLaunchKernelGGL((k_fft_strided_reg<T, N1, N2, C, INV>), dim3(nbx * nouter, 1, c->gz), dim3(C * N2), 0,
                       c->stream, spec, tw, sn, c0, c1, ostride, nbx, launch_slots(c, done != nullptr));
}

template <typename T, bool INV>
void launch_strided(fans_gpu_ctx *c, const FftLen &L, const cplx<T> *tw, int bc, size_t smem, cplx<T> *spec,
                    int64_t sn, int c0, int c1, int64_t ostride, int nouter, const int *done)
{
    switch (reg_len(L.n) || reg_len_y(L.n) ? L.n : 0) {
    case 96: launch_strided_reg<T, INV, 24, 4>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 80: launch_strided_reg<T, INV, 20, 4>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 48: launch_strided_reg<T, INV, 12, 4>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 256: launch_strided_reg<T, INV, 16, 16>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 128: launch_strided_reg<T, INV, 16, 8>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 64: launch_strided_reg<T, INV, 8, 8>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 32: launch_strided_reg<T, INV, 8, 4>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    case 16: launch_strided_reg<T, INV, 4, 4>(c, spec, tw, sn, c0, c1, ostride, nouter, done); break;
    default:
        hipLaunchKernelGGL((k_fft_strided<T, INV>), dim3((c1 - c0 + bc - 1) / bc, nouter, c->gz), dim3(kThreads),
                           smem, c->stream, spec, L, tw, sn, c0, c1, ostride, bc, launch_slots(c, done != nullptr));
    }
}

// Split transposes (fourstep_split) for the x+Green pass: measured at 256^3
// FP64 they took the pass from 0.29 to 0.50 ms (twice the barriers outweigh the
// doubled residency), so the complex tile stays
constexpr bool kXGreenSplit = false;

// x+Green launches return their block count (one partial of the spectral
// <r, s> per block when `partial` is set)
template <typename T, int H, int N1, int N2>
int launch_x_green_reg(fans_gpu_ctx *c, cplx<T> *spec, const cplx<T> *tw, const GreenView<T> &gv, int kz_lo,
                       int kz_hi, double *partial, const int *done)
{
    constexpr int C   = reg_cols<T, N1, N2>();
    constexpr int KB  = H == 3 ? (C == 16 ? 5 : C / 4) : C;
    const int     nbk = (kz_hi - kz_lo + KB - 1) / KB;
    // the complex transpose tile of the 256-point FP64 pass (70 KB) allows two
    // workgroups per CU; the split tile (35 KB) four
    constexpr bool SPLIT = kXGreenSplit && (int64_t) N1 * N2 * (C + 1) * sizeof(cplx<T>) > 48 * 1024;
    hipLaunchKernelGGL((k_fft_x_green_reg<T, H, N1, N2, KB, C, SPLIT>), dim3(nbk * c->ny, 1, c->gz), dim3(C * N2),
                       0, c->stream, spec, tw, gv, kz_lo, kz_hi, nbk, partial, launch_slots(c, done != nullptr));
    return nbk * c->ny;
}

template <typename T, int H>
int launch_x_green(fans_gpu_ctx *c, Fft3D<T> &P, cplx<T> *spec, const T *G, int kz_lo, int kz_hi, double *partial,
                   const int *done)
{
    const GreenView<T> gv{G, c->nx, c->ny, c->gxq, c->gyq, P.Mc};
    if constexpr (sizeof(T) == 4) {  // FP64 at 20 x 4 and 24 x 4 would spill (one wave)
        switch (P.Lx.n) {
        case 96: return launch_x_green_reg<T, H, 24, 4>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
        case 80: return launch_x_green_reg<T, H, 20, 4>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
        case 48: return launch_x_green_reg<T, H, 12, 4>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
        default: break;
        }
    }
    switch (reg_len(P.Lx.n) ? P.Lx.n : 0) {
    case 256: return launch_x_green_reg<T, H, 16, 16>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
    case 128: return launch_x_green_reg<T, H, 16, 8>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
    case 64: return launch_x_green_reg<T, H, 8, 8>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
    case 32: return launch_x_green_reg<T, H, 8, 4>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
    case 16: return launch_x_green_reg<T, H, 4, 4>(c, spec, P.twx, gv, kz_lo, kz_hi, partial, done);
    default: {
        dim3 g((kz_hi - kz_lo + P.bk - 1) / P.bk, c->ny, c->gz);
        hipLaunchKernelGGL((k_fft_x_green<T, H>), g, dim3(kThreads), P.smx, c->stream, spec, P.Lx, P.twx, gv, kz_lo,
                           kz_hi, P.bk, partial, launch_slots(c, done != nullptr));
        return (int) (g.x * g.y);
    }
    }
}

// blocks of one x+Green launch over all kz (the size of its partial array)
template <typename T>
int64_t x_green_blocks(const Fft3D<T> &P, int h, int ny)
{
    int kb = 0;
    if constexpr (sizeof(T) == 4)  // the same choice as launch_x_green
        switch (P.Lx.n) {
        case 96: kb = reg_cols<T, 24, 4>(); break;
        case 80: kb = reg_cols<T, 20, 4>(); break;
        case 48: kb = reg_cols<T, 12, 4>(); break;
        default: break;
        }
    switch (kb ? -1 : (reg_len(P.Lx.n) ? P.Lx.n : 0)) {
    case -1: break;
    case 256: kb = reg_cols<T, 16, 16>(); break;
    case 128: kb = reg_cols<T, 16, 8>(); break;
    case 64: kb = reg_cols<T, 8, 8>(); break;
    case 32: kb = reg_cols<T, 8, 4>(); break;
    case 16: kb = reg_cols<T, 4, 4>(); break;
    default: return (int64_t) ((P.Mc + P.bk - 1) / P.bk) * ny;
    }
    if (h == 3)
        kb = kb == 16 ? 5 : kb / 4;
    return (int64_t) ((P.Mc + kb - 1) / kb) * ny;
}

int reduce(fans_gpu_ctx *c, const double *partial, int nblk, int nq, int stage, const double *partial_b = nullptr,
           int nblk_b = 0);

// z passes: compile-time kernels where M = nz/2 is in {16, 24, 32, 40, 48, 64,
// 128} and the line count divides evenly, else the generic ones
template <typename T, int H, int PRO>
void launch_z_r2c(fans_gpu_ctx *c, Fft3D<T> &P, T *rin, cplx<T> *spec, const ProArgs<T> &pa, const int *done)
{
    const int64_t nlines = (int64_t) c->nx * c->ny;
    const Slots   sl     = launch_slots(c, done != nullptr);
#define ZF(MM)                                                                                                  \
    hipLaunchKernelGGL((k_fft_z_r2c_ct<T, H, MM, PRO>), dim3((unsigned) (nlines / zct_lpb<H, MM>()), 1, c->gz),  \
                       dim3(zct_threads<H, MM>()), 0, c->stream, rin, spec, P.twz, pa, sl)
    switch (P.zct ? P.M : 0) {
    case 128: ZF(128); break;
    case 64: ZF(64); break;
    case 48: ZF(48); break;
    case 40: ZF(40); break;
    case 32: ZF(32); break;
    case 24: ZF(24); break;
    case 16: ZF(16); break;
    default:
        hipLaunchKernelGGL((k_fft_z_r2c<T, H, PRO>), dim3(P.nblk_z, 1, c->gz), dim3(kThreads), P.smz, c->stream, rin,
                           spec, P.Lz, P.twz, c->nz, nlines, P.lpb, pa, sl);
    }
#undef ZF
}

template <typename T, int H, int EPI>
void launch_z_c2r(fans_gpu_ctx *c, Fft3D<T> &P, const cplx<T> *spec, const EpiArgs<T> &ep, const int *done)
{
    const int64_t nlines = (int64_t) c->nx * c->ny;
    const Slots   sl     = launch_slots(c, done != nullptr);
#define ZI(MM)                                                                                                  \
    hipLaunchKernelGGL((k_fft_z_c2r_ct<T, H, MM, EPI>), dim3((unsigned) (nlines / zct_lpb<H, MM>()), 1, c->gz),  \
                       dim3(zct_threads<H, MM>()), 0, c->stream, spec, P.twz, ep, sl)
    switch (P.zct ? P.M : 0) {
    case 128: ZI(128); break;
    case 64: ZI(64); break;
    case 48: ZI(48); break;
    case 40: ZI(40); break;
    case 32: ZI(32); break;
    case 24: ZI(24); break;
    case 16: ZI(16); break;
    default:
        hipLaunchKernelGGL((k_fft_z_c2r<T, H, EPI>), dim3(P.nblk_z, 1, c->gz), dim3(kThreads), P.smz, c->stream, spec,
                           P.Lz, P.twz, c->nz, nlines, P.lpb, ep, sl);
    }
#undef ZI
}

// ---------------------------------------------------------------------------
// FFT convolution: out = G * r  with the requested epilogue.  With `pro` the
// forward z pass first applies the previous CG residual update to r and the
// RED_ERR_NEXT reduction (error, stop test) follows it.
// ---------------------------------------------------------------------------
template <typename T>
int convolve_dev(fans_gpu_ctx *c, Fft3D<T> &P, T *rin, cplx<T> *spec, const T *Gp, int epi, EpiArgs<T> ep,
                 const int *done, const ProArgs<T> *pro = nullptr)
{
    const int h = c->h, Kc = P.Mc * h;
    {
        Timed            t(c, pro ? KF_CG_FFT_Z_FWD : KF_FFT_Z_FWD);
        const ProArgs<T> pa = pro ? *pro : ProArgs<T>{};
        if (h == 1) {
            if (pro)
                launch_z_r2c<T, 1, PRO_CG>(c, P, rin, spec, pa, done);
            else
                launch_z_r2c<T, 1, PRO_NONE>(c, P, rin, spec, pa, done);
        } else {
            if (pro)
                launch_z_r2c<T, 3, PRO_CG>(c, P, rin, spec, pa, done);
            else
                launch_z_r2c<T, 3, PRO_NONE>(c, P, rin, spec, pa, done);
        }
    }
    HIPCHK(c, hipGetLastError());
    if (pro) {
        const int rc = reduce(c, pro->partial, (int) P.nblk_z, 1, RED_ERR_NEXT);
        if (rc)
            return rc;
    }
    // The y / x+Green / y^-1 passes only couple values of equal kz, so they can
    // run on kz chunks one after the other; a chunk that fits the Infinity
    // Cache keeps the middle passes out of HBM.
    const int nch = std::max(1, std::min(c->fft_chunks, P.Mc));
    // EPI_CGD: delta = <r, s> from the spectrum (x+Green partials), then
    // beta (RED_DELTA) before the inverse z pass forms d = s + beta d_prev
    double *const xg_partial = c->partial;
    int           xg_blocks  = 0;
    for (int q = 0; q < nch; ++q) {
        const int kz_lo = (int) ((int64_t) P.Mc * q / nch), kz_hi = (int) ((int64_t) P.Mc * (q + 1) / nch);
        const int c0 = kz_lo * h, c1 = kz_hi * h;
        {
            Timed t(c, KF_FFT_Y_FWD);
            launch_strided<T, false>(c, P.Ly, P.twy, P.bc, P.smy, spec, (int64_t) Kc, c0, c1, (int64_t) c->ny * Kc,
                                     c->nx, done);
        }
        {
            Timed   t(c, KF_FFT_X_GREEN);
            double *xp = epi == EPI_CGD ? xg_partial + xg_blocks : nullptr;
            xg_blocks += h == 1 ? launch_x_green<T, 1>(c, P, spec, Gp, kz_lo, kz_hi, xp, done)
                                : launch_x_green<T, 3>(c, P, spec, Gp, kz_lo, kz_hi, xp, done);
        }
        {
            Timed t(c, KF_FFT_Y_INV);
            launch_strided<T, true>(c, P.Ly, P.twy, P.bc, P.smy, spec, (int64_t) Kc, c0, c1, (int64_t) c->ny * Kc,
                                    c->nx, done);
        }
    }
    if (epi == EPI_CGD) {
        HIPCHK(c, hipGetLastError());
        const int rc = reduce(c, xg_partial, xg_blocks, 1, RED_DELTA);
        if (rc)
            return rc;
    }
    {
        Timed t(c, KF_FFT_Z_INV);
        if (h == 1) {
            if (epi == EPI_STORE)
                launch_z_c2r<T, 1, EPI_STORE>(c, P, spec, ep, done);
            else if (epi == EPI_CG)
                launch_z_c2r<T, 1, EPI_CG>(c, P, spec, ep, done);
            else if (epi == EPI_CGD)
                launch_z_c2r<T, 1, EPI_CGD>(c, P, spec, ep, done);
            else
                launch_z_c2r<T, 1, EPI_FP>(c, P, spec, ep, done);
        } else {
            if (epi == EPI_STORE)
                launch_z_c2r<T, 3, EPI_STORE>(c, P, spec, ep, done);
            else if (epi == EPI_CG)
                launch_z_c2r<T, 3, EPI_CG>(c, P, spec, ep, done);
            else if (epi == EPI_CGD)
                launch_z_c2r<T, 3, EPI_CGD>(c, P, spec, ep, done);
            else
                launch_z_c2r<T, 3, EPI_FP>(c, P, spec, ep, done);
        }
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// assembly launches
// ---------------------------------------------------------------------------
template <typename T>
AsmArgs<T> asm_args(fans_gpu_ctx *c)
{
    AsmArgs<T> a{};
    a.nx      = c->nx;
    a.ny      = c->ny;
    a.nz      = c->nz;
    a.cx      = c->cx;
    a.nty     = c->nty;
    a.ntz     = c->ntz;
    a.geom    = c->geom;
    a.ms      = c->ms;
    a.ms16    = c->ms16;
    a.nphases = c->nphases;
    a.js      = J2State{c->js, c->nslot};
    a.partial = c->partial;
    a.measure = c->d.measure;
    a.sl      = launch_slots(c, false);
    if (c->voxel_moduli) {
        if constexpr (sizeof(T) == sizeof(double)) {
            a.vk = reinterpret_cast<const T *>(c->vk64);
            a.vg = reinterpret_cast<const T *>(c->vg64);
        } else {
            a.vk = reinterpret_cast<const T *>(c->vk32);
            a.vg = reinterpret_cast<const T *>(c->vg32);
        }
    }
    return a;
}

template <typename T, int MODE>
int launch_assemble(fans_gpu_ctx *c, AsmArgs<T> a)
{
    Timed t(c, MODE == ASM_CGLIN ? KF_ASSEMBLE_LIN : KF_ASSEMBLE_RES);
#define ASM(HH, GG)                                                                                             \
    do {                                                                                                        \
        if (c->all_iso && c->cubic)                                                                             \
            hipLaunchKernelGGL((k_assemble<T, HH, GG, MODE, ASM_ISO_CUBIC>), dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), \
                               0, c->stream, a);                                                                \
        else if (c->all_iso)                                                                                    \
            hipLaunchKernelGGL((k_assemble<T, HH, GG, MODE, ASM_ISO>), dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), 0,    \
                               c->stream, a);                                                                   \
        else if (MODE == ASM_RESID && HH == 3 && c->has_visco)                                                  \
            hipLaunchKernelGGL(                                                                                 \
                (k_assemble<T, HH, GG, MODE, MODE == ASM_RESID && HH == 3 ? ASM_GENERAL_VISCO : ASM_GENERAL>),   \
                dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), 0, c->stream, a);                                            \
        else if (MODE == ASM_RESID && c->has_pnl)                                                               \
            hipLaunchKernelGGL((k_assemble<T, HH, GG, MODE, MODE == ASM_RESID ? ASM_GENERAL_PNL : ASM_GENERAL>),   \
                               dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), 0, c->stream, a);                             \
        else                                                                                                    \
            hipLaunchKernelGGL((k_assemble<T, HH, GG, MODE, ASM_GENERAL>), dim3(c->nblk_asm, 1, c->gz), dim3(kThreads),   \
                               0, c->stream, a);                                                                \
    } while (0)
    if (c->nstr == 9) {
        // finite strain: only the residual (the models are nonlinear)
        if constexpr (MODE == ASM_RESID) {
            if (c->ngp == 8)
                hipLaunchKernelGGL((k_assemble<T, 3, 8, MODE, ASM_LARGE>), dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), 0,
                                   c->stream, a);
            else
                hipLaunchKernelGGL((k_assemble<T, 3, 1, MODE, ASM_LARGE>), dim3(c->nblk_asm, 1, c->gz), dim3(kThreads), 0,
                                   c->stream, a);
        } else {
            return fail(c, FANS_GPU_EINVAL, "linear matvec needs linear materials");
        }
    } else if (c->h == 1) {
        if (c->ngp == 8)
            ASM(1, 8);
        else
            ASM(1, 1);
    } else {
        if (c->ngp == 8)
            ASM(3, 8);
        else
            ASM(3, 1);
    }
#undef ASM
    HIPCHK(c, hipGetLastError());
    return 0;
}

// r = R(u) (material path with g0); partial slot 0 = norm(r), slot 1 = <r, dotv>
int residual64(fans_gpu_ctx *c, const double *uin, double *rout, const double *dotv)
{
    AsmArgs<double> a = asm_args<double>(c);
    a.phases          = c->dph64;
    a.in              = uin;
    a.out             = rout;
    a.dotv            = dotv;
    a.g0              = c->g0d;
    return launch_assemble<double, ASM_RESID>(c, a);
}

int reduce(fans_gpu_ctx *c, const double *partial, int nblk, int nq, int stage, const double *partial_b,
           int nblk_b)
{
    Timed t(c, KF_REDUCE);
    hipLaunchKernelGGL(k_reduce, dim3(1, 1, c->gz), dim3(kReduceThreads), 0, c->stream, partial, nblk, nq, stage, c->sc,
                       c->is, c->hist, partial_b, nblk_b, launch_slots(c, false));
    HIPCHK(c, hipGetLastError());
    return 0;
}

int read_scalars(fans_gpu_ctx *c, double *sc, int *is)
{
    if (sc)
        HIPCHK(c, hipMemcpyAsync(sc, c->sc, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, c->stream));
    if (is)
        HIPCHK(c, hipMemcpyAsync(is, c->is, sizeof(int) * IS_COUNT, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int init_scalars(fans_gpu_ctx *c, double tol, int err_type, int n_it)
{
    double sc[SC_COUNT] = {0};
    int    is[IS_COUNT] = {0};
    sc[SC_TOL]          = tol;
    sc[SC_DELTA]        = 1.0;
    is[IS_NIT]          = n_it;
    is[IS_ERRTYPE]      = err_type;
    is[IS_MEASURE]      = c->d.measure;
    HIPCHK(c, hipMemcpyAsync(c->sc, sc, sizeof(sc), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->is, is, sizeof(is), hipMemcpyHostToDevice, c->stream));
    return 0;
}

template <typename T>
int fill(fans_gpu_ctx *c, T *p, T v, int64_t n)
{
    hipLaunchKernelGGL((k_fill<T>), dim3(grid_vec(n)), dim3(kThreads), 0, c->stream, p, v, n);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int upload_g0(fans_gpu_ctx *c)
{
    float gf[9];
    for (int i = 0; i < 9; ++i)
        gf[i] = (float) c->g0[i];
    HIPCHK(c, hipMemcpyAsync(c->g0d, c->g0, sizeof(double) * 9, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->g0f, gf, sizeof(float) * 9, hipMemcpyHostToDevice, c->stream));
    return 0;
}

// Solver::get_homogenized_stress (solver.h:707-737)
int homog_stress(fans_gpu_ctx *c, double *out)
{
    ElemArgs<double> a{};
    a.nx      = c->nx;
    a.ny      = c->ny;
    a.nz      = c->nz;
    a.ntz     = c->ntz;
    a.geom    = c->geom;
    a.ms      = c->ms;
    a.ms16 = c->ms16;
    a.phases  = c->dph64;
    a.js      = J2State{c->js, c->nslot};
    a.u       = c->u;
    a.g0      = c->g0d;
    a.vk      = c->voxel_moduli ? c->vk64 : nullptr;
    a.vg      = c->voxel_moduli ? c->vg64 : nullptr;
    a.partial = c->partial;
    a.sl      = launch_slots(c, false);
    {
        Timed t(c, KF_ELEMENT);
#define EP(HH, GG)                                                                                              \
    hipLaunchKernelGGL((k_element_pass<double, HH, GG, ELEM_HOMOG>), dim3(c->nblk_el, 1, c->gz), dim3(kThreads), 0, \
                       c->stream, a)
        if (c->nstr == 9) {
            if (c->ngp == 8)
                hipLaunchKernelGGL((k_element_pass<double, 3, 8, ELEM_HOMOG, 9>), dim3(c->nblk_el, 1, c->gz),
                                   dim3(kThreads), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_element_pass<double, 3, 1, ELEM_HOMOG, 9>), dim3(c->nblk_el, 1, c->gz),
                                   dim3(kThreads), 0, c->stream, a);
        } else if (c->h == 1) {
            if (c->ngp == 8)
                EP(1, 8);
            else
                EP(1, 1);
        } else {
            if (c->ngp == 8)
                EP(3, 8);
            else
                EP(3, 1);
        }
#undef EP
    }
    HIPCHK(c, hipGetLastError());
    int rc = reduce(c, c->partial, c->nblk_el, c->nstr, RED_SUM);
    if (rc || !out)
        return rc;
    double sc[SC_COUNT];
    if ((rc = read_scalars(c, sc, nullptr)))
        return rc;
    for (int k = 0; k < c->nstr; ++k)
        out[k] = sc[SC_RES0 + k] / (double) c->N;
    return 0;
}

// commit the J2 state at the current u, g0 (J2PlasticityNew.h:37-41)
int commit_state(fans_gpu_ctx *c)
{
    if (!c->has_j2)
        return 0;
    ElemArgs<double> a{};
    a.nx     = c->nx;
    a.ny     = c->ny;
    a.nz     = c->nz;
    a.ntz    = c->ntz;
    a.geom   = c->geom;
    a.ms     = c->ms;
    a.ms16 = c->ms16;
    a.phases = c->dph64;
    a.js     = J2State{c->js, c->nslot};
    a.u      = c->u;
    a.g0     = c->g0d;
    a.vk      = c->voxel_moduli ? c->vk64 : nullptr;
    a.vg      = c->voxel_moduli ? c->vg64 : nullptr;
    Timed t(c, KF_ELEMENT);
    if (c->ngp == 8)
        hipLaunchKernelGGL((k_element_pass<double, 3, 8, ELEM_COMMIT>), dim3(c->nblk_slot), dim3(kThreads), 0,
                           c->stream, a);
    else
        hipLaunchKernelGGL((k_element_pass<double, 3, 1, ELEM_COMMIT>), dim3(c->nblk_slot), dim3(kThreads), 0,
                           c->stream, a);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// mixed boundary conditions (mixedBCs.h:150-227)
// ---------------------------------------------------------------------------
int mixed_update(fans_gpu_ctx *c)
{
    if (!c->mixed_active)
        return 0;
    double Pbar[9];
    int    rc = homog_stress(c, Pbar);
    if (rc)
        return rc;
    const int nF = (int) c->idxF.size();
    if (nF) {
        std::vector<double> rhs(nF), dE(nF, 0.0);
        for (int i = 0; i < nF; ++i)
            rhs[i] = c->PF[(size_t) c->step * nF + i] - Pbar[c->idxF[i]];
        for (int i = 0; i < nF; ++i)
            for (int j = 0; j < nF; ++j)
                dE[i] += c->M[i * nF + j] * rhs[j];
        for (int i = 0; i < nF; ++i)
            c->gvec[c->idxF[i]] += dE[i];
    }
    for (int i = 0; i < c->nstr; ++i)
        c->g0[i] = c->gvec[i];
    return upload_g0(c);
}

// symmetric pseudo-inverse of a small SPD block (Jacobi, host)
// inverse of a nonsingular, non-symmetric n x n block (n <= 9) by Gauss-Jordan
// elimination with partial pivoting: what the complete orthogonal
// decomposition pinv of mixedBCs.h:43 gives for such a block; nonzero if singular
int host_gauss_inverse(int n, const double *A, double *Ai)
{
    double a[9][18];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 2 * n; ++j)
            a[i][j] = j < n ? A[i * n + j] : (double) (j - n == i);
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(a[r][col]) > std::fabs(a[piv][col]))
                piv = r;
        if (a[piv][col] == 0.0)
            return -1;
        for (int j = 0; j < 2 * n; ++j)
            std::swap(a[col][j], a[piv][j]);
        const double d = a[col][col];
        for (int j = 0; j < 2 * n; ++j)
            a[col][j] /= d;
        for (int r = 0; r < n; ++r)
            if (r != col) {
                const double f = a[r][col];
                for (int j = 0; j < 2 * n; ++j)
                    a[r][j] -= f * a[col][j];
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            Ai[i * n + j] = a[i][n + j];
    return 0;
}

void host_sym_pinv(int n, const double *A, double *P)
{
    double a[81], V[81];
    std::memcpy(a, A, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            V[i * n + j] = i == j;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += a[i * n + j] * a[i * n + j];
                if (i != j)
                    off += a[i * n + j] * a[i * n + j];
            }
        if (off <= 1e-32 * tot || off == 0.0)
            break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (apq == 0.0)
                    continue;
                double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                double t     = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double cs = 1 / std::sqrt(t * t + 1), sn = t * cs;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = cs * akp - sn * akq;
                    a[k * n + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = cs * apk - sn * aqk;
                    a[q * n + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = cs * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + cs * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0;
            for (int k = 0; k < n; ++k)
                if (std::fabs(a[k * n + k]) > 1e-300)
                    acc += V[i * n + k] * V[j * n + k] / a[k * n + k];
            P[i * n + j] = acc;
        }
}

// ---------------------------------------------------------------------------
// solver loops
// ---------------------------------------------------------------------------

// Linear CG fast path (solverCG.h:84-117, islinear && !mixed); every scalar
// lives on the device, the host only polls the done flag once per chunk.
// Iteration j runs
//   z r2c [+ the residual update of iteration j-1: r -= alpha K d]
//   -> RED_ERR_NEXT (err; sets done) -> y, x+Green (delta = <r, s> by
//   Parseval), y^-1 -> RED_DELTA (beta from delta and deltamid)
//   -> z c2r (d_j = s + beta d_(j-1) with s = -G r, s itself not stored;
//      <d_j, K d_(j-1)> from the previous matvec's output)
//   -> matvec (K d_j, <d_j, K d_j>) -> RED_ALPHA (alpha_j, and deltamid of
//      iteration j+1 = <r - alpha K d, s> = delta - alpha (<d, K d> - beta
//      <d, K d_prev>), the reference's <r, s> of solverCG.h:87 in exact
//      arithmetic, K being symmetric)
// The directions go round a ring of kDirRing buffers (d_j in slot j % kDirRing)
// and the solution update u -= alpha_j d_j (solverCG.h:105) is applied for
// kDirRing iterations at a time, after the iteration that fills the last slot:
// one pass reading kDirRing directions instead of a read-modify-write of u per
// iteration.  After RED_ERR_NEXT sets `done` every later kernel returns at
// entry; r then holds the final residual and finish() applies the directions
// left in the ring.  ring[kDirRing - 1] must hold zeros on entry (d_(-1) = 0).
// integer scalars of every slot of a batched context (fields at slot 0)
int read_is_all(fans_gpu_ctx *c, std::vector<int> &is)
{
    is.resize((size_t) c->nb * IS_COUNT);
    HIPCHK(c, hipMemcpyAsync(is.data(), c->is, sizeof(int) * is.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Linear CG iterations (solverCG.h:84-117) until the stop test holds: of the
// selected RVE, or batched (c->use_map) of every slot of the map.  A batched
// loop runs its slots in lockstep, each stopping on its own test (its
// launches return at entry from then on); the finished slots leave the map
// once at most half of its slots still run.
// j0 > 0 (single RVE): continue a stopped loop at iteration j0 with a
// replaced residual r (the residual after iteration j0 - 1's update; the
// directions, K d_(j0-1) in rnew and the CG scalars are kept).
template <typename T>
int cg_linear_loop(fans_gpu_ctx *c, Fft3D<T> &P, T *r, T *const *ring, T *rnew, T *u, cplx<T> *spec, const T *Gp,
                   const Phase<T> *ph, int *iters_out, int j0 = 0)
{
    const int             *done  = c->is + IS_DONE;
    const bool             batch = c->use_map;
    const std::vector<int> slots0 = c->map_slots;
    int                    chunk = 4, rc;
    int                    is[IS_COUNT];
    std::vector<int>       isall;
    bool                   stop = false;
    // u -= alpha_m d_m for the ring slots 0 .. count-1 (iterations filled in
    // order); count < 0: each RVE slot's pending count; gated: finished RVE
    // slots skip it
    auto ring_update = [&](int count, bool gated) -> int {
        Timed      t(c, KF_CG_UPDATE);
        DirRing<T> dr{};
        for (int m = 0; m < kDirRing; ++m)
            dr.d[m] = ring[m];
        const Slots sl = launch_slots(c, true);
        if (count == kDirRing)
            hipLaunchKernelGGL((k_ring_update<T, true>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, u,
                               dr, c->sc + SC_ARING, count, c->Nh, sl, gated ? 1 : 0);
        else
            hipLaunchKernelGGL((k_ring_update<T, false>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, u,
                               dr, c->sc + SC_ARING, count, c->Nh, sl, gated ? 1 : 0);
        HIPCHK(c, hipGetLastError());
        return 0;
    };
    // the stop test after a run of iterations; batched, the finished slots
    // leave the map once the others fit a shallower grid
    auto poll = [&]() -> int {
        if (!batch) {
            if ((rc = read_scalars(c, nullptr, is)))
                return rc;
            stop = is[IS_DONE];
            return 0;
        }
        if ((rc = read_is_all(c, isall)))
            return rc;
        std::vector<int> act;
        for (int b : c->map_slots)
            if (!isall[(size_t) b * IS_COUNT + IS_DONE])
                act.push_back(b);
        stop = act.empty();
        // a new grid depth means a new captured graph: shrink by halves
        return (!stop && 2 * (int) act.size() <= c->gz) ? set_map(c, act) : 0;
    };
    // iterations 0 .. k-1 have run; those after the last full ring are pending
    auto finish = [&]() -> int {
        if (batch) {
            if (c->map_slots != slots0 && (rc = set_map(c, slots0)))
                return rc;
            return ring_update(-1, false);
        }
        const int k = is[IS_ITER];
        if (k % kDirRing) {
            const int e = ring_update(k % kDirRing, false);
            if (e)
                return e;
            // the flushed slots' alphas to zero: a continued loop's next full
            // ring update must not apply them again
            HIPCHK(c, hipMemsetAsync(c->sc + SC_ARING, 0, sizeof(double) * (k % kDirRing), c->stream));
        }
        if (iters_out)
            *iters_out = k;
        return 0;
    };
    // CG iteration j, with_pro: first the residual update of iteration j-1
    auto iteration = [&](int j, bool with_pro) -> int {
        T *const         dprev    = ring[(j + kDirRing - 1) % kDirRing];
        T *const         dnext    = ring[j % kDirRing];
        const ProArgs<T> pa{rnew, c->sc + SC_ALPHA, c->d.measure, c->partial};
        // delta and beta from the spectrum, then d = s + beta d_prev in the
        // inverse z pass, with <d, K d_prev> (read after the matvec); the first
        // iteration has d_prev = 0, and rnew holds no K d_prev yet
        EpiArgs<T> ep{};
        ep.partial = c->partial_e;
        ep.dprev   = dprev;
        ep.qprev   = j > 0 ? rnew : nullptr;
        ep.dnext   = dnext;
        ep.scal    = c->sc;  // SC_BETA == 0
        int e;
        if ((e = convolve_dev<T>(c, P, r, spec, Gp, EPI_CGD, ep, done, with_pro ? &pa : nullptr)))
            return e;
        // rnew = K d, <d, rnew>
        AsmArgs<T> a = asm_args<T>(c);
        a.phases     = ph;
        a.in         = dnext;
        a.out        = rnew;
        a.sl         = launch_slots(c, true);
        if ((e = launch_assemble<T, ASM_CGLIN>(c, a)))
            return e;
        // quantity 1 of the first iteration: the matvec's zero slot 1
        if ((e = reduce(c, c->partial, c->nblk_asm, 2, RED_ALPHA, j > 0 ? c->partial_e : nullptr,
                        (int) P.nblk_z)))
            return e;
        if (j % kDirRing == kDirRing - 1 && (e = ring_update(kDirRing, true)))
            return e;
        return 0;
    };
    // After iteration 0 the launches of iteration j depend on j % kDirRing
    // only, so kDirRing iterations are captured once into a hipGraph (per
    // direction buffer and grid depth) and replayed (one launch per replay
    // instead of ~10 per iteration).  Timing mode stays eager.
    static_assert(kDirRing == 8, "one captured graph holds one turn of the direction ring");
    if (c->use_graphs && !c->timing) {
        // eager up to the first iteration of a ring turn's graph (j = 1 mod 8)
        int j = j0;
        if ((rc = iteration(j++, false)))
            return rc;
        for (; j % kDirRing != 1; ++j)
            if ((rc = iteration(j, true)))
                return rc;
        if (j0 > 0) {
            if ((rc = poll()))
                return rc;
            if (stop)
                return finish();
        }
        for (;;) {
            const std::pair<const void *, int> key{(const void *) ring[0], c->gz};
            hipGraphExec_t                     exec = nullptr;
            for (auto &g : c->graphs)
                if (g.first == key)
                    exec = g.second;
            if (!exec) {
                hipGraph_t graph = nullptr;
                HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
                int crc = 0;
                for (int jj = 1; jj <= kDirRing && !crc; ++jj)
                    crc = iteration(jj, true);
                const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
                if (crc)
                    return crc;
                HIPCHK(c, ec);
                HIPCHK(c, hipGraphInstantiateWithFlags(&exec, graph, 0));
                (void) hipGraphDestroy(graph);
                c->graphs.push_back({key, exec});
            }
            HIPCHK(c, hipGraphLaunch(exec, c->stream));
            if ((rc = poll()))
                return rc;
            if (stop)
                break;
        }
        return finish();
    }
    for (int j = j0;;) {
        // timing mode checks `done` after every iteration so that launches that
        // returned at entry are not counted in the per-kernel durations
        const int n = c->timing ? 1 : chunk;
        for (int k = 0; k < n; ++k, ++j) {
            const size_t mark = c->ev_used.size();
            if ((rc = iteration(j, j > j0)))
                return rc;
            if (c->timing) {
                if ((rc = read_scalars(c, nullptr, is)))
                    return rc;
                // done was set by this iteration's RED_ERR_NEXT (the fused update
                // and that reduction did work) or before it (nothing did)
                if (is[IS_DONE])
                    drop_timing(c, j > j0 ? mark + 2 : mark);
            }
        }
        if ((rc = poll()))
            return rc;
        if (stop)
            break;
        chunk = std::min(chunk * 2, 16);
    }
    return finish();
}

int fetch_history(fans_gpu_ctx *c, int iters)
{
    c->err_hist.assign(iters + 1, 0.0);
    HIPCHK(c, hipMemcpyAsync(c->err_hist.data(), c->hist, sizeof(double) * (iters + 1), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// the direction ring of the linear CG in precision T (slots 0, 1 = dA, dB)
template <typename T>
int ensure_ring(fans_gpu_ctx *c, T *(&ring)[kDirRing], T *dA, T *dB)
{
    ring[0] = dA;
    ring[1] = dB;
    for (int m = 2; m < kDirRing; ++m)
        if (!ring[m])
            HIPCHK(c, hipMalloc(&ring[m], sizeof(T) * c->Nh));
    return 0;
}

int solve_cg_fp64(fans_gpu_ctx *c)
{
    int rc;
    if ((rc = init_scalars(c, c->tol, c->tol_type, c->d.n_it)))
        return rc;
    c->alpha_warm = 0.1;
    const bool linear = c->all_linear && !c->mixed_active;
    if (linear && (rc = ensure_ring<double>(c, c->ring64, c->dA, c->dB)))
        return rc;
    // d_(-1) = 0: dA for the nonlinear loop, the ring's last slot for the linear one
    if ((rc = fill<double>(c, c->s, 0.0, c->Nh)) ||
        (rc = fill<double>(c, linear ? c->ring64[kDirRing - 1] : c->dA, 0.0, c->Nh)))
        return rc;
    if ((rc = residual64(c, c->u, c->r, nullptr)))
        return rc;
    if ((rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)))
        return rc;
    if (linear) {
        int it = 0;
        if ((rc = cg_linear_loop<double>(c, c->fft64, c->r, c->ring64, c->rnew, c->u, c->spec64, c->G64, c->dph64,
                                         &it)))
            return rc;
        c->iter = it;
        return fetch_history(c, it);
    }
    // nonlinear CG with the secant line search (solverCG.h:84-160)
    int    is[IS_COUNT];
    double sc[SC_COUNT];
    if ((rc = read_scalars(c, nullptr, is)))
        return rc;
    double *d = c->dA;
    while (!is[IS_DONE]) {
        // deltamid = <r, s>
        hipLaunchKernelGGL((k_dot<double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->r, c->s, c->Nh,
                           c->d.measure, c->partial);
        if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_SUM)))
            return rc;
        if ((rc = read_scalars(c, sc, nullptr)))
            return rc;
        HIPCHK(c, hipMemcpyAsync(c->sc + SC_DELTAMID, &sc[SC_RES0], sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        if (c->d.precision == FANS_PREC_MIXED) {
            // mixed form of the nonlinear solve: the Green preconditioner s = -G r
            // in FP32 (FFTs and Green multiply); residuals, the J2 return map,
            // dot products and the line search stay FP64
            hipLaunchKernelGGL((k_convert<float, double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->r32,
                               c->r, c->Nh);
            EpiArgs<float> ep{};
            ep.out     = c->s32;
            ep.r       = c->r32;
            ep.partial = c->partial;
            if ((rc = convolve_dev<float>(c, c->fft32, c->r32, c->spec32, c->G32, EPI_CG, ep, nullptr)))
                return rc;
            hipLaunchKernelGGL((k_convert<double, float>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->s,
                               c->s32, c->Nh);
            hipLaunchKernelGGL((k_dot<double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->r, c->s, c->Nh,
                               c->d.measure, c->partial);
            if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_DELTA)))
                return rc;
        } else {
            EpiArgs<double> ep{};
            ep.out     = c->s;
            ep.r       = c->r;
            ep.partial = c->partial;
            if ((rc = convolve_dev<double>(c, c->fft64, c->r, c->spec64, c->G64, EPI_CG, ep, nullptr)))
                return rc;
            if ((rc = reduce(c, c->partial, (int) c->fft64.nblk_z, 1, RED_DELTA)))
                return rc;
        }
        // --- LineSearchSecant --- (d = s + beta d, rpd = <r, d> and the first
        // step u += alpha_warm d in one pass)
        const int    MaxIt = c->d.ls_max_iter;
        const double ltol  = c->d.ls_tol;
        double       lerr = 10.0, ap = 0.0, ac = c->alpha_warm;
        int          lit = 0;
        {
            Timed tv(c, KF_VECTOR);
            hipLaunchKernelGGL((k_ls_start<double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, d, c->s, c->r,
                               c->u, c->sc, ac, c->Nh, c->partial);
        }
        if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_SUM)) || (rc = read_scalars(c, sc, nullptr)))
            return rc;
        double rpd = sc[SC_RES0];
        if ((rc = mixed_update(c)) || (rc = residual64(c, c->u, c->rnew, d)))
            return rc;
        if ((rc = reduce(c, c->partial + c->nblk_asm, c->nblk_asm, 1, RED_SUM)) ||
            (rc = read_scalars(c, sc, nullptr)))
            return rc;
        double r1pd = sc[SC_RES0];
        while (lit < MaxIt && lerr > ltol) {
            double denom = r1pd - rpd;
            if (std::fabs(denom) < 1e-14 * (std::fabs(r1pd) + std::fabs(rpd)))
                break;
            double an = ac - r1pd * (ac - ap) / denom;
            if (an <= 0.0)
                an = 0.5 * (ap + ac);
            lerr = std::fabs(an - ac);
            hipLaunchKernelGGL((k_axpby<double, double>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream,
                               c->u, d, an - ac, 1.0, c->Nh, launch_slots(c, false));
            ap  = ac;
            rpd = r1pd;
            ac  = an;
            lit++;
            if ((rc = mixed_update(c)) || (rc = residual64(c, c->u, c->rnew, d)))
                return rc;
            if ((rc = reduce(c, c->partial + c->nblk_asm, c->nblk_asm, 1, RED_SUM)) ||
                (rc = read_scalars(c, sc, nullptr)))
                return rc;
            r1pd = sc[SC_RES0];
        }
        c->alpha_warm = (lit == MaxIt && lerr > ltol) ? 0.1 : ac;
        std::swap(c->r, c->rnew);
        // the last residual evaluation left norm(rnew) partials in slot 0
        if ((rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_NEXT)) || (rc = read_scalars(c, nullptr, is)))
            return rc;
    }
    c->iter = is[IS_ITER];
    return fetch_history(c, c->iter);
}

int solve_fp_fp64(fans_gpu_ctx *c)
{
    int rc;
    if ((rc = init_scalars(c, c->tol, c->tol_type, c->d.n_it)))
        return rc;
    if ((rc = residual64(c, c->u, c->r, nullptr)) || (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)))
        return rc;
    int is[IS_COUNT];
    if ((rc = read_scalars(c, nullptr, is)))
        return rc;
    const int *done = c->is + IS_DONE;
    while (!is[IS_DONE]) {
        const int chunk = c->mixed_active ? 1 : 8;
        for (int k = 0; k < chunk; ++k) {
            EpiArgs<double> ep{};
            ep.u = c->u;
            if ((rc = convolve_dev<double>(c, c->fft64, c->r, c->spec64, c->G64, EPI_FP, ep, done)))
                return rc;
            if ((rc = mixed_update(c)))
                return rc;
            AsmArgs<double> a = asm_args<double>(c);
            a.phases          = c->dph64;
            a.in              = c->u;
            a.out             = c->r;
            a.g0              = c->g0d;
            a.sl              = launch_slots(c, true);
            if ((rc = launch_assemble<double, ASM_RESID>(c, a)) ||
                (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_NEXT)))
                return rc;
        }
        if ((rc = read_scalars(c, nullptr, is)))
            return rc;
    }
    c->iter = is[IS_ITER];
    return fetch_history(c, c->iter);
}

// The mixed form's FP32 inner solves.
// Residual replacement (the default): one FP32 CG whose recursive residual is
// replaced by the true FP64 one, after u += x, whenever it has dropped by
// kReplaceEta against the last true error, or below kReplaceFloor times the
// error the outer test asks for; the directions and CG scalars carry over, so
// no Krylov information is lost.  Restarted inner solves (single RVE with
// FANS_OPT_MIXED_RESTART): each outer round starts a new CG from x = 0 to
// kRestartEta times the outer error, not below kRestartFloor times the
// target.  A restart loses the smooth modes' convergence, so at an equal
// final residual its homogenized stress is less accurate: the lower floor
// keeps it within 5e-9 of the FP64 solution on the 384-RVE batch.
constexpr double kReplaceFloor = 0.9, kReplaceEta = 1e-4;
constexpr double kRestartFloor = 0.5, kRestartEta = 1e-4;

// Mixed form with restarted inner solves: FP64 defect correction around FP32
// Green-preconditioned CG runs.
// outer: r = R(u) in FP64 (the true residual), stop when it meets the
// tolerance; inner: solve K x = r with FP32 operator/vectors and FP64 dots to a
// relative reduction eta (or the FP32 noise floor), then u += x.
int solve_cg_mixed_restart(fans_gpu_ctx *c)
{
    int rc;
    // outer residual
    if ((rc = init_scalars(c, c->tol, c->tol_type, c->d.n_it)))
        return rc;
    if ((rc = residual64(c, c->u, c->r, nullptr)) || (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)))
        return rc;
    double sc[SC_COUNT];
    int    is[IS_COUNT];
    if ((rc = read_scalars(c, sc, is)))
        return rc;
    std::vector<double> outer_hist{sc[SC_ERR]};
    double              err0  = sc[SC_ERR];
    int                 inner = 0, outer = 0;
    bool                done  = is[IS_DONE];
    const double        eta   = kRestartEta;
    // the absolute error the outer test asks for
    const double tol_abs = c->tol_type == FANS_ERR_ABSOLUTE ? c->tol : c->tol * err0;
    while (!done && inner < c->d.n_it) {
        // inner FP32 CG on K x = r, x = 0 (its own scalars/history)
        const double err_out = sc[SC_ERR];
        if ((rc = init_scalars(c, std::max(eta * err_out, kRestartFloor * tol_abs), FANS_ERR_ABSOLUTE,
                               c->d.n_it - inner)))
            return rc;
        if ((rc = ensure_ring<float>(c, c->ring32, c->dA32, c->dB32)))
            return rc;
        // x = 0 and d_(-1) = 0 in the ring's last slot
        hipLaunchKernelGGL(k_inner_start, dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->r32, c->x32,
                           c->ring32[kDirRing - 1], c->r, c->Nh, c->d.measure, c->partial, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
        if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_ERR_INIT)))
            return rc;
        int it = 0;
        if ((rc = cg_linear_loop<float>(c, c->fft32, c->r32, c->ring32, c->rnew32, c->x32, c->spec32, c->G32,
                                        c->dph32, &it)))
            return rc;
        inner += it;
        outer += 1;
        // u += x: the inner CG drove r(x) = R(u) + K x to zero (x -= alpha d)
        hipLaunchKernelGGL((k_axpby<double, float>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->u,
                           c->x32, 1.0, 1.0, c->Nh, launch_slots(c, false));
        // true residual
        if ((rc = residual64(c, c->u, c->r, nullptr)))
            return rc;
        if ((rc = init_scalars(c, c->tol, c->tol_type, 1 << 30)))
            return rc;
        if ((rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)) || (rc = read_scalars(c, sc, is)))
            return rc;
        outer_hist.push_back(sc[SC_ERR]);
        const double e = c->tol_type == FANS_ERR_ABSOLUTE ? sc[SC_ERR] : sc[SC_ERR] / err0;
        done           = !(e > c->tol);
        if (it == 0)
            break;  // FP32 floor reached without progress
    }
    c->iter        = inner;
    c->inner_iters = inner;
    c->outer_iters = outer;
    c->err_hist    = outer_hist;
    return 0;
}

// Mixed form (single RVE): FP64 residual replacement around one FP32
// Green-preconditioned CG (see kReplaceEta).
int solve_cg_mixed(fans_gpu_ctx *c)
{
    int rc;
    // nonlinear materials or mixed boundary conditions: the nonlinear CG with
    // an FP32 Green preconditioner (solve_cg_fp64 checks the precision)
    if (!(c->all_linear && !c->mixed_active))
        return solve_cg_fp64(c);
    if (c->mixed_restart)
        return solve_cg_mixed_restart(c);
    // outer residual
    if ((rc = init_scalars(c, c->tol, c->tol_type, c->d.n_it)))
        return rc;
    if ((rc = residual64(c, c->u, c->r, nullptr)) || (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)))
        return rc;
    double sc[SC_COUNT];
    int    is[IS_COUNT];
    if ((rc = read_scalars(c, sc, is)))
        return rc;
    std::vector<double> outer_hist{sc[SC_ERR]};
    const double        err0 = sc[SC_ERR];
    if (is[IS_DONE]) {
        c->iter = c->inner_iters = c->outer_iters = 0;
        c->err_hist = outer_hist;
        return 0;
    }
    // the absolute error the outer test asks for
    const double tol_abs = c->tol_type == FANS_ERR_ABSOLUTE ? c->tol : c->tol * err0;
    auto         phase_tol = [&](double err) { return std::max(kReplaceEta * err, kReplaceFloor * tol_abs); };
    if ((rc = init_scalars(c, phase_tol(err0), FANS_ERR_ABSOLUTE, c->d.n_it)) ||
        (rc = ensure_ring<float>(c, c->ring32, c->dA32, c->dB32)))
        return rc;
    // x = 0 and d_(-1) = 0 in the ring's last slot
    hipLaunchKernelGGL(k_inner_start, dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->r32, c->x32,
                       c->ring32[kDirRing - 1], c->r, c->Nh, c->d.measure, c->partial, launch_slots(c, false));
    HIPCHK(c, hipGetLastError());
    if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_ERR_INIT)))
        return rc;
    int k = 0, phases = 0;
    for (;;) {
        // one FP32 CG run until its residual meets the phase tolerance
        int kn = k;
        if ((rc = cg_linear_loop<float>(c, c->fft32, c->r32, c->ring32, c->rnew32, c->x32, c->spec32, c->G32,
                                        c->dph32, &kn, k)))
            return rc;
        const bool progress = kn > k;
        k                   = kn;
        phases += 1;
        // u += x, then the true residual replaces the FP32 recursive one
        hipLaunchKernelGGL((k_axpby<double, float>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->u,
                           c->x32, 1.0, 1.0, c->Nh, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
        if ((rc = residual64(c, c->u, c->r, nullptr)))
            return rc;
        // its error against the outer tolerance (recorded at iteration k)
        HIPCHK(c, hipMemcpyAsync(c->sc + SC_TOL, &tol_abs, sizeof(double), hipMemcpyHostToDevice, c->stream));
        if ((rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_SET)) || (rc = read_scalars(c, sc, is)))
            return rc;
        outer_hist.push_back(sc[SC_ERR]);
        const double e = c->tol_type == FANS_ERR_ABSOLUTE ? sc[SC_ERR] : sc[SC_ERR] / err0;
        if (!(e > c->tol) || k >= c->d.n_it || !progress)
            break;
        // continue the same CG from the replaced residual: r32 = (float) r,
        // x = 0 (directions, K d and scalars kept), the next phase tolerance
        const double tn = phase_tol(sc[SC_ERR]);
        const int    zero = 0;
        HIPCHK(c, hipMemcpyAsync(c->sc + SC_TOL, &tn, sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->is + IS_DONE, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_inner_start, dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->r32, c->x32,
                           (float *) nullptr, c->r, c->Nh, c->d.measure, c->partial, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));  // the host values copied above
    }
    c->iter        = k;
    c->inner_iters = k;
    c->outer_iters = phases;
    c->err_hist    = outer_hist;
    return 0;
}

// ---------------------------------------------------------------------------
// Batched solve: every RVE slot of a batched context, linear materials, CG in
// the FP64 or the mixed form.  Each slot takes the steps, and gets the
// results, of a solve of that RVE alone (same kernels, launch geometry and
// summation order; it stops on its own test), while each launch covers every
// slot still running.
// ---------------------------------------------------------------------------
struct BatchScalars {
    std::vector<double> sc;
    std::vector<int>    is;
};

int read_all_scalars(fans_gpu_ctx *c, BatchScalars &v)
{
    v.sc.resize((size_t) c->nb * SC_COUNT);
    v.is.resize((size_t) c->nb * IS_COUNT);
    HIPCHK(c, hipMemcpyAsync(v.sc.data(), c->sc, sizeof(double) * v.sc.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(v.is.data(), c->is, sizeof(int) * v.is.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int write_all_scalars(fans_gpu_ctx *c, const BatchScalars &v)
{
    HIPCHK(c, hipMemcpyAsync(c->sc, v.sc.data(), sizeof(double) * v.sc.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->is, v.is.data(), sizeof(int) * v.is.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// slot b's scalars as init_scalars sets a single context's
void init_slot_scalars(const fans_gpu_ctx *c, BatchScalars &v, int b, double tol, int err_type, int n_it)
{
    double *sc = &v.sc[(size_t) b * SC_COUNT];
    int    *is = &v.is[(size_t) b * IS_COUNT];
    std::fill(sc, sc + SC_COUNT, 0.0);
    std::fill(is, is + IS_COUNT, 0);
    sc[SC_TOL]     = tol;
    sc[SC_DELTA]   = 1.0;
    is[IS_NIT]     = n_it;
    is[IS_ERRTYPE] = err_type;
    is[IS_MEASURE] = c->d.measure;
}

// solve_cg_fp64's linear path for every slot
int solve_batch_cg64(fans_gpu_ctx *c, std::vector<SlotHost *> &H)
{
    const int        nb = c->nb;
    std::vector<int> all(nb);
    for (int b = 0; b < nb; ++b)
        all[b] = b;
    BatchScalars v;
    v.sc.assign((size_t) nb * SC_COUNT, 0.0);
    v.is.assign((size_t) nb * IS_COUNT, 0);
    for (int b = 0; b < nb; ++b)
        init_slot_scalars(c, v, b, c->tol, c->tol_type, c->d.n_it);
    int rc;
    if ((rc = write_all_scalars(c, v)) || (rc = ensure_ring<double>(c, c->ring64, c->dA, c->dB)))
        return rc;
    // d_(-1) = 0 in the ring's last slot, s = 0 (every slot at once)
    if ((rc = fill<double>(c, c->s, 0.0, c->Nh * nb)) ||
        (rc = fill<double>(c, c->ring64[kDirRing - 1], 0.0, c->Nh * nb)))
        return rc;
    if ((rc = set_map(c, all)) || (rc = residual64(c, c->u, c->r, nullptr)) ||
        (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)))
        return rc;
    if ((rc = cg_linear_loop<double>(c, c->fft64, c->r, c->ring64, c->rnew, c->u, c->spec64, c->G64, c->dph64,
                                     nullptr)))
        return rc;
    if ((rc = read_all_scalars(c, v)))
        return rc;
    std::vector<double> hist((size_t) nb * c->strides.hist);
    HIPCHK(c, hipMemcpyAsync(hist.data(), c->hist, sizeof(double) * hist.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < nb; ++b) {
        const int     it = v.is[(size_t) b * IS_COUNT + IS_ITER];
        const double *hb = &hist[(size_t) b * c->strides.hist];
        H[b]->iter        = it;
        H[b]->inner_iters = H[b]->outer_iters = 0;
        H[b]->err_hist.assign(hb, hb + it + 1);
    }
    return 0;
}

// solve_cg_mixed for every slot: one FP32 CG per slot with FP64 residual
// replacement.  A phase runs until every slot has met its phase tolerance;
// then each slot's true residual replaces its recursive one.  The slots that
// continue resume together at the largest iteration index J reached: a slot
// that stopped at k < J takes index J, its last direction moves to ring slot
// (J - 1) mod 8 and its iteration budget and error-history room follow.  The
// arithmetic of each slot, the order of its solution updates included, is
// the single-RVE solve's.
int solve_batch_mixed(fans_gpu_ctx *c, std::vector<SlotHost *> &H)
{
    const int        nb = c->nb;
    std::vector<int> all(nb);
    for (int b = 0; b < nb; ++b)
        all[b] = b;
    BatchScalars v;
    v.sc.assign((size_t) nb * SC_COUNT, 0.0);
    v.is.assign((size_t) nb * IS_COUNT, 0);
    for (int b = 0; b < nb; ++b)
        init_slot_scalars(c, v, b, c->tol, c->tol_type, c->d.n_it);
    int rc;
    if ((rc = write_all_scalars(c, v)) || (rc = ensure_ring<float>(c, c->ring32, c->dA32, c->dB32)))
        return rc;
    if ((rc = set_map(c, all)) || (rc = residual64(c, c->u, c->r, nullptr)) ||
        (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_INIT)) || (rc = read_all_scalars(c, v)))
        return rc;
    auto scal = [&](int b, int k) -> double & { return v.sc[(size_t) b * SC_COUNT + k]; };
    auto ints = [&](int b, int k) -> int & { return v.is[(size_t) b * IS_COUNT + k]; };
    std::vector<double>              err0(nb), tol_abs(nb);
    std::vector<std::vector<double>> ohist(nb);
    std::vector<int>                 inner(nb, 0), outer(nb, 0), kstart(nb, 0);
    std::vector<int>                 act;
    for (int b = 0; b < nb; ++b) {
        err0[b]    = scal(b, SC_ERR);
        tol_abs[b] = c->tol_type == FANS_ERR_ABSOLUTE ? c->tol : c->tol * err0[b];
        ohist[b]   = {err0[b]};
        if (!ints(b, IS_DONE))
            act.push_back(b);
    }
    auto phase_tol = [&](int b, double err) { return std::max(kReplaceEta * err, kReplaceFloor * tol_abs[b]); };
    int  J         = 0;
    if (!act.empty()) {
        for (int b : act)
            init_slot_scalars(c, v, b, phase_tol(b, err0[b]), FANS_ERR_ABSOLUTE, c->d.n_it);
        if ((rc = write_all_scalars(c, v)) || (rc = set_map(c, act)))
            return rc;
        // x = 0 and d_(-1) = 0 in the ring's last slot
        hipLaunchKernelGGL(k_inner_start, dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->r32, c->x32,
                           c->ring32[kDirRing - 1], c->r, c->Nh, c->d.measure, c->partial, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
        if ((rc = reduce(c, c->partial, c->nblk_vec, 1, RED_ERR_INIT)))
            return rc;
    }
    while (!act.empty()) {
        if ((rc = cg_linear_loop<float>(c, c->fft32, c->r32, c->ring32, c->rnew32, c->x32, c->spec32, c->G32,
                                        c->dph32, nullptr, J)))
            return rc;
        if ((rc = read_all_scalars(c, v)))
            return rc;
        std::vector<int> kend(nb, 0);
        for (int b : act) {
            kend[b] = ints(b, IS_ITER);
            inner[b] += kend[b] - kstart[b];
            outer[b] += 1;
        }
        // u += x and the true residual of the slots that ran, tested against
        // the outer tolerance
        hipLaunchKernelGGL((k_axpby<double, float>), dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->u,
                           c->x32, 1.0, 1.0, c->Nh, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
        if ((rc = residual64(c, c->u, c->r, nullptr)))
            return rc;
        for (int b : act)
            scal(b, SC_TOL) = tol_abs[b];
        if ((rc = write_all_scalars(c, v)) || (rc = reduce(c, c->partial, c->nblk_asm, 1, RED_ERR_SET)) ||
            (rc = read_all_scalars(c, v)))
            return rc;
        std::vector<int> next;
        for (int b : act) {
            const double err = scal(b, SC_ERR);
            ohist[b].push_back(err);
            const double e = c->tol_type == FANS_ERR_ABSOLUTE ? err : err / err0[b];
            if (e > c->tol && inner[b] < c->d.n_it && kend[b] > kstart[b])
                next.push_back(b);
        }
        act = next;
        if (act.empty())
            break;
        // resume together at the largest iteration index reached
        J = 0;
        for (int b : act)
            J = std::max(J, kend[b]);
        for (int b : act) {
            const int from = (kend[b] + kDirRing - 1) % kDirRing, to = (J + kDirRing - 1) % kDirRing;
            if (from != to)
                HIPCHK(c, hipMemcpyAsync(c->ring32[to] + (size_t) b * c->strides.field,
                                         c->ring32[from] + (size_t) b * c->strides.field,
                                         sizeof(float) * c->Nh, hipMemcpyDeviceToDevice, c->stream));
            ints(b, IS_ITER) = J;
            ints(b, IS_NIT)  = J + (c->d.n_it - inner[b]);
            ints(b, IS_DONE) = 0;
            scal(b, SC_TOL)  = phase_tol(b, scal(b, SC_ERR));
            for (int m = 0; m < kDirRing; ++m)
                scal(b, SC_ARING + m) = 0.0;  // every direction is in x (flushed) or was moved
            kstart[b] = J;
        }
        if ((rc = write_all_scalars(c, v)) || (rc = set_map(c, act)))
            return rc;
        // r32 = (float) r, x = 0; the directions stay
        hipLaunchKernelGGL(k_inner_start, dim3(c->nblk_vec, 1, c->gz), dim3(kThreads), 0, c->stream, c->r32, c->x32,
                           (float *) nullptr, c->r, c->Nh, c->d.measure, c->partial, launch_slots(c, false));
        HIPCHK(c, hipGetLastError());
    }
    for (int b = 0; b < nb; ++b) {
        H[b]->iter        = inner[b];
        H[b]->inner_iters = inner[b];
        H[b]->outer_iters = outer[b];
        H[b]->err_hist    = ohist[b];
    }
    return 0;
}

// Every slot's host state in `parked` and the fields at slot 0 for a batched
// operation; the slots must be loaded alike (linear materials, same element
// variant and Green storage).  batch_end restores the selected slot.
int batch_begin(fans_gpu_ctx *c, std::vector<SlotHost *> &H, int &sel)
{
    sel = c->cur;
    swap_host(c, c->parked[sel]);
    set_views(c, 0);
    H.resize(c->nb);
    for (int b = 0; b < c->nb; ++b)
        H[b] = &c->parked[b];
    int nph = 0;
    for (int b = 0; b < c->nb; ++b) {
        const SlotHost &h = *H[b], &h0 = *H[0];
        if (!h.have_ms || !h.have_mat)
            return fail(c, FANS_GPU_ESTATE, "slot %d: set_microstructure and set_materials must precede this call", b);
        if (h.ms_max >= h.nphases)
            return fail(c, FANS_GPU_EINVAL, "slot %d: MaterialManager: Phase %d not assigned", b, h.nphases);
        if (!h.all_linear || h.voxel_moduli)
            return fail(c, FANS_GPU_EINVAL, "slot %d: a batched context solves linear phase tables only", b);
        if (h.all_iso != h0.all_iso || h.ms16 != h0.ms16 || h.gxq != h0.gxq || h.gyq != h0.gyq)
            return fail(c, FANS_GPU_EINVAL,
                        "slot %d: the slots of a batch need the same kind of phase table, id width and Green storage",
                        b);
        nph = std::max(nph, h.nphases);
    }
    // the launch arguments of the batch (captured graphs hold them by value)
    const SlotHost &h0 = *H[0];
    c->nphases      = nph;
    c->all_iso      = h0.all_iso;
    c->ms16         = h0.ms16;
    c->gxq          = h0.gxq;
    c->gyq          = h0.gyq;
    c->all_linear   = true;
    c->has_j2       = false;
    c->has_pnl      = false;
    c->has_visco    = false;
    c->voxel_moduli = false;
    const std::array<int, 7> sig = {c->all_iso, c->cubic, c->ms16, c->gxq, c->gyq, c->nphases, 1};
    if (sig != c->batch_sig) {
        drop_graphs(c);
        c->batch_sig = sig;
    }
    return 0;
}

void batch_end(fans_gpu_ctx *c, int sel)
{
    clear_map(c);
    swap_host(c, c->parked[sel]);
    set_views(c, sel);
}

int solve_batch(fans_gpu_ctx *c)
{
    if (c->d.method != FANS_METHOD_CG)
        return fail(c, FANS_GPU_EINVAL, "a batched context solves with the CG method");
    if (c->timing)
        return fail(c, FANS_GPU_EINVAL, "per-kernel timing is not available on a batched context");
    std::vector<SlotHost *> H;
    int                     sel = 0;
    int                     rc  = batch_begin(c, H, sel);
    if (!rc)
        rc = c->d.precision == FANS_PREC_MIXED ? solve_batch_mixed(c, H) : solve_batch_cg64(c, H);
    batch_end(c, sel);
    return rc;
}

// iteration counts, final error and stop test of one RVE's last solve
struct SolveResult {
    int    iterations = 0, inner = 0, outer = 0;
    double final_error = 0.0;
    bool   converged   = false;
};
template <typename S>
SolveResult solve_result(const fans_gpu_ctx *c, const S &h)
{
    SolveResult r;
    r.iterations    = h.iter;
    r.inner         = h.inner_iters ? h.inner_iters : h.iter;
    r.outer         = h.outer_iters;
    r.final_error   = h.err_hist.empty() ? 0.0 : h.err_hist.back();
    const double e0 = h.err_hist.empty() ? 1.0 : h.err_hist[0];
    const double ev = c->tol_type == FANS_ERR_ABSOLUTE ? r.final_error : r.final_error / e0;
    r.converged     = ev <= c->tol;
    return r;
}
SolveResult slot_result(const fans_gpu_ctx *c, int b)
{
    return b == c->cur ? solve_result(c, *c) : solve_result(c, c->parked[b]);
}

// partial sums: the largest grid of any reducing kernel times its quantities;
// any_tiles: for every assembly tile depth and FFT LDS budget (one line or
// frequency per FFT workgroup, one x plane per assembly workgroup)
int64_t partial_need(const fans_gpu_ctx *c, bool any_tiles)
{
    const int64_t asm_blocks = any_tiles ? (int64_t) c->nx * c->nty * c->ntz : c->nblk_asm;
    int64_t       fft_blocks = std::max<int64_t>({c->fft64.nblk_z, c->fft32.nblk_z, x_green_blocks(c->fft64, c->h, c->ny),
                                                  x_green_blocks(c->fft32, c->h, c->ny)});
    if (any_tiles)
        fft_blocks = std::max<int64_t>({fft_blocks, (int64_t) c->nx * c->ny, (int64_t) (c->nz / 2 + 1) * c->ny});
    int64_t need = 2 * std::max<int64_t>({asm_blocks, (int64_t) c->nblk_vec, fft_blocks});
    need         = std::max<int64_t>(need, (int64_t) c->nstr * c->nblk_el);
    return std::max<int64_t>(need, 6 * 1024);  // k_moduli_stats: 6 quantities, <= 1024 blocks
}

// single-RVE contexts: grow the partial buffers to a new launch shape
int ensure_partials(fans_gpu_ctx *c)
{
    const int64_t need   = partial_need(c, false);
    const int64_t need_e = std::max<int64_t>(c->fft64.nblk_z, c->fft32.nblk_z);
    if (c->nb > 1)
        return (need > c->partial_len || need_e > c->partial_e_len)
                   ? fail(c, FANS_GPU_EINVAL, "launch shape outside the batched context's partial buffers")
                   : 0;
    if (need > c->partial_len || need_e > c->partial_e_len) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (need > c->partial_len) {
        HIPCHK(c, hipFree(c->partial));
        c->partial = nullptr;
        HIPCHK(c, hipMalloc(&c->partial, sizeof(double) * need));
        c->partial_len   = need;
        c->strides.part  = need;
    }
    if (need_e > c->partial_e_len) {
        HIPCHK(c, hipFree(c->partial_e));
        c->partial_e = nullptr;
        HIPCHK(c, hipMalloc(&c->partial_e, sizeof(double) * need_e));
        c->partial_e_len = need_e;
    }
    return 0;
}

// every per-RVE buffer nb times (batched contexts: slot b at b * stride);
// a batched context also allocates now what a single one allocates on first
// use (Green operators, phase tables, the CG direction ring), so that no slot
// buffer moves later
int alloc_fields(fans_gpu_ctx *c)
{
    const int64_t nb = c->nb;
    const size_t  vb = sizeof(double) * c->Nh * nb;
    HIPCHK(c, hipMalloc(&c->u, vb));
    HIPCHK(c, hipMalloc(&c->u_prev, vb));
    HIPCHK(c, hipMalloc(&c->r, vb));
    HIPCHK(c, hipMalloc(&c->s, vb));
    HIPCHK(c, hipMalloc(&c->dA, vb));
    HIPCHK(c, hipMalloc(&c->dB, vb));
    HIPCHK(c, hipMalloc(&c->rnew, vb));
    HIPCHK(c, hipMemsetAsync(c->u, 0, vb, c->stream));
    HIPCHK(c, hipMemsetAsync(c->u_prev, 0, vb, c->stream));
    HIPCHK(c, hipMemsetAsync(c->r, 0, vb, c->stream));
    HIPCHK(c, hipMalloc(&c->spec64, sizeof(cplx<double>) * c->Nf * c->h * nb));
    HIPCHK(c, hipMalloc(&c->g0d, sizeof(double) * 9 * nb));
    HIPCHK(c, hipMalloc(&c->g0f, sizeof(float) * 9 * nb));
    HIPCHK(c, hipMalloc(&c->sc, sizeof(double) * SC_COUNT * nb));
    HIPCHK(c, hipMalloc(&c->is, sizeof(int) * IS_COUNT * nb));
    HIPCHK(c, hipMemsetAsync(c->is, 0, sizeof(int) * IS_COUNT * nb, c->stream));
    // error history; a batched mixed solve moves a slot's iteration index
    // ahead by up to n_it (solve_batch_mixed)
    const int64_t hist_len = nb > 1 ? 2 * (int64_t) c->d.n_it + 2 : c->d.n_it + 2;
    HIPCHK(c, hipMalloc(&c->hist, sizeof(double) * hist_len * nb));
    // phase ids: bytes, room for 16-bit ids per slot of a batch
    c->ms_bytes = nb > 1 ? 2 * c->N : c->N;
    HIPCHK(c, hipMalloc(&c->ms, c->ms_bytes * nb));
    const bool mixed = c->d.precision == FANS_PREC_MIXED;
    if (mixed) {
        const size_t fb = sizeof(float) * c->Nh * nb;
        HIPCHK(c, hipMalloc(&c->r32, fb));
        HIPCHK(c, hipMalloc(&c->s32, fb));
        HIPCHK(c, hipMalloc(&c->dA32, fb));
        HIPCHK(c, hipMalloc(&c->dB32, fb));
        HIPCHK(c, hipMalloc(&c->rnew32, fb));
        HIPCHK(c, hipMalloc(&c->x32, fb));
        HIPCHK(c, hipMalloc(&c->spec32, sizeof(cplx<float>) * c->Nf * c->h * nb));
    }
    // partial sums (batched: for any tile setting, so that the per-slot
    // stride never changes); partial_e holds the per-block partials of the
    // linear CG's inverse z pass, kept apart because the matvec that follows
    // writes `partial` (batched: at the same stride)
    const int64_t need = partial_need(c, nb > 1);
    c->partial_len     = need;
    HIPCHK(c, hipMalloc(&c->partial, sizeof(double) * need * nb));
    c->partial_e_len = nb > 1 ? need : std::max<int64_t>(c->fft64.nblk_z, c->fft32.nblk_z);
    HIPCHK(c, hipMalloc(&c->partial_e, sizeof(double) * c->partial_e_len * nb));
    const int npk      = c->h * (c->h + 1) / 2;
    c->strides.field   = c->Nh;
    c->strides.spec    = c->Nf * c->h;
    c->strides.green   = c->Nf * npk;
    c->strides.part    = need;
    c->strides.ms      = c->ms_bytes;
    c->strides.hist    = hist_len;
    if (nb > 1) {
        c->nph_cap       = 256;
        c->strides.ph    = c->nph_cap;
        HIPCHK(c, hipMalloc(&c->dph64, sizeof(Phase<double>) * c->nph_cap * nb));
        HIPCHK(c, hipMalloc(&c->dph32, sizeof(Phase<float>) * c->nph_cap * nb));
        HIPCHK(c, hipMalloc(&c->G64, sizeof(double) * c->Nf * npk * nb));
        if (mixed)
            HIPCHK(c, hipMalloc(&c->G32, sizeof(float) * c->Nf * npk * nb));
        for (int m = 2; m < kDirRing; ++m) {
            if (mixed)
                HIPCHK(c, hipMalloc(&c->ring32[m], sizeof(float) * c->Nh * nb));
            else
                HIPCHK(c, hipMalloc(&c->ring64[m], sizeof(double) * c->Nh * nb));
        }
        c->map_cap = 1;
        while (c->map_cap < c->nb)
            c->map_cap *= 2;
        HIPCHK(c, hipMalloc(&c->slot_map, sizeof(int) * c->map_cap));
#define X(T, name, kind) c->base.name = c->name;
        FANS_SLOT_BUFFERS(X)
#undef X
        for (int m = 2; m < kDirRing; ++m) {
            c->base.ring64[m] = c->ring64[m];
            c->base.ring32[m] = c->ring32[m];
        }
        c->parked.assign(c->nb, SlotHost{});
        set_views(c, 0);
    }
    return 0;
}

void free_all(fans_gpu_ctx *c)
{
    if (c->nb > 1 && c->base.u)
        set_views(c, 0);  // the fields hold the allocations again
    if (c->slot_map)
        (void) hipFree(c->slot_map);
    for (void *p : {(void *) c->ms, (void *) c->dph64, (void *) c->dph32, (void *) c->u, (void *) c->u_prev,
                    (void *) c->r, (void *) c->s, (void *) c->dA, (void *) c->dB, (void *) c->rnew,
                    (void *) c->spec64, (void *) c->G64, (void *) c->keref_d, (void *) c->G32, (void *) c->js,
                    (void *) c->g0d, (void *) c->g0f, (void *) c->partial, (void *) c->partial_e, (void *) c->sc, (void *) c->is,
                    (void *) c->hist, (void *) c->r32, (void *) c->s32, (void *) c->dA32, (void *) c->dB32,
                    (void *) c->rnew32, (void *) c->x32, (void *) c->spec32, (void *) c->vk64, (void *) c->vg64,
                    (void *) c->vk32, (void *) c->vg32, (void *) c->ub_d, (void *) c->sens_d})
        if (p)
            (void) hipFree(p);
    for (int m = 2; m < kDirRing; ++m) {
        if (c->ring64[m])
            (void) hipFree(c->ring64[m]);
        if (c->ring32[m])
            (void) hipFree(c->ring32[m]);
    }
    drop_graphs(c);
    c->fft64.release();
    c->fft32.release();
    for (auto &p : c->ev_used) {
        (void) hipEventDestroy(p.second.first);
        (void) hipEventDestroy(p.second.second);
    }
    for (auto e : c->ev_pool)
        (void) hipEventDestroy(e);
}

// ---- host material tables ---------------------------------------------------
void linear_C(int nstr, int kind, const double *p, double *C)
{
    std::memset(C, 0, sizeof(double) * nstr * nstr);
    if (kind == FANS_MODEL_THERMAL_ISO) {
        for (int i = 0; i < 3; ++i)
            C[i * 3 + i] = p[0];
    } else if (kind == FANS_MODEL_THERMAL_TRI) {
        const double K[9] = {p[0], p[1], p[2], p[1], p[3], p[4], p[2], p[4], p[5]};
        std::memcpy(C, K, sizeof(K));
    } else if (kind == FANS_MODEL_ELASTIC_TRI) {
        int k = 0;
        for (int row = 0; row < 6; ++row)
            for (int col = row; col < 6; ++col) {
                C[row * 6 + col] = p[k];
                C[col * 6 + row] = p[k];
                ++k;
            }
    } else {
        const double lam = p[0] - (2.0 / 3.0) * p[1], mu = p[1];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                C[i * 6 + j] = lam;
        for (int i = 0; i < 6; ++i)
            C[i * 6 + i] += 2 * mu;
    }
}

// B_int[p] (nstr x 8h) of matmodel.h:104-155 / 157-188 / 267-304 for the
// reference element stiffness (matmodel.h:237-253)
void host_B(const fans_gpu_ctx *c, double x, double y, double z, double *out)
{
    const double lx = c->d.L[0] / c->nx, ly = c->d.L[1] / c->ny, lz = c->d.L[2] / c->nz;
    double       b[3][8];
    b[0][0]        = -(1. - y) * (1. - z) / lx;
    b[0][1]        = (1. - y) * (1. - z) / lx;
    b[0][2]        = -y * (1. - z) / lx;
    b[0][3]        = y * (1. - z) / lx;
    b[0][4]        = -(1. - y) * z / lx;
    b[0][5]        = (1. - y) * z / lx;
    b[0][6]        = -y * z / lx;
    b[0][7]        = y * z / lx;
    b[1][0]        = -(1. - x) * (1. - z) / ly;
    b[1][1]        = -x * (1. - z) / ly;
    b[1][2]        = (1. - x) * (1. - z) / ly;
    b[1][3]        = x * (1. - z) / ly;
    b[1][4]        = -(1. - x) * z / ly;
    b[1][5]        = -x * z / ly;
    b[1][6]        = (1. - x) * z / ly;
    b[1][7]        = x * z / ly;
    b[2][0]        = -(1. - x) * (1. - y) / lz;
    b[2][1]        = -x * (1. - y) / lz;
    b[2][2]        = -(1. - x) * y / lz;
    b[2][3]        = -x * y / lz;
    b[2][4]        = (1. - x) * (1. - y) / lz;
    b[2][5]        = x * (1. - y) / lz;
    b[2][6]        = (1. - x) * y / lz;
    b[2][7]        = x * y / lz;
    const int ncol = 8 * c->h;
    std::memset(out, 0, sizeof(double) * c->nstr * ncol);
    if (c->nstr == 9) {  // F_iJ rows (LargeStrainMechModel.h:205-222)
        for (int i = 0; i < 3; ++i)
            for (int J = 0; J < 3; ++J)
                for (int q = 0; q < 8; ++q)
                    out[(3 * i + J) * ncol + 3 * q + i] = b[J][q];
        return;
    }
    if (c->h == 1) {
        for (int i = 0; i < 3; ++i)
            for (int q = 0; q < 8; ++q)
                out[i * ncol + q] = b[i][q];
        return;
    }
    const double sh = kSqrtHalf;
    for (int q = 0; q < 8; ++q) {
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

void ref_element_stiffness(fans_gpu_ctx *c)
{
    const int    ne = 8 * c->h, ns = c->nstr;
    const double xp = 0.5 + std::sqrt(3.) / 6., xm = 0.5 - std::sqrt(3.) / 6.;
    const double xi[8][3] = {{xm, xm, xm}, {xp, xm, xm}, {xm, xp, xm}, {xp, xp, xm},
                             {xm, xm, xp}, {xp, xm, xp}, {xm, xp, xp}, {xp, xp, xp}};
    const double ve       = (c->d.L[0] / c->nx) * (c->d.L[1] / c->ny) * (c->d.L[2] / c->nz);
    double       tmp[576] = {0};
    double       Bc[9 * 24];
    host_B(c, 0.5, 0.5, 0.5, Bc);
    for (int p = 0; p < c->ngp; ++p) {
        double B[9 * 24];
        if (c->ngp == 1)
            host_B(c, 0.5, 0.5, 0.5, B);
        else
            host_B(c, xi[p][0], xi[p][1], xi[p][2], B);
        if (c->geom.bbar)  // volumetric rows from the centre (matmodel.h:118-131)
            for (int j = 0; j < ne; ++j) {
                const double dv = (Bc[j] + Bc[ne + j] + Bc[2 * ne + j] - B[j] - B[ne + j] - B[2 * ne + j]) / 3.0;
                for (int i = 0; i < 3; ++i)
                    B[i * ne + j] += dv;
            }
        double CB[9 * 24];
        for (int i = 0; i < ns; ++i)
            for (int j = 0; j < ne; ++j) {
                double a = 0;
                for (int k = 0; k < ns; ++k)
                    a += c->kref[i * ns + k] * B[k * ne + j];
                CB[i * ne + j] = a;
            }
        for (int i = 0; i < ne; ++i)
            for (int j = 0; j < ne; ++j) {
                double a = 0;
                for (int k = 0; k < ns; ++k)
                    a += B[k * ne + i] * CB[k * ne + j];
                tmp[i * ne + j] += a * ve / c->ngp;
            }
    }
    for (int i = 0; i < ne; ++i)
        for (int j = 0; j < ne; ++j)
            c->keref[((i % c->h) * 8 + i / c->h) * ne + (j % c->h) * 8 + j / c->h] = tmp[i * ne + j];
}

template <typename T>
int build_fft(fans_gpu_ctx *c, Fft3D<T> &P)
{
    hipError_t e = P.init(c->nx, c->ny, c->nz, c->h);
    if (e == hipErrorInvalidValue)
        return fail(c, FANS_GPU_EUNSUP, "grid sizes must be products of 2, 3 and 5 (got %d x %d x %d)", c->nx, c->ny,
                    c->nz);
    HIPCHK(c, e);
    return 0;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int fans_gpu_abi_version(void) { return FANS_GPU_ABI_VERSION; }

const char *fans_gpu_last_error(const fans_gpu_ctx *ctx)
{
    if (ctx)
        return ctx->err.c_str();
    // context-free entry points (fans_gpu_fft_r2c_noalloc) leave their message here
    return g_noalloc_err[0] ? g_noalloc_err : "null context";
}

static int create_ctx(fans_gpu_ctx **out, const fans_gpu_desc *d, int slots)
{
    if (!out || !d)
        return FANS_GPU_EINVAL;
    *out   = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0)
        return FANS_GPU_ENODEV;
    if (d->howmany != 1 && d->howmany != 3)
        return FANS_GPU_EINVAL;
    if (d->n_x < 2 || d->n_y < 2 || d->n_z < 2 || d->n_z % 2 != 0)
        return FANS_GPU_EINVAL;
    if (d->fe_type != FANS_FE_HEX8 && d->fe_type != FANS_FE_HEX8R && d->fe_type != FANS_FE_BBAR)
        return FANS_GPU_EINVAL;
    if (d->method != FANS_METHOD_CG && d->method != FANS_METHOD_FP)
        return FANS_GPU_EINVAL;
    if (d->n_it < 0 || d->device < 0 || d->device >= nd)
        return FANS_GPU_EINVAL;
    if ((d->large_strain != 0 && d->large_strain != 1) || (d->large_strain && d->howmany != 3))
        return FANS_GPU_EINVAL;
    fans_gpu_ctx *c = new fans_gpu_ctx();
    c->d            = *d;
    c->nb           = slots;
    if (hipSetDevice(d->device) != hipSuccess) {
        delete c;
        return FANS_GPU_EHIP;
    }
    c->nx       = d->n_x;
    c->ny       = d->n_y;
    c->nz       = d->n_z;
    c->h        = d->howmany;
    c->nstr     = d->howmany == 1 ? 3 : (d->large_strain ? 9 : 6);
    c->ngp      = d->fe_type == FANS_FE_HEX8R ? 1 : 8;
    c->N        = (int64_t) c->nx * c->ny * c->nz;
    c->Nh       = c->N * c->h;
    c->Nf       = (int64_t) c->nx * c->ny * (c->nz / 2 + 1);
    c->tol      = d->tol;
    c->tol_type = d->err_type;
    const double lx = d->L[0] / c->nx, ly = d->L[1] / c->ny, lz = d->L[2] / c->nz;
    c->geom.ilx     = 1.0 / lx;
    c->geom.ily     = 1.0 / ly;
    c->geom.ilz     = 1.0 / lz;
    c->geom.w       = lx * ly * lz / c->ngp;
    c->geom.q0      = 0.5 - std::sqrt(3.) / 6.;
    c->geom.q1      = 0.5 + std::sqrt(3.) / 6.;
    c->geom.bbar    = d->fe_type == FANS_FE_BBAR && c->h == 3;
    c->cubic        = lx == ly && ly == lz;
    *out            = c;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(c, FANS_GPU_EHIP, "hipStreamCreateWithFlags failed");
    // launch geometry
    c->nty      = (c->ny + NOUT - 1) / NOUT;
    c->ntz      = (c->nz + NOUT - 1) / NOUT;
    c->cx       = c->nx >= 64 ? 16 : std::max(1, c->nx / 4);
    const int nbx = (c->nx + c->cx - 1) / c->cx;
    c->nblk_asm = nbx * c->nty * c->ntz;
    c->nblk_vec = grid_vec(c->Nh);
    c->nblk_el  = (int) ((c->N + kThreads - 1) / kThreads);
    // J2 state slots (J2State); the state commit walks them
    c->nslot    = (int64_t) c->nx * c->ny * c->ntz * 16;
    c->nblk_slot = (int) ((c->nslot + kThreads - 1) / kThreads);
    int rc;
    if ((rc = build_fft<double>(c, c->fft64)))
        return rc;
    if (d->precision == FANS_PREC_MIXED && (rc = build_fft<float>(c, c->fft32)))
        return rc;
    if ((rc = alloc_fields(c)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fans_gpu_create(fans_gpu_ctx **out, const fans_gpu_desc *d) { return create_ctx(out, d, 1); }

int fans_gpu_create_batch(fans_gpu_ctx **out, const fans_gpu_desc *d, int32_t slots)
{
    if (out)
        *out = nullptr;
    if (slots < 1 || slots > kMaxSlots || (d && d->method != FANS_METHOD_CG))
        return FANS_GPU_EINVAL;
    return create_ctx(out, d, slots);
}

int fans_gpu_select_slot(fans_gpu_ctx *c, int32_t slot)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (slot < 0 || slot >= c->nb)
        return fail(c, FANS_GPU_EINVAL, "slot %d outside [0, %d)", slot, c->nb);
    select_slot(c, slot);
    return 0;
}

int fans_gpu_batch_results(fans_gpu_ctx *c, int32_t *iterations, int32_t *converged, double *final_error,
                           int32_t count)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (count != c->nb)
        return fail(c, FANS_GPU_EINVAL, "count must be the context's slot count %d", c->nb);
    for (int b = 0; b < c->nb; ++b) {
        const SolveResult r = slot_result(c, b);
        if (iterations)
            iterations[b] = r.inner;
        if (converged)
            converged[b] = r.converged;
        if (final_error)
            final_error[b] = r.final_error;
    }
    return 0;
}

int fans_gpu_batch_homogenized_stress(fans_gpu_ctx *c, double *out, int32_t n_str, int32_t count)
{
    if (!c || !out || n_str != c->nstr)
        return FANS_GPU_EINVAL;
    if (count != c->nb)
        return fail(c, FANS_GPU_EINVAL, "count must be the context's slot count %d", c->nb);
    if (c->nb == 1)
        return fans_gpu_homogenized_stress(c, out, n_str);
    std::vector<SlotHost *> H;
    int                     sel = 0;
    int                     rc  = batch_begin(c, H, sel);
    std::vector<int>        all(c->nb);
    for (int b = 0; b < c->nb; ++b)
        all[b] = b;
    BatchScalars v;
    if (!rc && !(rc = set_map(c, all)) && !(rc = homog_stress(c, nullptr)) && !(rc = read_all_scalars(c, v)))
        for (int b = 0; b < c->nb; ++b)
            for (int k = 0; k < c->nstr; ++k)
                out[(size_t) b * c->nstr + k] = v.sc[(size_t) b * SC_COUNT + SC_RES0 + k] / (double) c->N;
    batch_end(c, sel);
    return rc;
}

int fans_gpu_destroy(fans_gpu_ctx *c)
{
    if (!c)
        return 0;
    if (c->stream)
        (void) hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->stream)
        (void) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int fans_gpu_set_microstructure(fans_gpu_ctx *c, const uint16_t *ms, int64_t count)
{
    if (!c || !ms || count != c->N)
        return fail(c, FANS_GPU_EINVAL, "microstructure must hold n_x*n_y*n_z entries");
    int mx = -1;
    for (int64_t i = 0; i < c->N; ++i)
        mx = std::max(mx, (int) ms[i]);
    // bytes on the device when every id fits (1 B per voxel of phase traffic),
    // else the 16-bit ids as given
    const bool    wide  = mx > 255;
    const int64_t bytes = wide ? 2 * c->N : c->N;
    if (bytes > c->ms_bytes) {
        (void) hipFree(c->ms);
        c->ms       = nullptr;
        c->ms_bytes = 0;
        HIPCHK(c, hipMalloc(&c->ms, bytes));
        c->ms_bytes = bytes;
    }
    if (wide) {
        HIPCHK(c, copy_on(c, c->ms, ms, bytes, hipMemcpyHostToDevice));
    } else {
        std::vector<uint8_t> m8(c->N);
        for (int64_t i = 0; i < c->N; ++i)
            m8[i] = (uint8_t) ms[i];
        HIPCHK(c, copy_on(c, c->ms, m8.data(), bytes, hipMemcpyHostToDevice));
    }
    if (wide != c->ms16)
        drop_graphs(c);  // the captured launches carry the id width by value
    c->ms16    = wide;
    c->ms_max  = mx;
    c->have_ms = true;
    return 0;
}

// reference element stiffness (from c->kref) and the Green operator on the device
int rebuild_green(fans_gpu_ctx *c)
{
    ref_element_stiffness(c);
    const int ne = 8 * c->h, npk = c->h * (c->h + 1) / 2;
    if (!c->keref_d)
        HIPCHK(c, hipMalloc(&c->keref_d, sizeof(double) * 576));
    HIPCHK(c, copy_on(c, c->keref_d, c->keref, sizeof(double) * ne * ne, hipMemcpyHostToDevice));
    if (!c->G64)
        HIPCHK(c, hipMalloc(&c->G64, sizeof(double) * c->Nf * npk));
    // quarter storage when the reference element is mirror symmetric in x and y
    const bool quarter = c->nx % 2 == 0 && c->ny % 2 == 0 && keref_mirror_symmetric(c->keref, c->h, 0) &&
                         keref_mirror_symmetric(c->keref, c->h, 1);
    c->gxq            = quarter ? c->nx / 2 + 1 : c->nx;
    c->gyq            = quarter ? c->ny / 2 + 1 : c->ny;
    const int64_t ngs = (int64_t) c->gxq * c->gyq * (c->nz / 2 + 1);  // stored frequencies
    const int     gb  = (int) ((ngs + kThreads - 1) / kThreads);
    if (c->h == 1)
        hipLaunchKernelGGL((k_build_green<1>), dim3(gb), dim3(kThreads), 0, c->stream, c->keref_d, c->G64, c->nx,
                           c->ny, c->nz, c->gxq, c->gyq);
    else
        hipLaunchKernelGGL((k_build_green<3>), dim3(gb), dim3(kThreads), 0, c->stream, c->keref_d, c->G64, c->nx,
                           c->ny, c->nz, c->gxq, c->gyq);
    HIPCHK(c, hipGetLastError());
    if (c->d.precision == FANS_PREC_MIXED) {
        if (!c->G32)
            HIPCHK(c, hipMalloc(&c->G32, sizeof(float) * c->Nf * npk));
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_vec(ngs * npk)), dim3(kThreads), 0, c->stream,
                           c->G32, c->G64, ngs * npk);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Cholesky factorisation of the lower triangle of the row-major n x n matrix A
// (n <= 9), the test behind MaterialManager::compute_reference_stiffness's
// "must be symmetric positive definite" (MaterialManager.h:189-192, an Eigen
// LLT, which reads the lower triangle only): false on a non-positive or NaN pivot.
static bool cholesky_lower_ok(const double *A, int n)
{
    double Lm[81] = {};
    for (int j = 0; j < n; ++j) {
        double d = A[j * n + j];
        for (int k = 0; k < j; ++k)
            d -= Lm[j * n + k] * Lm[j * n + k];
        if (!(d > 0.0))
            return false;
        const double ljj = std::sqrt(d);
        Lm[j * n + j]    = ljj;
        for (int i = j + 1; i < n; ++i) {
            double v = A[i * n + j];
            for (int k = 0; k < j; ++k)
                v -= Lm[i * n + k] * Lm[j * n + k];
            Lm[i * n + j] = v / ljj;
        }
    }
    return true;
}

int fans_gpu_set_materials(fans_gpu_ctx *c, int32_t ng, const int32_t *kind, const int32_t *nphase,
                           const int32_t *phase_ids, const double *params, int64_t n_params, const double *ref)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (ng <= 0 || !kind || !nphase || !phase_ids || !params)
        return fail(c, FANS_GPU_EINVAL, "MaterialManager: 'materials' must be non-empty array");
    int     maxp = -1, tot = 0;
    int64_t np = 0;
    for (int g = 0; g < ng; ++g) {
        if (nparams(kind[g]) < 0)
            return fail(c, FANS_GPU_EINVAL, "unknown material model kind %d", kind[g]);
        const bool thermal = kind[g] <= 2;
        if (thermal != (c->h == 1))
            return fail(c, FANS_GPU_EINVAL, "material model kind %d does not match howmany=%d", kind[g], c->h);
        // createMatmodel<3,6> / <3,9> (setup.h:21-73)
        if (c->h == 3 && model_large(kind[g]) != (c->nstr == 9))
            return fail(c, FANS_GPU_EINVAL, "material model kind %d is not a valid %s strain material model", kind[g],
                        c->nstr == 9 ? "large" : "small");
        for (int i = 0; i < nphase[g]; ++i)
            maxp = std::max(maxp, phase_ids[tot + i]);
        tot += nphase[g];
        np += (int64_t) nphase[g] * nparams(kind[g]);
    }
    if (np != n_params)
        return fail(c, FANS_GPU_EINVAL, "expected %lld material parameters, got %lld", (long long) np,
                    (long long) n_params);
    if (maxp < 0)
        return fail(c, FANS_GPU_EINVAL, "MaterialManager: No phases defined");
    if (maxp > 65535)
        return fail(c, FANS_GPU_EINVAL, "phase ids are unsigned 16-bit (reader.h:53): %d is too large", maxp);
    c->nphases = maxp + 1;
    c->phase_kind.assign(c->nphases, 0);
    std::vector<int>          group_of(c->nphases, -1), local_of(c->nphases, 0);
    std::vector<const double *> pp(c->nphases, nullptr);
    c->all_linear = true;
    c->has_j2     = false;
    c->has_pnl    = false;
    c->has_visco  = false;
    int     ofs   = 0;
    int64_t pofs  = 0;
    for (int g = 0; g < ng; ++g) {
        if (!model_linear(kind[g]))
            c->all_linear = false;
        if (kind[g] == FANS_MODEL_J2_NEW || model_visco(kind[g]))
            c->has_j2 = true;
        if (model_visco(kind[g]))
            c->has_visco = true;
        if (kind[g] == FANS_MODEL_PSEUDO_PLASTIC_NL)
            c->has_pnl = true;
        for (int i = 0; i < nphase[g]; ++i) {
            const int p = phase_ids[ofs + i];
            if (p < 0 || group_of[p] >= 0)
                return fail(c, FANS_GPU_EINVAL, "MaterialManager: Invalid or duplicate phase %d", p);
            group_of[p]      = g;
            local_of[p]      = i;
            c->phase_kind[p] = kind[g];
            pp[p]            = params + pofs + (int64_t) i * nparams(kind[g]);
        }
        ofs += nphase[g];
        pofs += (int64_t) nphase[g] * nparams(kind[g]);
    }
    for (int p = 0; p < c->nphases; ++p)
        if (group_of[p] < 0)
            return fail(c, FANS_GPU_EINVAL, "MaterialManager: Phase %d not assigned", p);
    // reference stiffness (MaterialManager.h:177-205)
    const int ns = c->nstr;
    if (ref && !cholesky_lower_ok(ref, ns))
        return fail(c, FANS_GPU_EINVAL, "reference_material must be symmetric positive definite");
    c->user_ref  = ref != nullptr;
    if (ref) {
        std::memcpy(c->kref, ref, sizeof(double) * ns * ns);
    } else {
        std::memset(c->kref, 0, sizeof(c->kref));
        ofs  = 0;
        pofs = 0;
        for (int g = 0; g < ng; ++g) {
            double       C[81] = {0};
            const int    npg   = nphase[g], npk = nparams(kind[g]);
            const double *gp   = params + pofs;
            if (model_large(kind[g])) {
                // get_reference_stiffness of both finite strain models: dP/dF at
                // F = I averaged over the phases (SaintVenantKirchhoff.h:49-66,
                // CompressibleNeoHookean.h:89-104).  There S = 0 and dS/dE is the
                // isotropic lam m m^T + 2 mu I, so the tensor is
                // lam d_iJ d_kL + mu (d_ik d_JL + d_iL d_Jk); the reference's sum over
                // the symmetric pair (P, Q) with P <= Q only
                // (LargeStrainMechModel.h:142-170) halves the off-diagonal F_kL columns.
                for (int q = 0; q < npg; ++q) {
                    const double mu = gp[q * npk + 1], lam = gp[q * npk] - (2.0 / 3.0) * mu;
                    for (int i = 0; i < 3; ++i)
                        for (int J = 0; J < 3; ++J)
                            for (int k = 0; k < 3; ++k)
                                for (int L = 0; L < 3; ++L) {
                                    const double t = lam * (i == J) * (k == L) +
                                                     mu * ((i == k) * (J == L) + (i == L) * (J == k));
                                    C[(3 * i + J) * 9 + 3 * k + L] += t * (k == L ? 1.0 : 0.5) / npg;
                                }
                }
            } else if (kind[g] == FANS_MODEL_ELASTIC_ISO) {
                double lmax = -INFINITY, lmin = INFINITY, mmax = -INFINITY, mmin = INFINITY;
                for (int i = 0; i < npg; ++i) {
                    const double lam = gp[i * npk] - (2.0 / 3.0) * gp[i * npk + 1];
                    lmax = std::max(lmax, lam), lmin = std::min(lmin, lam);
                    mmax = std::max(mmax, gp[i * npk + 1]), mmin = std::min(mmin, gp[i * npk + 1]);
                }
                const double lr = (lmax + lmin) / 2, mr = (mmax + mmin) / 2;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        C[i * 6 + j] = lr;
                for (int i = 0; i < 6; ++i)
                    C[i * 6 + i] += 2 * mr;
            } else if (kind[g] == FANS_MODEL_J2_NEW || kind[g] == FANS_MODEL_PSEUDO_PLASTIC ||
                       kind[g] == FANS_MODEL_PSEUDO_PLASTIC_NL || model_visco(kind[g])) {
                // mean K, mean G (J2PlasticityNew.h:111-121, PseudoPlastic.h:43-54,
                // J2Plasticity.h:110-120)
                double Kb = 0, Gb = 0;
                for (int i = 0; i < npg; ++i) {
                    Kb += gp[i * npk];
                    Gb += gp[i * npk + 1];
                }
                Kb /= npg;
                Gb /= npg;
                const double lb = Kb - 2.0 * Gb / 3.0;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        C[i * 6 + j] = lb;
                for (int i = 0; i < 6; ++i)
                    C[i * 6 + i] += 2.0 * Gb;
            } else {
                double Ci[81];
                for (int i = 0; i < npg; ++i) {
                    linear_C(ns, kind[g], gp + i * npk, Ci);
                    for (int k = 0; k < ns * ns; ++k)
                        C[k] += Ci[k];
                }
                for (int k = 0; k < ns * ns; ++k)
                    C[k] /= npg;
            }
            for (int k = 0; k < ns * ns; ++k)
                c->kref[k] += C[k];
            ofs += npg;
            pofs += (int64_t) npg * npk;
        }
        for (int k = 0; k < ns * ns; ++k)
            c->kref[k] /= (double) ng;
    }
    // per-phase device tables
    c->ph64.assign(c->nphases, Phase<double>{});
    c->ph32.assign(c->nphases, Phase<float>{});
    for (int p = 0; p < c->nphases; ++p) {
        Phase<double> &P = c->ph64[p];
        P.kind           = c->phase_kind[p];
        const double *q  = pp[p];
        switch (P.kind) {
        case FANS_MODEL_THERMAL_ISO: P.p[0] = q[0]; break;
        case FANS_MODEL_THERMAL_TRI:
        case FANS_MODEL_ELASTIC_TRI: linear_C(ns, P.kind, q, P.p); break;
        case FANS_MODEL_ELASTIC_ISO:
            P.p[0] = q[0] - (2.0 / 3.0) * q[1];
            P.p[1] = q[1];
            break;
        case FANS_MODEL_J2_NEW:
            for (int i = 0; i < 4; ++i)
                P.p[i] = q[i];
            break;
        case FANS_MODEL_J2_VISCO_LIN:
        case FANS_MODEL_J2_VISCO_NL:
            for (int i = 0; i < nparams(P.kind); ++i)
                P.p[i] = q[i];
            break;
        case FANS_MODEL_PSEUDO_PLASTIC:
            for (int i = 0; i < 4; ++i)
                P.p[i] = q[i];
            P.p[4] = std::sqrt(2. / 3.) * q[2] / (2. * q[1]);
            P.p[5] = (3. * q[1]) / (3. * q[1] + q[3]);
            P.p[6] = local_of[p];           // plastic_flag: index in the group ...
            P.p[7] = nphase[group_of[p]];   // ... and the group's phase count
            break;
        case FANS_MODEL_PSEUDO_PLASTIC_NL:  // eps_crit of PseudoPlastic.h:137-140
            for (int i = 0; i < 4; ++i)
                P.p[i] = q[i];
            P.p[4] = q[4] * std::pow(q[2] / (3.0 * q[1] * q[4]), 1.0 / (1.0 - q[3]));
            P.p[5] = q[4];
            P.p[6] = local_of[p];
            P.p[7] = nphase[group_of[p]];
            break;
        case FANS_MODEL_SVK:
        case FANS_MODEL_NEO_HOOKE:  // lambda = K - 2/3 G, mu = G (SaintVenantKirchhoff.h:26-31)
            P.p[0] = q[0] - (2.0 / 3.0) * q[1];
            P.p[1] = q[1];
            break;
        }
        c->ph32[p].kind = P.kind;
        for (int i = 0; i < 36; ++i)
            c->ph32[p].p[i] = (float) P.p[i];
    }
    // a batched context keeps its fixed per-slot tables and no J2 state
    if (c->nb > 1 && (c->nphases > c->nph_cap || !c->all_linear)) {
        c->have_mat = false;
        return fail(c, FANS_GPU_EINVAL, "a batched context takes at most %d phases, all with linear models",
                    c->nph_cap);
    }
    if (c->nphases > c->nph_cap) {
        (void) hipFree(c->dph64);
        (void) hipFree(c->dph32);
        c->dph64 = nullptr;
        c->dph32 = nullptr;
        c->nph_cap = 0;
        const int cap = std::max(256, c->nphases);
        HIPCHK(c, hipMalloc(&c->dph64, sizeof(Phase<double>) * cap));
        HIPCHK(c, hipMalloc(&c->dph32, sizeof(Phase<float>) * cap));
        c->nph_cap = cap;
        drop_graphs(c);  // captured launches hold the old table addresses
    }
    HIPCHK(c, copy_on(c, c->dph64, c->ph64.data(), sizeof(Phase<double>) * c->nphases, hipMemcpyHostToDevice));
    HIPCHK(c, copy_on(c, c->dph32, c->ph32.data(), sizeof(Phase<float>) * c->nphases, hipMemcpyHostToDevice));
    // J2 state (zero, Solver ctor solver.h:139); the visco models' kinematic
    // and current variables widen it
    const int ncomp = c->has_visco ? kViscoComps : kJ2Comps;
    if (c->has_j2 && (!c->js || c->js_comps < ncomp)) {
        (void) hipFree(c->js);
        c->js       = nullptr;
        c->js_comps = 0;
        const size_t bytes = sizeof(double) * ncomp * c->ngp * c->nslot;
        HIPCHK(c, hipMalloc(&c->js, bytes));
        HIPCHK(c, fill_on(c, c->js, bytes));
        c->js_comps = ncomp;
        drop_graphs(c);
    }
    c->voxel_moduli = false;  // new phase tables replace per-voxel moduli
    if (const int rcg = rebuild_green(c))
        return rcg;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->all_iso = true;
    for (int p = 0; p < c->nphases; ++p)
        if (c->phase_kind[p] != (c->h == 1 ? FANS_MODEL_THERMAL_ISO : FANS_MODEL_ELASTIC_ISO))
            c->all_iso = false;
    c->have_mat = true;
    if (c->nb == 1)  // batched solves check their graphs themselves
        refresh_graphs(c);
    return 0;
}

int fans_gpu_set_gradient(fans_gpu_ctx *c, const double *g0, int32_t n)
{
    if (!c || !g0 || n != c->nstr)
        return fail(c, FANS_GPU_EINVAL, "Invalid length of loading vector: expected %d", c ? c->nstr : 0);
    for (int i = 0; i < n; ++i)
        c->g0[i] = g0[i];
    int rc = upload_g0(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fans_gpu_get_gradient(fans_gpu_ctx *c, double *g0, int32_t n)
{
    if (!c || !g0 || n != c->nstr)
        return FANS_GPU_EINVAL;
    for (int i = 0; i < n; ++i)
        g0[i] = c->g0[i];
    return 0;
}

int fans_gpu_enable_mixed_bc(fans_gpu_ctx *c, int32_t nE, const int32_t *iE, int32_t nF, const int32_t *iF,
                             int32_t nsteps, const double *FE, const double *PF, int32_t step)
{
    if (c && c->nb > 1)
        return fail(c, FANS_GPU_EINVAL, "mixed boundary conditions are not available on a batched context");
    if (!c)
        return FANS_GPU_EINVAL;
    if (!c->have_mat)
        return fail(c, FANS_GPU_ESTATE, "set_materials must precede enable_mixed_bc");
    const int n = c->nstr;
    if (nE + nF != n || step < 0 || step >= nsteps)
        return fail(c, FANS_GPU_EINVAL, "each component must be either strain- or stress-controlled");
    std::vector<char> present(n, 0);
    for (int i = 0; i < nE; ++i) {
        if (iE[i] < 0 || iE[i] >= n)
            return fail(c, FANS_GPU_EINVAL, "strain index out of range");
        present[iE[i]] = 1;
    }
    for (int i = 0; i < nF; ++i) {
        if (iF[i] < 0 || iF[i] >= n)
            return fail(c, FANS_GPU_EINVAL, "stress index out of range");
        if (present[iF[i]])
            return fail(c, FANS_GPU_EINVAL, "index appears in both strain_indices and stress_indices");
        present[iF[i]] = 1;
    }
    c->mixed_active = true;
    c->idxE.assign(iE, iE + nE);
    c->idxF.assign(iF, iF + nF);
    if (FE)
        c->FE.assign(FE, FE + (size_t) nsteps * nE);
    else
        c->FE.clear();
    if (PF)
        c->PF.assign(PF, PF + (size_t) nsteps * nF);
    else
        c->PF.clear();
    c->nsteps = nsteps;
    c->step   = step;
    if (nF) {  // M = pinv(Q_F^T C0 Q_F) (mixedBCs.h:252-253)
        std::vector<double> A(nF * nF);
        bool                sym = true;
        for (int i = 0; i < nF; ++i)
            for (int j = 0; j < nF; ++j) {
                A[i * nF + j] = c->kref[iF[i] * n + iF[j]];
                sym           = sym && c->kref[iF[i] * n + iF[j]] == c->kref[iF[j] * n + iF[i]];
            }
        c->M.assign(nF * nF, 0.0);
        if (sym || host_gauss_inverse(nF, A.data(), c->M.data()))
            host_sym_pinv(nF, A.data(), c->M.data());
    }
    if (step == 0) {
        std::memset(c->gvec, 0, sizeof(c->gvec));
        if (n == 9)  // finite strain starts from F = I (mixedBCs.h:191-197)
            c->gvec[0] = c->gvec[4] = c->gvec[8] = 1.0;
        std::memcpy(c->gvec_prev, c->gvec, sizeof(c->gvec));
    } else {
        double delta[9];
        for (int i = 0; i < n; ++i)
            delta[i] = c->gvec[i] - c->gvec_prev[i];
        std::memcpy(c->gvec_prev, c->gvec, sizeof(c->gvec));
        for (int i = 0; i < nF; ++i)
            c->gvec[iF[i]] += delta[iF[i]];
    }
    for (int i = 0; i < nE; ++i)
        c->gvec[iE[i]] = c->FE[(size_t) step * nE + i];
    for (int i = 0; i < n; ++i)
        c->g0[i] = c->gvec[i];
    int rc = upload_g0(c);
    if (rc)
        return rc;
    rc = mixed_update(c);
    if (rc)
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fans_gpu_disable_mixed_bc(fans_gpu_ctx *c)
{
    if (!c)
        return FANS_GPU_EINVAL;
    c->mixed_active = false;
    return 0;
}

// microstructure and materials set, and every phase id of the microstructure
// has a material (MaterialManager.h:104-107): the kernels index the phase
// table with the ids unchecked
int check_ready(fans_gpu_ctx *c)
{
    if (!c->have_ms || !c->have_mat)
        return fail(c, FANS_GPU_ESTATE, "set_microstructure and set_materials must precede this call");
    if (c->ms_max >= c->nphases)
        return fail(c, FANS_GPU_EINVAL, "MaterialManager: Phase %d not assigned", c->nphases);
    return 0;
}

int fans_gpu_solve(fans_gpu_ctx *c, fans_gpu_stats *st)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (c->nb == 1)
        if (int rc = check_ready(c))
            return rc;
    for (int k = 0; k < KF_COUNT; ++k) {
        c->fam_ms[k] = 0;
        c->fam_n[k]  = 0;
    }
    c->inner_iters = c->outer_iters = 0;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    int rc;
    if (c->nb > 1)
        rc = solve_batch(c);
    else if (c->d.precision == FANS_PREC_MIXED && c->d.method == FANS_METHOD_CG)
        rc = solve_cg_mixed(c);
    else if (c->d.method == FANS_METHOD_FP)
        rc = solve_fp_fp64(c);
    else
        rc = solve_cg_fp64(c);
    if (rc)
        return rc;
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    c->solve_ms = ms;
    if (c->nb > 1) {
        // the batch's time for every slot; the statistics: the largest
        // iteration counts and error, converged if every slot is
        for (auto &h : c->parked)
            h.solve_ms = ms;
        if (st) {
            *st           = fans_gpu_stats{};
            st->converged = 1;
            for (int b = 0; b < c->nb; ++b) {
                const SolveResult r  = slot_result(c, b);
                st->iterations       = std::max(st->iterations, r.iterations);
                st->inner_iterations = std::max(st->inner_iterations, r.inner);
                st->outer_iterations = std::max(st->outer_iterations, r.outer);
                st->final_error      = std::max(st->final_error, r.final_error);
                st->converged        = st->converged && r.converged;
            }
            st->solve_ms = ms;
        }
        return 0;
    }
    if (c->timing)
        harvest_timing(c);
    // MaterialManager::update_internal_variables (solver.h:299)
    if ((rc = commit_state(c)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (st) {
        const SolveResult r  = solve_result(c, *c);
        st->iterations       = r.iterations;
        st->inner_iterations = r.inner;
        st->outer_iterations = r.outer;
        st->solve_ms         = c->solve_ms;
        st->final_error      = r.final_error;
        st->converged        = r.converged;
    }
    return 0;
}

int fans_gpu_error_history(fans_gpu_ctx *c, double *out, int32_t cap, int32_t *count)
{
    if (!c)
        return FANS_GPU_EINVAL;
    const int n = (int) c->err_hist.size();
    if (count)
        *count = n;
    if (out)
        std::memcpy(out, c->err_hist.data(), sizeof(double) * std::min(n, cap));
    return 0;
}

int fans_gpu_homogenized_stress(fans_gpu_ctx *c, double *out, int32_t n)
{
    if (!c || !out || n != c->nstr)
        return FANS_GPU_EINVAL;
    if (int rc = check_ready(c))
        return rc;
    return homog_stress(c, out);
}

int fans_gpu_homogenized_tangent(fans_gpu_ctx *c, double pert, double *out, int32_t n)
{
    if (c && c->nb > 1)
        return fail(c, FANS_GPU_EINVAL, "the homogenized tangent is not available on a batched context");
    if (!c || !out || n != c->nstr)
        return FANS_GPU_EINVAL;
    if (int rc0 = check_ready(c))
        return rc0;
    const int ns = c->nstr;
    double    unpert[9], g0[9], pe[9], ps[9], C[81];
    int       rc = homog_stress(c, unpert);
    if (rc)
        return rc;
    std::memcpy(g0, c->g0, sizeof(g0));
    // solver.h:749-750: later solves use the relative error with TOL >= 1e-6
    c->tol_type = FANS_ERR_RELATIVE;
    c->tol      = std::max(1e-6, c->tol);
    for (int i = 0; i < ns; ++i) {
        if (c->all_linear) {
            std::memset(pe, 0, sizeof(pe));
            pe[i] = 1.0;
        } else {
            std::memcpy(pe, g0, sizeof(pe));
            pe[i] += pert;
        }
        if ((rc = fans_gpu_set_gradient(c, pe, ns)))
            return rc;
        c->mixed_active = false;
        if ((rc = fans_gpu_solve(c, nullptr)) || (rc = homog_stress(c, ps)))
            return rc;
        for (int j = 0; j < ns; ++j)
            C[j * ns + i] = c->all_linear ? ps[j] : (ps[j] - unpert[j]) / pert;
    }
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < ns; ++j)
            out[i * ns + j] = 0.5 * (C[i * ns + j] + C[j * ns + i]);
    return 0;
}

int fans_gpu_extrapolate_displacement(fans_gpu_ctx *c)
{
    if (!c)
        return FANS_GPU_EINVAL;
    hipLaunchKernelGGL(k_extrapolate, dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->u, c->u_prev, c->Nh);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fans_gpu_reset_state(fans_gpu_ctx *c)
{
    if (!c)
        return FANS_GPU_EINVAL;
    int rc;
    if ((rc = fill<double>(c, c->u, 0.0, c->Nh)) || (rc = fill<double>(c, c->u_prev, 0.0, c->Nh)) ||
        (rc = fill<double>(c, c->r, 0.0, c->Nh)))
        return rc;
    if (c->js && (rc = fill<double>(c, c->js, 0.0, (int64_t) c->js_comps * c->ngp * c->nslot)))
        return rc;
    c->mixed_active = false;
    std::memset(c->gvec, 0, sizeof(c->gvec));
    std::memset(c->gvec_prev, 0, sizeof(c->gvec_prev));
    c->tol      = c->d.tol;
    c->tol_type = c->d.err_type;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int fans_gpu_postprocess(fans_gpu_ctx *c, double *strain_elem, double *stress_elem, double *strain_gp,
                         double *stress_gp, double *stress_avg, double *strain_avg, double *phase_stress_avg,
                         double *phase_strain_avg, int32_t *phase_counts, int32_t n_slots)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (int rc = check_ready(c))
        return rc;
    const int ns = c->nstr, ng = c->ngp;
    double   *de = nullptr, *ds = nullptr, *dge = nullptr, *dgs = nullptr;
    HIPCHK(c, hipMalloc(&de, sizeof(double) * c->N * ns));
    HIPCHK(c, hipMalloc(&ds, sizeof(double) * c->N * ns));
    if (strain_gp)
        HIPCHK(c, hipMalloc(&dge, sizeof(double) * c->N * ns * ng));
    if (stress_gp)
        HIPCHK(c, hipMalloc(&dgs, sizeof(double) * c->N * ns * ng));
    ElemArgs<double> a{};
    a.nx          = c->nx;
    a.ny          = c->ny;
    a.nz          = c->nz;
    a.ntz         = c->ntz;
    a.geom        = c->geom;
    a.ms          = c->ms;
    a.ms16 = c->ms16;
    a.phases      = c->dph64;
    a.js          = J2State{c->js, c->nslot};
    a.u           = c->u;
    a.g0          = c->g0d;
    a.vk      = c->voxel_moduli ? c->vk64 : nullptr;
    a.vg      = c->voxel_moduli ? c->vg64 : nullptr;
    a.strain_elem = de;
    a.stress_elem = ds;
    a.strain_gp   = dge;
    a.stress_gp   = dgs;
#define EP(HH, GG)                                                                                               \
    hipLaunchKernelGGL((k_element_pass<double, HH, GG, ELEM_FIELDS>), dim3(c->nblk_el), dim3(kThreads), 0, \
                       c->stream, a)
    if (c->nstr == 9) {
        if (ng == 8)
            hipLaunchKernelGGL((k_element_pass<double, 3, 8, ELEM_FIELDS, 9>), dim3(c->nblk_el), dim3(kThreads), 0,
                               c->stream, a);
        else
            hipLaunchKernelGGL((k_element_pass<double, 3, 1, ELEM_FIELDS, 9>), dim3(c->nblk_el), dim3(kThreads), 0,
                               c->stream, a);
    } else if (c->h == 1) {
        if (ng == 8)
            EP(1, 8);
        else
            EP(1, 1);
    } else {
        if (ng == 8)
            EP(3, 8);
        else
            EP(3, 1);
    }
#undef EP
    HIPCHK(c, hipGetLastError());
    std::vector<double>  he(c->N * ns), hs(c->N * ns);
    std::vector<uint8_t> hm(c->ms16 ? 2 * c->N : c->N);
    HIPCHK(c, hipMemcpyAsync(he.data(), de, sizeof(double) * c->N * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hs.data(), ds, sizeof(double) * c->N * ns, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hm.data(), c->ms, hm.size(), hipMemcpyDeviceToHost, c->stream));
    if (strain_gp)
        HIPCHK(c, hipMemcpyAsync(strain_gp, dge, sizeof(double) * c->N * ns * ng, hipMemcpyDeviceToHost, c->stream));
    if (stress_gp)
        HIPCHK(c, hipMemcpyAsync(stress_gp, dgs, sizeof(double) * c->N * ns * ng, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (void *p : {(void *) de, (void *) ds, (void *) dge, (void *) dgs})
        if (p)
            (void) hipFree(p);
    if (strain_elem)
        std::memcpy(strain_elem, he.data(), sizeof(double) * c->N * ns);
    if (stress_elem)
        std::memcpy(stress_elem, hs.data(), sizeof(double) * c->N * ns);
    // averages in element order like solver.h:546-574
    double sa[9] = {0}, ea[9] = {0};
    if (phase_stress_avg)
        std::memset(phase_stress_avg, 0, sizeof(double) * n_slots * ns);
    if (phase_strain_avg)
        std::memset(phase_strain_avg, 0, sizeof(double) * n_slots * ns);
    if (phase_counts)
        std::memset(phase_counts, 0, sizeof(int32_t) * n_slots);
    for (int64_t e = 0; e < c->N; ++e) {
        const int p = c->ms16 ? (int) reinterpret_cast<const uint16_t *>(hm.data())[e] : (int) hm[e];
        for (int j = 0; j < ns; ++j) {
            sa[j] += hs[e * ns + j];
            ea[j] += he[e * ns + j];
            if (p < n_slots) {
                if (phase_stress_avg)
                    phase_stress_avg[p * ns + j] += hs[e * ns + j];
                if (phase_strain_avg)
                    phase_strain_avg[p * ns + j] += he[e * ns + j];
            }
        }
        if (phase_counts && p < n_slots)
            phase_counts[p]++;
    }
    for (int j = 0; j < ns; ++j) {
        if (stress_avg)
            stress_avg[j] = sa[j] / (double) c->N;
        if (strain_avg)
            strain_avg[j] = ea[j] / (double) c->N;
    }
    if (phase_counts)
        for (int p = 0; p < n_slots; ++p)
            if (phase_counts[p] > 0)
                for (int j = 0; j < ns; ++j) {
                    if (phase_stress_avg)
                        phase_stress_avg[p * ns + j] /= phase_counts[p];
                    if (phase_strain_avg)
                        phase_strain_avg[p * ns + j] /= phase_counts[p];
                }
    return 0;
}

int fans_gpu_get_field(fans_gpu_ctx *c, int32_t which, double *host, int64_t count)
{
    if (!c || !host)
        return FANS_GPU_EINVAL;
    const double *src = nullptr;
    int64_t       n   = c->Nh;
    if (which == FANS_FIELD_U)
        src = c->u;
    else if (which == FANS_FIELD_R)
        src = c->r;
    else if (which == FANS_FIELD_U_PREV)
        src = c->u_prev;
    else if (which == FANS_FIELD_PLASTIC_STRAIN || which == FANS_FIELD_HARDENING ||
             which == FANS_FIELD_KIN_HARDENING) {
        const int ng = c->ngp;
        const int nc = which == FANS_FIELD_HARDENING ? 1 : 6;  // components per Gauss point
        const int c0 = which == FANS_FIELD_PLASTIC_STRAIN ? 0 : (which == FANS_FIELD_HARDENING ? 6 : 7);
        n            = c->N * ng * nc;
        if (count != n)
            return fail(c, FANS_GPU_EINVAL, "field size mismatch (expected %lld)", (long long) n);
        if (!c->js || c0 + nc > c->js_comps) {
            std::memset(host, 0, sizeof(double) * n);
            return 0;
        }
        std::vector<double> st((size_t) nc * ng * c->nslot);
        HIPCHK(c, copy_on(c, st.data(), c->js + (size_t) c0 * ng * c->nslot, sizeof(double) * st.size(),
                          hipMemcpyDeviceToHost));
        // element (x, y, z) from its canonical slot: tile z / kStateTile, position z % kStateTile + 1
        for (int64_t e = 0; e < c->N; ++e) {
            const int     z  = (int) (e % c->nz);
            const int64_t xy = e / c->nz;
            const int64_t sl = (xy * c->ntz + z / kStateTile) * 16 + z % kStateTile + 1;
            for (int g = 0; g < ng; ++g)
                for (int k = 0; k < nc; ++k)
                    host[(e * ng + g) * nc + k] = st[((size_t) k * ng + g) * c->nslot + sl];
        }
        return 0;
    } else if (which == FANS_FIELD_PLASTIC_FLAG) {
        if (count != c->N)
            return fail(c, FANS_GPU_EINVAL, "field size mismatch (expected %lld)", (long long) c->N);
        if (int rc = check_ready(c))
            return rc;
        if (c->h != 3 || c->nstr != 6) {
            std::memset(host, 0, sizeof(double) * c->N);
            return 0;
        }
        double *df = nullptr;
        HIPCHK(c, hipMalloc(&df, sizeof(double) * c->N));
        ElemArgs<double> a{};
        a.nx           = c->nx;
        a.ny           = c->ny;
        a.nz           = c->nz;
        a.ntz          = c->ntz;
        a.geom         = c->geom;
        a.ms           = c->ms;
        a.ms16 = c->ms16;
        a.phases       = c->dph64;
        a.js           = J2State{c->js, c->nslot};
        a.u            = c->u;
        a.g0           = c->g0d;
        a.vk           = c->voxel_moduli ? c->vk64 : nullptr;
        a.vg           = c->voxel_moduli ? c->vg64 : nullptr;
        a.plastic_flag = df;
        if (c->ngp == 8)
            hipLaunchKernelGGL((k_element_pass<double, 3, 8, ELEM_FIELDS>), dim3(c->nblk_el), dim3(kThreads), 0,
                               c->stream, a);
        else
            hipLaunchKernelGGL((k_element_pass<double, 3, 1, ELEM_FIELDS>), dim3(c->nblk_el), dim3(kThreads), 0,
                               c->stream, a);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = copy_on(c, host, df, sizeof(double) * c->N, hipMemcpyDeviceToHost);
        (void) hipFree(df);
        HIPCHK(c, e);
        return 0;
    } else
        return fail(c, FANS_GPU_EINVAL, "unknown field %d", which);
    if (count != n)
        return fail(c, FANS_GPU_EINVAL, "field size mismatch (expected %lld)", (long long) n);
    // the nodal fields may go to host or device memory
    HIPCHK(c, copy_on(c, host, src, sizeof(double) * n, hipMemcpyDefault));
    return 0;
}

int fans_gpu_set_field(fans_gpu_ctx *c, int32_t which, const double *host, int64_t count)
{
    if (!c || !host)
        return FANS_GPU_EINVAL;
    double *dst = which == FANS_FIELD_U ? c->u : (which == FANS_FIELD_U_PREV ? c->u_prev : nullptr);
    if (!dst)
        return fail(c, FANS_GPU_EINVAL, "field %d is not writable", which);
    if (count != c->Nh)
        return fail(c, FANS_GPU_EINVAL, "field size mismatch");
    HIPCHK(c, copy_on(c, dst, host, sizeof(double) * count, hipMemcpyDefault));
    return 0;
}

int fans_gpu_residual(fans_gpu_ctx *c, const double *u, double *r, int32_t linear_kd)
{
    if (!c || !u || !r)
        return FANS_GPU_EINVAL;
    if (int rc = check_ready(c))
        return rc;
    if (linear_kd && !c->all_linear)
        return fail(c, FANS_GPU_EINVAL, "linear matvec needs linear materials");
    HIPCHK(c, copy_on(c, c->dB, u, sizeof(double) * c->Nh, hipMemcpyHostToDevice));
    int rc;
    if (c->op_fp32 && linear_kd) {
        // the FP32 inner solve's matvec: rnew32 = K s32, s32 = (float) u
        hipLaunchKernelGGL((k_convert<float, double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->s32,
                           c->dB, c->Nh);
        HIPCHK(c, hipGetLastError());
        if ((rc = init_scalars(c, 0.0, 0, 1)))
            return rc;
        AsmArgs<float> a = asm_args<float>(c);
        a.phases         = c->dph32;
        a.in             = c->s32;
        a.out            = c->rnew32;
        if ((rc = launch_assemble<float, ASM_CGLIN>(c, a)))
            return rc;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->rnew,
                           c->rnew32, c->Nh);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, copy_on(c, r, c->rnew, sizeof(double) * c->Nh, hipMemcpyDeviceToHost));
        return 0;
    }
    if (linear_kd) {
        // rnew = K u (the CG matvec with d = u)
        AsmArgs<double> a = asm_args<double>(c);
        a.phases          = c->dph64;
        a.in              = c->dB;
        a.out             = c->rnew;
        rc                = launch_assemble<double, ASM_CGLIN>(c, a);
    } else {
        rc = residual64(c, c->dB, c->rnew, nullptr);
    }
    if (rc)
        return rc;
    HIPCHK(c, copy_on(c, r, c->rnew, sizeof(double) * c->Nh, hipMemcpyDeviceToHost));
    return 0;
}

int fans_gpu_convolve(fans_gpu_ctx *c, const double *r, double *out)
{
    if (!c || !r || !out)
        return FANS_GPU_EINVAL;
    if (!c->have_mat)
        return fail(c, FANS_GPU_ESTATE, "materials not set");
    HIPCHK(c, copy_on(c, c->dB, r, sizeof(double) * c->Nh, hipMemcpyHostToDevice));
    if (c->op_fp32) {
        // the FP32 inner solve's convolution (FP32 FFT passes and Green operator)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->r32,
                           c->dB, c->Nh);
        HIPCHK(c, hipGetLastError());
        EpiArgs<float> ep{};
        ep.out = c->s32;
        if (int rc = convolve_dev<float>(c, c->fft32, c->r32, c->spec32, c->G32, EPI_STORE, ep, nullptr))
            return rc;
        hipLaunchKernelGGL((k_convert<double, float>), dim3(c->nblk_vec), dim3(kThreads), 0, c->stream, c->rnew,
                           c->s32, c->Nh);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, copy_on(c, out, c->rnew, sizeof(double) * c->Nh, hipMemcpyDeviceToHost));
        return 0;
    }
    EpiArgs<double> ep{};
    ep.out = c->rnew;
    int rc = convolve_dev<double>(c, c->fft64, c->dB, c->spec64, c->G64, EPI_STORE, ep, nullptr);
    if (rc)
        return rc;
    HIPCHK(c, copy_on(c, out, c->rnew, sizeof(double) * c->Nh, hipMemcpyDeviceToHost));
    return 0;
}

int fans_gpu_get_green(fans_gpu_ctx *c, double *out, int64_t count)
{
    if (!c || !out)
        return FANS_GPU_EINVAL;
    const int     np = c->h * (c->h + 1) / 2, Mc = c->nz / 2 + 1;
    const int64_t n  = c->Nf * np;
    if (count != n || !c->G64)
        return fail(c, FANS_GPU_EINVAL, "green size mismatch or not built");
    std::vector<double> st((size_t) c->gxq * c->gyq * Mc * np);
    HIPCHK(c, copy_on(c, st.data(), c->G64, sizeof(double) * st.size(), hipMemcpyDeviceToHost));
    // expand the stored frequencies to the full [nx][ny][nz/2+1] layout
    for (int ix = 0; ix < c->nx; ++ix)
        for (int iy = 0; iy < c->ny; ++iy) {
            const bool   fx = ix >= c->gxq, fy = iy >= c->gyq;
            const int    qx = fx ? c->nx - ix : ix, qy = fy ? c->ny - iy : iy;
            const double sx = fx ? -1.0 : 1.0, sy = fy ? -1.0 : 1.0;
            for (int iz = 0; iz < Mc; ++iz) {
                const double *g = &st[(((size_t) qx * c->gyq + qy) * Mc + iz) * np];
                double       *o = out + (((int64_t) ix * c->ny + iy) * Mc + iz) * np;
                if (np == 1) {
                    o[0] = g[0];
                } else {
                    const double sg[6] = {1.0, sx * sy, sx, 1.0, sy, 1.0};
                    for (int k = 0; k < 6; ++k)
                        o[k] = g[k] * sg[k];
                }
            }
        }
    return 0;
}

int fans_gpu_set_voxel_moduli(fans_gpu_ctx *c, const double *k, const double *g, int64_t count)
{
    if (c && c->nb > 1)
        return fail(c, FANS_GPU_EINVAL, "per-voxel moduli are not available on a batched context");
    if (!c || !k || (c->h == 3 && !g))
        return FANS_GPU_EINVAL;
    if (!c->have_mat)
        return fail(c, FANS_GPU_ESTATE, "set_materials must precede set_voxel_moduli");
    if (count != c->N)
        return fail(c, FANS_GPU_EINVAL, "voxel moduli: expected %lld values, got %lld", (long long) c->N,
                    (long long) count);
    if (!c->all_iso)
        return fail(c, FANS_GPU_EINVAL, "per-voxel moduli need LinearElasticIsotropic / LinearThermalIsotropic phases");
    const size_t b64 = sizeof(double) * c->N, b32 = sizeof(float) * c->N;
    if (!c->vk64) {
        HIPCHK(c, hipMalloc(&c->vk64, b64));
        HIPCHK(c, hipMalloc(&c->vg64, b64));
        HIPCHK(c, hipMalloc(&c->vk32, b32));
        HIPCHK(c, hipMalloc(&c->vg32, b32));
    }
    // k and g may be host or device arrays (the runtime tells them apart)
    HIPCHK(c, copy_on(c, c->vk64, k, b64, hipMemcpyDefault));
    HIPCHK(c, copy_on(c, c->vg64, c->h == 3 ? g : k, b64, hipMemcpyDefault));
    // input check and reference-stiffness statistics on the device
    const int nb = (int) std::min<int64_t>(1024, (c->N + kThreads - 1) / kThreads);  // partial_len >= 6 * 1024
    hipLaunchKernelGGL(k_moduli_stats, dim3(nb), dim3(kThreads), 0, c->stream, c->vk64, c->vg64, c->N, c->partial);
    HIPCHK(c, hipGetLastError());
    std::vector<double> ps((size_t) 6 * nb);
    HIPCHK(c, copy_on(c, ps.data(), c->partial, sizeof(double) * ps.size(), hipMemcpyDeviceToHost));
    double st[6] = {0.0, -INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0};
    for (int b = 0; b < nb; ++b) {
        st[0] += ps[b];
        for (int q = 1; q < 5; ++q)
            st[q] = std::max(st[q], ps[(size_t) q * nb + b]);
        st[5] += ps[(size_t) 5 * nb + b];
    }
    if (st[5] > 0.0)
        return fail(c, FANS_GPU_EINVAL, "voxel moduli must be positive (%lld voxels are not)", (long long) st[5]);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_vec(c->N)), dim3(kThreads), 0, c->stream, c->vk32,
                       c->vk64, c->N);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_vec(c->N)), dim3(kThreads), 0, c->stream, c->vg32,
                       c->vg64, c->N);
    HIPCHK(c, hipGetLastError());
    c->voxel_moduli = true;
    // reference stiffness as if every voxel were its own phase:
    // LinearElasticIsotropic (max + min) / 2 of lambda and mu (LinearElastic.h:55-69),
    // LinearThermalIsotropic mean conductivity (LinearThermal.h:19-45)
    if (!c->user_ref) {
        std::memset(c->kref, 0, sizeof(c->kref));
        if (c->h == 1) {
            const double m = st[0] / (double) count;
            for (int i = 0; i < 3; ++i)
                c->kref[i * 3 + i] = m;
        } else {
            const double lmin = -st[1], lmax = st[2], mmin = -st[3], mmax = st[4];
            const double lr = (lmax + lmin) / 2, mr = (mmax + mmin) / 2;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    c->kref[i * 6 + j] = lr;
            for (int i = 0; i < 6; ++i)
                c->kref[i * 6 + i] += 2 * mr;
        }
    }
    if (const int rcg = rebuild_green(c))
        return rcg;
    refresh_graphs(c);
    return 0;
}

int fans_gpu_sensitivity(fans_gpu_ctx *c, const double *ub, const double *g0b, double *sens, int64_t count)
{
    if (c && c->nb > 1)
        return fail(c, FANS_GPU_EINVAL, "the sensitivity is not available on a batched context");
    if (!c || !ub || !g0b || !sens)
        return FANS_GPU_EINVAL;
    if (int rc = check_ready(c))
        return rc;
    const int nq = c->h == 1 ? 1 : 2;
    if (count != c->N * nq)
        return fail(c, FANS_GPU_EINVAL, "sensitivity: expected %lld values", (long long) (c->N * nq));
    if (!c->all_iso)
        return fail(c, FANS_GPU_EUNSUP, "sensitivities are implemented for isotropic linear phases");
    if (!c->ub_d) {
        HIPCHK(c, hipMalloc(&c->ub_d, sizeof(double) * (c->Nh + 16)));
        HIPCHK(c, hipMalloc(&c->sens_d, sizeof(double) * c->N * 2));
    }
    // ub and sens may be host or device arrays; g0b is host memory
    HIPCHK(c, copy_on(c, c->ub_d, ub, sizeof(double) * c->Nh, hipMemcpyDefault));
    HIPCHK(c, copy_on(c, c->ub_d + c->Nh, g0b, sizeof(double) * c->nstr, hipMemcpyHostToDevice));
    ElemArgs<double> a{};
    a.nx     = c->nx;
    a.ny     = c->ny;
    a.nz     = c->nz;
    a.ntz    = c->ntz;
    a.geom   = c->geom;
    a.ms     = c->ms;
    a.ms16 = c->ms16;
    a.phases = c->dph64;
    a.js     = J2State{c->js, c->nslot};
    a.u      = c->u;
    a.g0     = c->g0d;
    a.ub     = c->ub_d;
    a.g0b    = c->ub_d + c->Nh;
    a.sens   = c->sens_d;
#define EP(HH, GG)                                                                                                  \
    hipLaunchKernelGGL((k_element_pass<double, HH, GG, ELEM_SENS>), dim3(c->nblk_el), dim3(kThreads), 0, c->stream, \
                       a)
    if (c->h == 1) {
        if (c->ngp == 8)
            EP(1, 8);
        else
            EP(1, 1);
    } else {
        if (c->ngp == 8)
            EP(3, 8);
        else
            EP(3, 1);
    }
#undef EP
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, copy_on(c, sens, c->sens_d, sizeof(double) * count, hipMemcpyDefault));
    return 0;
}

int fans_gpu_get_reference_stiffness(fans_gpu_ctx *c, double *out, int32_t n)
{
    if (!c || !out || n != c->nstr)
        return FANS_GPU_EINVAL;
    std::memcpy(out, c->kref, sizeof(double) * n * n);
    return 0;
}

static int fft_standalone(int dev, int nx, int ny, int nz, int h, const double *in, double *out, bool fwd)
{
    fans_gpu_desc d{};
    d.n_x = nx, d.n_y = ny, d.n_z = nz, d.howmany = h, d.n_it = 1, d.device = dev;
    d.L[0] = d.L[1] = d.L[2] = 1.0;
    fans_gpu_ctx *c = nullptr;
    int           rc = fans_gpu_create(&c, &d);
    if (rc) {
        fans_gpu_destroy(c);
        return rc;
    }
    const int64_t nr = (int64_t) nx * ny * nz * h, nc = (int64_t) nx * ny * (nz / 2 + 1) * h;
    auto         &P  = c->fft64;
    const int     Kc     = P.Mc * h;
    if (fwd) {
        rc = copy_on(c, c->dB, in, sizeof(double) * nr, hipMemcpyHostToDevice) == hipSuccess ? 0 : FANS_GPU_EHIP;
        if (!rc) {
            const ProArgs<double> pa{};
            if (h == 1)
                launch_z_r2c<double, 1, PRO_NONE>(c, P, c->dB, c->spec64, pa, nullptr);
            else
                launch_z_r2c<double, 3, PRO_NONE>(c, P, c->dB, c->spec64, pa, nullptr);
            launch_strided<double, false>(c, P.Ly, P.twy, P.bc, P.smy, c->spec64, (int64_t) Kc, 0, Kc,
                                          (int64_t) ny * Kc, nx, nullptr);
            // x pass: strided with columns = the whole (y, kz, h) block
            launch_strided<double, false>(c, P.Lx, P.twx, P.bcx, P.smx_plain, c->spec64, (int64_t) ny * Kc, 0,
                                          ny * Kc, (int64_t) 0, 1, nullptr);
            rc = copy_on(c, out, c->spec64, sizeof(cplx<double>) * nc / 1, hipMemcpyDeviceToHost) == hipSuccess
                     ? 0
                     : FANS_GPU_EHIP;
        }
    } else {
        rc = copy_on(c, c->spec64, in, sizeof(cplx<double>) * nc, hipMemcpyHostToDevice) == hipSuccess ? 0
                                                                                                      : FANS_GPU_EHIP;
        if (!rc) {
            launch_strided<double, true>(c, P.Lx, P.twx, P.bcx, P.smx_plain, c->spec64, (int64_t) ny * Kc, 0,
                                         ny * Kc, (int64_t) 0, 1, nullptr);
            launch_strided<double, true>(c, P.Ly, P.twy, P.bc, P.smy, c->spec64, (int64_t) Kc, 0, Kc,
                                         (int64_t) ny * Kc, nx, nullptr);
            EpiArgs<double> ep{};
            ep.out = c->dB;
            if (h == 1)
                launch_z_c2r<double, 1, EPI_STORE>(c, P, c->spec64, ep, nullptr);
            else
                launch_z_c2r<double, 3, EPI_STORE>(c, P, c->spec64, ep, nullptr);
            rc = copy_on(c, out, c->dB, sizeof(double) * nr, hipMemcpyDeviceToHost) == hipSuccess ? 0 : FANS_GPU_EHIP;
        }
    }
    if (!rc && hipGetLastError() != hipSuccess)
        rc = FANS_GPU_EHIP;
    fans_gpu_destroy(c);
    return rc;
}

int fans_gpu_fft_r2c(int32_t dev, int32_t nx, int32_t ny, int32_t nz, int32_t h, const double *in, double *out)
{
    if (!in || !out || (h != 1 && h != 3))
        return FANS_GPU_EINVAL;
    return fft_standalone(dev, nx, ny, nz, h, in, out, true);
}

int fans_gpu_fft_c2r(int32_t dev, int32_t nx, int32_t ny, int32_t nz, int32_t h, const double *in, double *out)
{
    if (!in || !out || (h != 1 && h != 3))
        return FANS_GPU_EINVAL;
    return fft_standalone(dev, nx, ny, nz, h, in, out, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// fans_gpu_fft_r2c_noalloc: fans_gpu_fft_r2c on caller-owned device buffers and
// the caller's stream, for callers that cannot allocate.  The only memory
// beyond input and output is the three twiddle tables, kept in `scratch`, a
// file-scope device array (each device holds its own copy for the life of the
// process).  Host code never passes `scratch` itself to a kernel: its device
// address comes from hipGetSymbolAddress, once per device with that device
// current, and is kept in g_scratch_dev.
// ---------------------------------------------------------------------------
namespace {
constexpr int kScratchCplx = 16384;  // complex FP64 twiddles: 256 KiB per device
constexpr int kScratchAlign = 32;    // table offsets in complex entries (512 B)
constexpr int kMaxDevices   = 64;
}  // namespace

static __device__ __align__(256) double scratch[2 * kScratchCplx];

namespace {
std::mutex g_scratch_mu;
double    *g_scratch_dev[kMaxDevices] = {};

int noalloc_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_noalloc_err, sizeof(g_noalloc_err), fmt, ap);
    va_end(ap);
    return code;
}

// device address of `scratch` on the current device `dev`
int scratch_address(int dev, double **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void      *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noalloc_fail(FANS_GPU_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return 0;
}

int round_up(int v, int a) { return (v + a - 1) / a * a; }

// w[k] = exp(-2 pi i k / n) for the three tables at their offsets
__global__ void k_twiddles3(cplx<double> *w, int nz, int oy, int ny, int ox, int nx)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int       n, k, o;
    if (t < nz)
        n = nz, k = t, o = 0;
    else if (t < nz + ny)
        n = ny, k = t - nz, o = oy;
    else if (t < nz + ny + nx)
        n = nx, k = t - nz - ny, o = ox;
    else
        return;
    double sn, cs;
    sincospi(-2.0 * (double) k / (double) n, &sn, &cs);
    w[o + k] = mk(cs, sn);
}
}  // namespace

extern "C" int fans_gpu_fft_r2c_noalloc(int32_t dev, int32_t nx, int32_t ny, int32_t nz, int32_t h,
                                        const double *in, double *out_cplx, void *stream)
{
    g_noalloc_err[0] = '\0';
    if (!in || !out_cplx || (h != 1 && h != 3) || nx < 1 || ny < 1 || nz < 2 || nz % 2)
        return noalloc_fail(FANS_GPU_EINVAL, "fft_r2c_noalloc: bad arguments (null buffer, h not 1 or 3, or "
                                             "n_z not even)");
    const int oy = round_up(nz, kScratchAlign), ox = oy + round_up(ny, kScratchAlign);
    if ((int64_t) ox + nx > kScratchCplx)
        return noalloc_fail(FANS_GPU_EINVAL,
                            "fft_r2c_noalloc: twiddle tables for %d x %d x %d need %lld complex entries, the "
                            "scratch array holds %d (n_x + n_y + n_z up to about %d)",
                            nx, ny, nz, (long long) ox + nx, kScratchCplx, kScratchCplx - 2 * kScratchAlign);
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0)
        return noalloc_fail(FANS_GPU_ENODEV, "no HIP device");
    if (dev < 0 || dev >= nd || dev >= kMaxDevices)
        return noalloc_fail(FANS_GPU_EINVAL, "fft_r2c_noalloc: device %d out of range", dev);
    // the host-side fields the launch helpers read: grid, stream, plan
    fans_gpu_ctx c;
    c.nx = nx, c.ny = ny, c.nz = nz, c.h = h;
    c.stream    = static_cast<hipStream_t>(stream);
    auto     &P = c.fft64;
    if (P.plan(nx, ny, nz, h) != hipSuccess)
        return noalloc_fail(FANS_GPU_EINVAL, "fft_r2c_noalloc: unsupported grid %d x %d x %d (lengths must be "
                                             "products of 2, 3 and 5 within the kernels' limits)",
                            nx, ny, nz);
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess)
        return noalloc_fail(FANS_GPU_EHIP, "hipGetDevice failed");
    if (prev != dev && hipSetDevice(dev) != hipSuccess)
        return noalloc_fail(FANS_GPU_EHIP, "hipSetDevice(%d) failed", dev);
    int     rc = 0;
    double *w  = nullptr;
    if (!(rc = scratch_address(dev, &w))) {
        auto *tw = reinterpret_cast<cplx<double> *>(w);
        P.twz = tw, P.twy = tw + oy, P.twx = tw + ox;
        const int nt = nz + ny + nx;
        hipLaunchKernelGGL(k_twiddles3, dim3((nt + 255) / 256), dim3(256), 0, c.stream, tw, nz, oy, ny, ox, nx);
        auto     *spec = reinterpret_cast<cplx<double> *>(out_cplx);
        const int Kc   = P.Mc * h;
        const ProArgs<double> pa{};
        if (h == 1)
            launch_z_r2c<double, 1, PRO_NONE>(&c, P, const_cast<double *>(in), spec, pa, nullptr);
        else
            launch_z_r2c<double, 3, PRO_NONE>(&c, P, const_cast<double *>(in), spec, pa, nullptr);
        launch_strided<double, false>(&c, P.Ly, P.twy, P.bc, P.smy, spec, (int64_t) Kc, 0, Kc, (int64_t) ny * Kc,
                                      nx, nullptr);
        launch_strided<double, false>(&c, P.Lx, P.twx, P.bcx, P.smx_plain, spec, (int64_t) ny * Kc, 0, ny * Kc,
                                      (int64_t) 0, 1, nullptr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            rc = noalloc_fail(FANS_GPU_EHIP, "fft_r2c_noalloc: launch failed: %s", hipGetErrorString(e));
    }
    if (prev != dev)
        (void) hipSetDevice(prev);
    return rc;
}

extern "C" {

int fans_gpu_stream_bandwidth(int32_t dev, int64_t bytes, int32_t reps, double *ms_per_copy)
{
    if (bytes < (1 << 20) || reps < 1 || !ms_per_copy)
        return FANS_GPU_EINVAL;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0)
        return FANS_GPU_ENODEV;
    if (dev < 0 || dev >= nd || hipSetDevice(dev) != hipSuccess)
        return FANS_GPU_EINVAL;
    const int64_t n = bytes / 16;
    double2      *a = nullptr, *b = nullptr;
    hipEvent_t    e0 = nullptr, e1 = nullptr;
    int           rc = FANS_GPU_EHIP;
    if (hipMalloc(&a, n * 16) == hipSuccess && hipMalloc(&b, n * 16) == hipSuccess &&
        hipMemset(a, 0, n * 16) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        // 8 blocks per CU-sized share of work: 2048 blocks, contiguous chunks
        const int     nb  = 2048;
        const int64_t per = (n + nb - 1) / nb;
        hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, nullptr, b, a, n, per);  // warm-up
        (void) hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_stream_copy, dim3(nb), dim3(256), 0, nullptr, (r & 1) ? a : b, (r & 1) ? b : a, n,
                               per);
        (void) hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
            hipGetLastError() == hipSuccess) {
            *ms_per_copy = ms / reps;
            rc           = 0;
        }
    }
    if (e0)
        (void) hipEventDestroy(e0);
    if (e1)
        (void) hipEventDestroy(e1);
    (void) hipFree(a);
    (void) hipFree(b);
    return rc;
}

int fans_gpu_set_option(fans_gpu_ctx *c, int32_t opt, int64_t value)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (opt == FANS_OPT_TIMING) {
        c->timing = value != 0;
        return 0;
    }
    if (opt == FANS_OPT_FFT_CHUNKS) {
        if (value < 1)
            return fail(c, FANS_GPU_EINVAL, "FFT chunk count must be >= 1");
        c->fft_chunks = (int) value;
        drop_graphs(c);
        return 0;
    }
    if (opt == FANS_OPT_GRAPHS) {
        c->use_graphs = value != 0;
        return 0;
    }
    if (opt == FANS_OPT_ASM_PLANES) {
        // x planes per assembly workgroup (tile shape of the stencil pass)
        if (value < 1 || value > c->nx)
            return fail(c, FANS_GPU_EINVAL, "assembly planes per block must be in [1, n_x]");
        c->cx           = (int) value;
        const int nbx   = (c->nx + c->cx - 1) / c->cx;
        c->nblk_asm     = nbx * c->nty * c->ntz;
        drop_graphs(c);
        return ensure_partials(c);
    }
    if (opt == FANS_OPT_MIXED_RESTART) {
        c->mixed_restart = value != 0;
        return 0;
    }
    if (opt == FANS_OPT_FFT_LDS_KIB) {
        // LDS budget of the generic FFT passes' workgroups (lengths other
        // than the powers of two 16..256, whose kernels fix their tiles)
        if (value < 8 || value > 64)
            return fail(c, FANS_GPU_EINVAL, "FFT LDS budget must be in [8, 64] KiB");
        c->fft64.lds_kib = (int) value;
        c->fft32.lds_kib = (int) value;
        if (c->fft64.plan(c->nx, c->ny, c->nz, c->h) != hipSuccess ||
            (c->fft32.twz && c->fft32.plan(c->nx, c->ny, c->nz, c->h) != hipSuccess))
            return fail(c, FANS_GPU_EINVAL, "FFT plan failed");
        drop_graphs(c);
        return ensure_partials(c);
    }
    if (opt == FANS_OPT_OPERATOR_FP32) {
        if (value && !c->r32)
            return fail(c, FANS_GPU_EINVAL, "FP32 operators need a context created with FANS_PREC_MIXED");
        c->op_fp32 = value != 0;
        return 0;
    }
    return fail(c, FANS_GPU_EINVAL, "unknown option %d", opt);
}

const char *fans_gpu_kernel_family_name(int32_t k) { return (k >= 0 && k < KF_COUNT) ? kFamilyNames[k] : ""; }

int fans_gpu_kernel_times(fans_gpu_ctx *c, double *out, int32_t cap, int32_t *count)
{
    if (!c)
        return FANS_GPU_EINVAL;
    if (count)
        *count = KF_COUNT;
    for (int k = 0; k < KF_COUNT && 2 * k + 1 < cap; ++k) {
        out[2 * k]     = c->fam_ms[k];
        out[2 * k + 1] = c->fam_n[k];
    }
    return 0;
}

}  // extern "C"
