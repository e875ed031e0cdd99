"""ctypes binding of libfans_gpu.so (include/fans_gpu.h).

There is no CPU fallback: loading fails loudly when the HIP library is not
built, and creating a solver fails when no HIP device is present.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .problem import MODEL_KIND, Problem

_HERE = os.path.dirname(os.path.abspath(__file__))
# FANS_GPU_LIB selects another build of the library (e.g. a variant for A/B timing)
LIB_PATH = os.environ.get("FANS_GPU_LIB") or os.path.join(_HERE, "libfans_gpu.so")
_LIB = None

FE = {"HEX8": 0, "HEX8R": 1, "BBAR": 2}
METHOD = {"cg": 0, "fp": 1}
MEASURE = {"L1": 0, "L2": 1, "Linfinity": 2}
ERR_TYPE = {"absolute": 0, "relative": 1}
PRECISION = {"fp64": 0, "mixed": 1}
FIELD = {"u": 0, "r": 1, "u_prev": 2, "plastic_strain": 3, "hardening": 4, "plastic_flag": 5,
         "kinematic_hardening": 6}
OPT_TIMING = 1
OPT_FFT_CHUNKS = 2
OPT_GRAPHS = 3
OPT_ASM_PLANES = 4
OPT_OPERATOR_FP32 = 5
OPT_FFT_LDS_KIB = 6
OPT_MIXED_RESTART = 7
ERRORS = {-1: "EINVAL", -2: "ENODEV", -3: "EHIP", -4: "ESTATE", -5: "EUNSUP"}


class FansGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Desc(C.Structure):
    _fields_ = [
        ("n_x", C.c_int32), ("n_y", C.c_int32), ("n_z", C.c_int32), ("howmany", C.c_int32),
        ("fe_type", C.c_int32), ("method", C.c_int32), ("measure", C.c_int32), ("err_type", C.c_int32),
        ("n_it", C.c_int32), ("ls_max_iter", C.c_int32), ("L", C.c_double * 3), ("tol", C.c_double),
        ("ls_tol", C.c_double), ("precision", C.c_int32), ("device", C.c_int32),
        ("large_strain", C.c_int32), ("reserved", C.c_int32 * 5),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("inner_iterations", C.c_int32), ("outer_iterations", C.c_int32),
        ("converged", C.c_int32), ("solve_ms", C.c_double), ("final_error", C.c_double),
    ]


EXPORTS = [
    "fans_gpu_abi_version", "fans_gpu_last_error", "fans_gpu_create", "fans_gpu_destroy",
    "fans_gpu_set_microstructure", "fans_gpu_set_materials", "fans_gpu_set_gradient", "fans_gpu_get_gradient",
    "fans_gpu_enable_mixed_bc", "fans_gpu_disable_mixed_bc", "fans_gpu_solve", "fans_gpu_error_history",
    "fans_gpu_homogenized_stress", "fans_gpu_homogenized_tangent", "fans_gpu_extrapolate_displacement",
    "fans_gpu_postprocess", "fans_gpu_get_field", "fans_gpu_set_field", "fans_gpu_residual", "fans_gpu_convolve",
    "fans_gpu_get_green", "fans_gpu_get_reference_stiffness", "fans_gpu_fft_r2c", "fans_gpu_fft_c2r",
    "fans_gpu_fft_r2c_noalloc", "fans_gpu_stream_bandwidth",
    "fans_gpu_set_option", "fans_gpu_kernel_times", "fans_gpu_kernel_family_name", "fans_gpu_reset_state",
    "fans_gpu_set_voxel_moduli", "fans_gpu_sensitivity", "fans_gpu_create_batch", "fans_gpu_select_slot",
    "fans_gpu_batch_results", "fans_gpu_batch_homogenized_stress",
]


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run python -m fans_amd.build")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, i64 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.c_int64
        L.fans_gpu_last_error.restype = C.c_char_p
        L.fans_gpu_last_error.argtypes = [vp]
        L.fans_gpu_create.argtypes = [C.POINTER(vp), C.POINTER(Desc)]
        if hasattr(L, "fans_gpu_create_batch"):  # (FANS_GPU_LIB may name an older build)
            L.fans_gpu_create_batch.argtypes = [C.POINTER(vp), C.POINTER(Desc), C.c_int32]
            L.fans_gpu_select_slot.argtypes = [vp, C.c_int32]
            L.fans_gpu_batch_results.argtypes = [vp, ip, ip, dp, C.c_int32]
            L.fans_gpu_batch_homogenized_stress.argtypes = [vp, dp, C.c_int32, C.c_int32]
        L.fans_gpu_destroy.argtypes = [vp]
        L.fans_gpu_set_microstructure.argtypes = [vp, C.POINTER(C.c_uint16), i64]
        L.fans_gpu_set_materials.argtypes = [vp, C.c_int32, ip, ip, ip, dp, i64, dp]
        L.fans_gpu_set_gradient.argtypes = [vp, dp, C.c_int32]
        L.fans_gpu_get_gradient.argtypes = [vp, dp, C.c_int32]
        L.fans_gpu_enable_mixed_bc.argtypes = [vp, C.c_int32, ip, C.c_int32, ip, C.c_int32, dp, dp, C.c_int32]
        L.fans_gpu_disable_mixed_bc.argtypes = [vp]
        L.fans_gpu_solve.argtypes = [vp, C.POINTER(Stats)]
        L.fans_gpu_error_history.argtypes = [vp, dp, C.c_int32, ip]
        L.fans_gpu_homogenized_stress.argtypes = [vp, dp, C.c_int32]
        L.fans_gpu_homogenized_tangent.argtypes = [vp, C.c_double, dp, C.c_int32]
        L.fans_gpu_extrapolate_displacement.argtypes = [vp]
        L.fans_gpu_reset_state.argtypes = [vp]
        L.fans_gpu_postprocess.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, ip, C.c_int32]
        L.fans_gpu_get_field.argtypes = [vp, C.c_int32, dp, i64]
        L.fans_gpu_set_field.argtypes = [vp, C.c_int32, dp, i64]
        L.fans_gpu_residual.argtypes = [vp, dp, dp, C.c_int32]
        L.fans_gpu_convolve.argtypes = [vp, dp, dp]
        L.fans_gpu_get_green.argtypes = [vp, dp, i64]
        L.fans_gpu_get_reference_stiffness.argtypes = [vp, dp, C.c_int32]
        for f in (L.fans_gpu_fft_r2c, L.fans_gpu_fft_c2r):
            f.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp]
        L.fans_gpu_set_option.argtypes = [vp, C.c_int32, C.c_int64]
        L.fans_gpu_kernel_times.argtypes = [vp, dp, C.c_int32, ip]
        L.fans_gpu_kernel_family_name.restype = C.c_char_p
        L.fans_gpu_kernel_family_name.argtypes = [C.c_int32]
        L.fans_gpu_set_voxel_moduli.argtypes = [vp, dp, dp, i64]
        L.fans_gpu_sensitivity.argtypes = [vp, dp, dp, dp, i64]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _on_device(x):
    """True for a torch tensor in GPU memory (the library then reads or writes it
    in place; torch must have been imported before the library was loaded so
    that both share one HIP runtime)."""
    return bool(getattr(x, "is_cuda", False))


def _hip_runtimes():
    """Files of the HIP runtimes mapped into this process (Linux)."""
    try:
        with open("/proc/self/maps") as f:
            return {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    except OSError:
        return set()


def _dev_dp(t, n):
    """Device pointer of a contiguous float64 torch tensor of n values, after
    torch's queued work on it has finished (the library uses its own stream).
    Refuses when torch and the library run on two different HIP runtimes (the
    library was loaded before torch): the pointer would mean nothing to it."""
    import torch

    lib()
    if len(_hip_runtimes()) > 1:
        raise RuntimeError("torch and libfans_gpu.so use different HIP runtimes; import torch before the "
                           "library is first used to pass GPU tensors")
    if t.dtype != torch.float64 or not t.is_contiguous() or t.numel() != n:
        raise ValueError(f"device array: need {n} contiguous float64 values")
    torch.cuda.synchronize(t.device)
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def stream_bandwidth(nbytes=1 << 30, reps=20, device=0):
    """(ms per copy, bytes moved per copy) of the library's streaming copy."""
    L = lib()
    L.fans_gpu_stream_bandwidth.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    ms = C.c_double()
    rc = L.fans_gpu_stream_bandwidth(device, int(nbytes), int(reps), C.byref(ms))
    if rc:
        raise FansGpuError(rc, "fans_gpu_stream_bandwidth failed")
    return ms.value, 2 * (int(nbytes) // 16) * 16


def fft_r2c(x, device=0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny, nz, h = x.shape
    out = np.zeros((nx, ny, nz // 2 + 1, h), dtype=np.complex128)
    rc = lib().fans_gpu_fft_r2c(device, nx, ny, nz, h, _dp(x), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise FansGpuError(rc, "fans_gpu_fft_r2c failed")
    return out


def fft_c2r(X, nz, device=0):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    nx, ny, nc, h = X.shape
    out = np.zeros((nx, ny, nz, h))
    rc = lib().fans_gpu_fft_c2r(device, nx, ny, nz, h, X.ctypes.data_as(C.POINTER(C.c_double)), _dp(out))
    if rc:
        raise FansGpuError(rc, "fans_gpu_fft_c2r failed")
    return out


# settings fans_gpu_create fixes for the life of a context (grid, physics,
# element weights and inverse lengths, solver and error settings)
_CONTEXT_FIELDS = ("dims", "howmany", "strain_type", "L", "fe_type", "method", "measure", "err_type", "tol",
                   "n_it", "ls_max_iter", "ls_tol")


def context_key(problem: Problem):
    """The part of a problem a context is created for: problems with equal keys
    can share one context through GpuSolver.load."""
    return (tuple(int(v) for v in problem.dims), int(problem.howmany), problem.strain_type,
            tuple(float(v) for v in problem.L), problem.fe_type, problem.method, problem.measure,
            problem.err_type, float(problem.tol), int(problem.n_it), int(problem.ls_max_iter),
            float(problem.ls_tol))


class GpuSolver:
    """One load case on the GPU: the reference's (MaterialManager, Solver) pair."""

    def __init__(self, problem: Problem, precision="fp64", device=0, timing=False, slots=1):
        self.p = problem
        self.L = lib()
        self.h = problem.howmany
        self.nstr = problem.n_str
        self.dims = tuple(problem.dims)
        self.N = int(np.prod(self.dims))
        self.ngp = 1 if problem.fe_type == "HEX8R" else 8
        d = Desc()
        d.n_x, d.n_y, d.n_z = self.dims
        d.howmany = self.h
        d.large_strain = int(self.nstr == 9)
        d.fe_type = FE[problem.fe_type]
        d.method = METHOD[problem.method]
        d.measure = MEASURE[problem.measure]
        d.err_type = ERR_TYPE[problem.err_type]
        d.n_it = problem.n_it
        d.ls_max_iter = problem.ls_max_iter
        d.L[:] = list(problem.L)
        d.tol = problem.tol
        d.ls_tol = problem.ls_tol
        d.precision = PRECISION[precision]
        d.device = device
        self.ptr = C.c_void_p()
        if slots > 1:
            rc = self.L.fans_gpu_create_batch(C.byref(self.ptr), C.byref(d), int(slots))
        else:
            rc = self.L.fans_gpu_create(C.byref(self.ptr), C.byref(d))
        if rc:
            msg = self._err()
            self.close()
            raise FansGpuError(rc, msg or "fans_gpu_create failed")
        if timing:
            self._check(self.L.fans_gpu_set_option(self.ptr, OPT_TIMING, 1))
        self.stats = Stats()
        self._key = context_key(problem)
        self.load(problem)

    def load(self, problem: Problem):
        """Microstructure and material tables of `problem` (same grid, physics and
        solver settings as the context was created with); fresh solver state."""
        if context_key(problem) != self._key:
            diff = [n for n, a, b in zip(_CONTEXT_FIELDS, context_key(problem), self._key) if a != b]
            raise ValueError("load(): the problem differs from the context in " + ", ".join(diff) +
                             " (fixed when the context was created)")
        self.p = problem
        ms = np.ascontiguousarray(problem.ms, dtype=np.uint16).reshape(-1)
        self._check(self.L.fans_gpu_set_microstructure(self.ptr, ms.ctypes.data_as(C.POINTER(C.c_uint16)), ms.size))
        kinds, nph, pids, params = problem.flat_materials()
        self._ref = None if problem.ref_material is None else np.ascontiguousarray(problem.ref_material, np.float64)
        self._check(self.L.fans_gpu_set_materials(self.ptr, len(kinds), _ip(kinds), _ip(nph), _ip(pids), _dp(params),
                                                  params.size, _dp(self._ref)))
        self._check(self.L.fans_gpu_reset_state(self.ptr))

    def _err(self):
        if self.ptr:
            s = self.L.fans_gpu_last_error(self.ptr)
            return s.decode() if s else ""
        return ""

    def _check(self, rc):
        if rc:
            raise FansGpuError(rc, self._err())

    def close(self):
        if getattr(self, "ptr", None):
            self.L.fans_gpu_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- reference Solver surface ---------------------------------------------
    def set_gradient(self, g0):
        g = np.ascontiguousarray(g0, dtype=np.float64)
        self._check(self.L.fans_gpu_set_gradient(self.ptr, _dp(g), g.size))

    def macro_gradient(self):
        g = np.zeros(self.nstr)
        self._check(self.L.fans_gpu_get_gradient(self.ptr, _dp(g), self.nstr))
        return g

    def enable_mixed_bc(self, mbc, step):
        iE = np.ascontiguousarray(mbc["strain_indices"], dtype=np.int32)
        iF = np.ascontiguousarray(mbc["stress_indices"], dtype=np.int32)
        FE_ = np.ascontiguousarray(mbc["strain_path"], dtype=np.float64).reshape(-1)
        PF_ = np.ascontiguousarray(mbc["stress_path"], dtype=np.float64).reshape(-1)
        self._check(self.L.fans_gpu_enable_mixed_bc(self.ptr, iE.size, _ip(iE) if iE.size else None, iF.size,
                                                     _ip(iF) if iF.size else None, int(mbc["n_steps"]),
                                                     _dp(FE_) if FE_.size else None,
                                                     _dp(PF_) if PF_.size else None, int(step)))

    def disable_mixed_bc(self):
        self._check(self.L.fans_gpu_disable_mixed_bc(self.ptr))

    def solve(self):
        self._check(self.L.fans_gpu_solve(self.ptr, C.byref(self.stats)))
        return self.stats.iterations

    def error_history(self):
        n = C.c_int32()
        self._check(self.L.fans_gpu_error_history(self.ptr, None, 0, C.byref(n)))
        out = np.zeros(n.value)
        self._check(self.L.fans_gpu_error_history(self.ptr, _dp(out), n.value, C.byref(n)))
        return out

    def homogenized_stress(self):
        out = np.zeros(self.nstr)
        self._check(self.L.fans_gpu_homogenized_stress(self.ptr, _dp(out), self.nstr))
        return out

    def homogenized_tangent(self, pert=1e-6):
        out = np.zeros((self.nstr, self.nstr))
        self._check(self.L.fans_gpu_homogenized_tangent(self.ptr, float(pert), _dp(out), self.nstr))
        return out

    def extrapolate_displacement(self):
        self._check(self.L.fans_gpu_extrapolate_displacement(self.ptr))

    def reset_state(self):
        self._check(self.L.fans_gpu_reset_state(self.ptr))

    def set_timing(self, on=True):
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_TIMING, int(bool(on))))

    def set_fft_chunks(self, k):
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_FFT_CHUNKS, int(k)))

    def set_asm_planes(self, p):
        """x planes per assembly workgroup (stencil tile depth)."""
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_ASM_PLANES, int(p)))

    def set_fft_lds_kib(self, kib):
        """LDS budget (KiB) of one workgroup of the generic FFT passes (tile
        width of lengths other than the powers of two 16..256)."""
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_FFT_LDS_KIB, int(kib)))

    def set_mixed_restart(self, on=True):
        """Mixed form: restart the FP32 CG in every outer round (the batched
        contexts' algorithm) instead of replacing its residual."""
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_MIXED_RESTART, int(bool(on))))

    def set_operator_fp32(self, on=True):
        """convolve() and residual(linear_kd=True) run the mixed form's FP32
        kernels (needs precision="mixed")."""
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_OPERATOR_FP32, int(bool(on))))

    def set_graphs(self, on=True):
        """Replay captured CG iterations (hipGraph; default) or launch eagerly."""
        self._check(self.L.fans_gpu_set_option(self.ptr, OPT_GRAPHS, int(bool(on))))

    def get_field(self, name):
        if name in ("plastic_strain", "kinematic_hardening"):
            out = np.zeros((self.N, self.ngp, 6))
        elif name == "hardening":
            out = np.zeros((self.N, self.ngp))
        elif name == "plastic_flag":
            out = np.zeros(self.N)
        else:
            out = np.zeros(self.dims + (self.h,))
        self._check(self.L.fans_gpu_get_field(self.ptr, FIELD[name], _dp(out), out.size))
        return out

    def get_u(self, out=None):
        """u as a numpy array, or written into `out` (a float64 torch tensor in GPU
        memory with N * h values: a device-to-device copy)."""
        if out is not None and _on_device(out):
            self._check(self.L.fans_gpu_get_field(self.ptr, FIELD["u"], _dev_dp(out, self.N * self.h), out.numel()))
            return out
        return self.get_field("u")

    def get_r(self):
        return self.get_field("r")

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self._check(self.L.fans_gpu_set_field(self.ptr, FIELD["u"], _dp(u), u.size))

    def get_state(self):
        return self.get_field("plastic_strain"), self.get_field("hardening")

    def plastic_flag(self):
        """PseudoPlastic::postprocess plastic_flag at the current u (element means)."""
        return self.get_field("plastic_flag")

    def element_fields(self, n_slots=None, gp=True):
        ns, ng = self.nstr, self.ngp
        n_slots = n_slots or (int(np.max(self.p.ms)) + 1)
        out = dict(
            strain=np.zeros((self.N, ns)), stress=np.zeros((self.N, ns)),
            strain_gp=np.zeros((self.N, ng, ns)) if gp else None, stress_gp=np.zeros((self.N, ng, ns)) if gp else None,
            stress_average=np.zeros(ns), strain_average=np.zeros(ns),
            phase_stress_average=np.zeros((n_slots, ns)), phase_strain_average=np.zeros((n_slots, ns)),
            phase_counts=np.zeros(n_slots, dtype=np.int32))
        self._check(self.L.fans_gpu_postprocess(
            self.ptr, _dp(out["strain"]), _dp(out["stress"]), _dp(out["strain_gp"]), _dp(out["stress_gp"]),
            _dp(out["stress_average"]), _dp(out["strain_average"]), _dp(out["phase_stress_average"]),
            _dp(out["phase_strain_average"]), _ip(out["phase_counts"]), n_slots))
        return out

    # --- single operators -----------------------------------------------------
    def residual(self, u, linear_kd=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.dims + (self.h,))
        self._check(self.L.fans_gpu_residual(self.ptr, _dp(u), _dp(r), int(linear_kd)))
        return r

    def convolve(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(self.dims + (self.h,))
        self._check(self.L.fans_gpu_convolve(self.ptr, _dp(r), _dp(out)))
        return out

    def green(self):
        nx, ny, nz = self.dims
        npk = self.h * (self.h + 1) // 2
        out = np.zeros((nx, ny, nz // 2 + 1, npk))
        self._check(self.L.fans_gpu_get_green(self.ptr, _dp(out), out.size))
        return out

    def set_voxel_moduli(self, k, g=None):
        """Per-voxel isotropic moduli [x][y][z]: bulk k and shear g (mechanics) or
        conductivity k (thermal); rebuilds the reference stiffness and Green operator.
        k, g: numpy arrays, or float64 torch tensors in GPU memory (copied on the
        device, nothing crosses to the host)."""
        N = int(np.prod(self.p.dims))
        if _on_device(k):
            kp = _dev_dp(k.reshape(-1), N)
            gp = None if g is None or self.p.howmany == 1 else _dev_dp(g.reshape(-1), N)
            self._check(self.L.fans_gpu_set_voxel_moduli(self.ptr, kp, gp, N))
            return
        kk = np.ascontiguousarray(np.asarray(k, dtype=np.float64).reshape(-1))
        gg = None if g is None else np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1))
        if kk.size != N or (gg is not None and gg.size != N):
            raise ValueError(f"voxel moduli need {N} values each")
        self._check(self.L.fans_gpu_set_voxel_moduli(self.ptr, _dp(kk), None if gg is None else _dp(gg), N))

    def sensitivity(self, ub, g0b):
        """Per-element Gauss integrals of eps_a^T dC/dp eps_b (p = K, G; thermal:
        k) between the current solution and (ub, g0b); [N, 2] or [N, 1].  With ub a
        float64 torch tensor in GPU memory the result is one too (no host copy)."""
        N = int(np.prod(self.p.dims))
        nq = 1 if self.p.howmany == 1 else 2
        if _on_device(ub):
            import torch

            gb = np.ascontiguousarray(np.asarray(g0b, dtype=np.float64).reshape(-1))
            if gb.size != self.p.n_str:
                raise ValueError("sensitivity: gradient size mismatch")
            out = torch.empty((N, nq), dtype=torch.float64, device=ub.device)
            self._check(self.L.fans_gpu_sensitivity(self.ptr, _dev_dp(ub.reshape(-1), N * self.p.howmany), _dp(gb),
                                                    _dev_dp(out, N * nq), out.numel()))
            return out
        ubb = np.ascontiguousarray(np.asarray(ub, dtype=np.float64).reshape(-1))
        gb = np.ascontiguousarray(np.asarray(g0b, dtype=np.float64).reshape(-1))
        if ubb.size != N * self.p.howmany or gb.size != self.p.n_str:
            raise ValueError("sensitivity: field or gradient size mismatch")
        out = np.empty(N * nq)
        self._check(self.L.fans_gpu_sensitivity(self.ptr, _dp(ubb), _dp(gb), _dp(out), out.size))
        return out.reshape(N, nq)

    def reference_stiffness(self):
        out = np.zeros((self.nstr, self.nstr))
        self._check(self.L.fans_gpu_get_reference_stiffness(self.ptr, _dp(out), self.nstr))
        return out

    def kernel_times(self):
        buf = np.zeros(64)
        n = C.c_int32()
        self._check(self.L.fans_gpu_kernel_times(self.ptr, _dp(buf), buf.size, C.byref(n)))
        res = {}
        for k in range(n.value):
            name = self.L.fans_gpu_kernel_family_name(k).decode()
            res[name] = (float(buf[2 * k]), int(buf[2 * k + 1]))
        return res


class GpuBatch(GpuSolver):
    """Several RVEs of one creation setting (context_key) in one batched
    context: solve() runs every RVE at once, each kernel launch covering every
    RVE still iterating, with per-RVE results equal to solving each alone.
    The GpuSolver methods address the selected RVE (select())."""

    def __init__(self, problems, precision="mixed", device=0):
        problems = list(problems)
        if not problems:
            raise ValueError("GpuBatch needs at least one problem")
        self.slots = len(problems)
        self.problems = list(problems)
        super().__init__(problems[0], precision=precision, device=device, slots=self.slots)
        for b in range(1, self.slots):
            self.load_slot(b, problems[b])
        self.select(0)

    def select(self, slot):
        self._check(self.L.fans_gpu_select_slot(self.ptr, int(slot)))
        self.p = self.problems[slot]

    def load_slot(self, slot, problem, g0=None):
        """Microstructure, materials (fresh state) and optionally the gradient of
        RVE `slot`; leaves it selected."""
        self.select(slot)
        self.load(problem)
        self.problems[slot] = problem
        if g0 is not None:
            self.set_gradient(g0)

    def results(self):
        """Per RVE of the last solve: iterations (mixed form: FP32 inner
        iterations), converged flags, final errors."""
        it = np.zeros(self.slots, dtype=np.int32)
        cv = np.zeros(self.slots, dtype=np.int32)
        fe = np.zeros(self.slots)
        self._check(self.L.fans_gpu_batch_results(self.ptr, _ip(it), _ip(cv), _dp(fe), self.slots))
        return it, cv.astype(bool), fe

    def stresses(self):
        """Homogenized stress (flux) of every RVE, [slots, n_str]."""
        out = np.zeros((self.slots, self.nstr))
        self._check(self.L.fans_gpu_batch_homogenized_stress(self.ptr, _dp(out), self.nstr, self.slots))
        return out
