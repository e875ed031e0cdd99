"""ctypes binding of the CPU FP64 restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the parity checker.  The product (fans_amd)
never imports this package.

The API mirrors the reference's Solver<howmany,n_str> surface for one load case
(include/solver.h:11-103, include/solverCG.h, include/solverFP.h,
include/mixedBCs.h:150-227): set_gradient / enable_mixed_bc / solve /
homogenized_stress / homogenized_tangent / extrapolate_displacement.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FE = {"HEX8": 0, "HEX8R": 1, "BBAR": 2}
METHOD = {"cg": 0, "fp": 1}
MEASURE = {"L1": 0, "L2": 1, "Linfinity": 2}
ERR_TYPE = {"absolute": 0, "relative": 1}
MODEL = {
    "LinearThermalIsotropic": 1,
    "LinearThermalTriclinic": 2,
    "LinearElasticIsotropic": 3,
    "LinearElasticTriclinic": 4,
    "J2PlasticityNew_LinearIsotropicHardening": 5,
    "PseudoPlasticLinearHardening": 6,
    "SaintVenantKirchhoff": 7,
    "CompressibleNeoHookean": 8,
    "PseudoPlasticNonLinearHardening": 9,
}


class OrDesc(C.Structure):
    _fields_ = [
        ("n_x", C.c_int32), ("n_y", C.c_int32), ("n_z", C.c_int32),
        ("howmany", C.c_int32), ("fe_type", C.c_int32), ("method", C.c_int32),
        ("measure", C.c_int32), ("err_type", C.c_int32), ("n_it", C.c_int32),
        ("ls_max_iter", C.c_int32), ("L", C.c_double * 3), ("tol", C.c_double),
        ("ls_tol", C.c_double), ("large_strain", C.c_int32),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(OrDesc)]
        L.or_destroy.argtypes = [vp]
        L.or_set_microstructure.argtypes = [vp, C.POINTER(C.c_uint16)]
        L.or_set_materials.argtypes = [vp, C.c_int, ip, ip, ip, dp, dp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_gradient.argtypes = [vp, dp]
        L.or_enable_mixed_bc.argtypes = [vp, C.c_int, ip, C.c_int, ip, C.c_int, dp, dp, C.c_int]
        L.or_disable_mixed_bc.argtypes = [vp]
        L.or_mixed_active.argtypes = [vp]
        L.or_get_macro_gradient.argtypes = [vp, dp]
        L.or_solve.argtypes = [vp]
        L.or_error_history.argtypes = [vp, dp, C.c_int]
        L.or_homogenized_stress.argtypes = [vp, dp]
        L.or_homogenized_tangent.argtypes = [vp, C.c_double, dp]
        L.or_extrapolate_displacement.argtypes = [vp]
        for f in ("or_get_u", "or_set_u", "or_get_r", "or_get_green", "or_get_reference_stiffness",
                  "or_get_ref_element_stiffness", "or_get_B"):
            getattr(L, f).argtypes = [vp, dp]
        L.or_get_state.argtypes = [vp, dp, dp, C.c_int]
        L.or_element_fields.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, ip, C.c_int]
        L.or_plastic_flag.argtypes = [vp, dp]
        L.or_residual.argtypes = [vp, dp, dp, C.c_int]
        L.or_convolve.argtypes = [vp, dp, dp]
        L.or_get_phase_stiffness.argtypes = [vp, C.c_int, dp]
        L.or_n_gp.argtypes = [vp]
        L.or_fft_r2c.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.or_fft_c2r.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
        L.or_material_point.argtypes = [C.c_int, dp, dp, dp, C.c_double, dp, dp, dp]
        L.or_visco_point.argtypes = [C.c_int, dp, dp, dp, C.c_double, dp, dp, dp, dp, dp]
        L.or_get_kinematic_state.argtypes = [vp, dp, C.c_int]
        L.or_model_nparams.argtypes = [C.c_int]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def set_threads(n):
    lib().or_set_threads(int(n))


def fft_r2c(x):
    """x: real [nx,ny,nz,h] -> complex [nx,ny,nz//2+1,h] (unnormalised)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    nx, ny, nz, h = x.shape
    out = np.zeros((nx, ny, nz // 2 + 1, h), dtype=np.complex128)
    lib().or_fft_r2c(nx, ny, nz, h, _dp(x), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def fft_c2r(X, nz):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    nx, ny, nc, h = X.shape
    out = np.zeros((nx, ny, nz, h))
    lib().or_fft_c2r(nx, ny, nz, h, X.ctypes.data_as(C.POINTER(C.c_double)), _dp(out))
    return out


def material_point(kind, params, eps, ep_t=None, q_t=0.0):
    kind = MODEL[kind] if isinstance(kind, str) else kind
    params = np.ascontiguousarray(params, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    ep_t = np.zeros(6) if ep_t is None else np.ascontiguousarray(ep_t, dtype=np.float64)
    sig = np.zeros(len(eps))
    ep = np.zeros(6)
    q = np.zeros(1)
    lib().or_material_point(kind, _dp(params), _dp(eps), _dp(ep_t), float(q_t), _dp(sig), _dp(ep), _dp(q))
    return sig, ep, float(q[0])


def visco_point(kind, params, eps, ep_t=None, psi_t=0.0, pb_t=None, psi=0.0, pb=None):
    """One J2Plasticity::get_sigma call: (sigma, ep, psi, psi_bar) with psi and
    psi_bar accumulated from the given current values."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    eps = np.ascontiguousarray(eps, dtype=np.float64)
    ep_t = np.zeros(6) if ep_t is None else np.ascontiguousarray(ep_t, dtype=np.float64)
    pb_t = np.zeros(6) if pb_t is None else np.ascontiguousarray(pb_t, dtype=np.float64)
    pb = np.zeros(6) if pb is None else np.array(pb, dtype=np.float64)
    sig, ep, ps = np.zeros(6), np.zeros(6), np.array([float(psi)])
    lib().or_visco_point(kind, _dp(params), _dp(eps), _dp(ep_t), float(psi_t), _dp(pb_t), _dp(sig), _dp(ep),
                         _dp(ps), _dp(pb))
    return sig, ep, float(ps[0]), pb


class Oracle:
    """One load case of the reference solver on the CPU (a Solver + MaterialManager pair)."""

    def __init__(self, problem):
        self.p = problem
        d = OrDesc()
        d.n_x, d.n_y, d.n_z = problem.dims
        d.howmany = problem.howmany
        d.fe_type = FE[problem.fe_type]
        d.method = METHOD[problem.method]
        d.measure = MEASURE[problem.measure]
        d.err_type = ERR_TYPE[problem.err_type]
        d.n_it = problem.n_it
        d.ls_max_iter = problem.ls_max_iter
        d.L[:] = list(problem.L)
        d.tol = problem.tol
        d.ls_tol = problem.ls_tol
        d.large_strain = int(problem.n_str == 9)
        self.L = lib()
        self.h = problem.howmany
        self.nstr = problem.n_str
        self.dims = tuple(problem.dims)
        self.N = int(np.prod(self.dims))
        self.ptr = self.L.or_create(C.byref(d))
        if not self.ptr:
            raise ValueError("or_create rejected the problem description")
        ms = np.ascontiguousarray(problem.ms, dtype=np.uint16).reshape(-1)
        self.L.or_set_microstructure(self.ptr, ms.ctypes.data_as(C.POINTER(C.c_uint16)))
        kinds, nph, pids, params = problem.flat_materials()
        ref = problem.ref_material
        refp = _dp(np.ascontiguousarray(ref, dtype=np.float64)) if ref is not None else None
        if ref is not None:
            self._ref = np.ascontiguousarray(ref, dtype=np.float64)
            refp = _dp(self._ref)
        rc = self.L.or_set_materials(self.ptr, len(kinds), _ip(kinds), _ip(nph), _ip(pids), _dp(params), refp)
        if rc != 0:
            raise ValueError(f"or_set_materials failed ({rc})")
        self.ngp = self.L.or_n_gp(self.ptr)
        self.iters = 0

    def __del__(self):
        try:
            if self.ptr:
                self.L.or_destroy(self.ptr)
                self.ptr = None
        except Exception:
            pass

    # --- reference Solver surface -------------------------------------------------
    def set_gradient(self, g0):
        g = np.ascontiguousarray(g0, dtype=np.float64)
        self.L.or_set_gradient(self.ptr, _dp(g))

    def enable_mixed_bc(self, mbc, step):
        iE = np.ascontiguousarray(mbc["strain_indices"], dtype=np.int32)
        iF = np.ascontiguousarray(mbc["stress_indices"], dtype=np.int32)
        FE_ = np.ascontiguousarray(mbc["strain_path"], dtype=np.float64).reshape(-1)
        PF_ = np.ascontiguousarray(mbc["stress_path"], dtype=np.float64).reshape(-1)
        n_steps = int(mbc["n_steps"])
        rc = self.L.or_enable_mixed_bc(self.ptr, len(iE), _ip(iE), len(iF), _ip(iF), n_steps,
                                       _dp(FE_ if FE_.size else np.zeros(1)),
                                       _dp(PF_ if PF_.size else np.zeros(1)), int(step))
        if rc != 0:
            raise ValueError("or_enable_mixed_bc failed")

    def disable_mixed_bc(self):
        self.L.or_disable_mixed_bc(self.ptr)

    def macro_gradient(self):
        g = np.zeros(self.nstr)
        self.L.or_get_macro_gradient(self.ptr, _dp(g))
        return g

    def solve(self):
        self.iters = self.L.or_solve(self.ptr)
        return self.iters

    def error_history(self):
        out = np.zeros(self.p.n_it + 1)
        n = self.L.or_error_history(self.ptr, _dp(out), len(out))
        return out[:n]

    def homogenized_stress(self):
        out = np.zeros(self.nstr)
        self.L.or_homogenized_stress(self.ptr, _dp(out))
        return out

    def homogenized_tangent(self, pert=1e-6):
        out = np.zeros((self.nstr, self.nstr))
        self.L.or_homogenized_tangent(self.ptr, float(pert), _dp(out))
        return out

    def extrapolate_displacement(self):
        self.L.or_extrapolate_displacement(self.ptr)

    def get_u(self):
        out = np.zeros(self.dims + (self.h,))
        self.L.or_get_u(self.ptr, _dp(out))
        return out

    def set_u(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        self.L.or_set_u(self.ptr, _dp(u))

    def get_r(self):
        out = np.zeros(self.dims + (self.h,))
        self.L.or_get_r(self.ptr, _dp(out))
        return out

    def get_state(self, committed=True):
        ep = np.zeros((self.N, self.ngp, 6))
        q = np.zeros((self.N, self.ngp))
        self.L.or_get_state(self.ptr, _dp(ep), _dp(q), int(committed))
        return ep, q

    def get_kinematic_state(self, committed=True):
        """J2ViscoPlastic psi_bar per element and Gauss point, [N][n_gp][6]."""
        pb = np.zeros((self.N, self.ngp, 6))
        self.L.or_get_kinematic_state(self.ptr, _dp(pb), int(committed))
        return pb

    def element_fields(self, n_slots=None):
        ns, ng = self.nstr, self.ngp
        n_slots = n_slots or (int(np.max(self.p.ms)) + 1)
        out = dict(
            strain=np.zeros((self.N, ns)), stress=np.zeros((self.N, ns)),
            strain_gp=np.zeros((self.N, ng, ns)), stress_gp=np.zeros((self.N, ng, ns)),
            stress_average=np.zeros(ns), strain_average=np.zeros(ns),
            phase_stress_average=np.zeros((n_slots, ns)), phase_strain_average=np.zeros((n_slots, ns)),
            phase_counts=np.zeros(n_slots, dtype=np.int32))
        self.L.or_element_fields(self.ptr, _dp(out["strain"]), _dp(out["stress"]), _dp(out["strain_gp"]),
                                 _dp(out["stress_gp"]), _dp(out["stress_average"]), _dp(out["strain_average"]),
                                 _dp(out["phase_stress_average"]), _dp(out["phase_strain_average"]),
                                 _ip(out["phase_counts"]), n_slots)
        return out

    def plastic_flag(self):
        """PseudoPlastic::postprocess plastic_flag: element means of the Gauss-point flags."""
        out = np.zeros(self.N)
        self.L.or_plastic_flag(self.ptr, _dp(out))
        return out

    # --- single operators ---------------------------------------------------------
    def residual(self, u, linear_kd=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        r = np.zeros(self.dims + (self.h,))
        self.L.or_residual(self.ptr, _dp(u), _dp(r), int(linear_kd))
        return r

    def convolve(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(self.dims + (self.h,))
        self.L.or_convolve(self.ptr, _dp(r), _dp(out))
        return out

    def green(self):
        nx, ny, nz = self.dims
        np_ = self.h * (self.h + 1) // 2
        out = np.zeros((nx, ny, nz // 2 + 1, np_))
        self.L.or_get_green(self.ptr, _dp(out))
        return out

    def reference_stiffness(self):
        out = np.zeros((self.nstr, self.nstr))
        self.L.or_get_reference_stiffness(self.ptr, _dp(out))
        return out

    def ref_element_stiffness(self):
        ne = 8 * self.h
        out = np.zeros((ne, ne))
        self.L.or_get_ref_element_stiffness(self.ptr, _dp(out))
        return out

    def B(self):
        out = np.zeros((self.ngp, self.nstr, 8 * self.h))
        self.L.or_get_B(self.ptr, _dp(out))
        return out

    def phase_stiffness(self, phase):
        ne = 8 * self.h
        out = np.zeros((ne, ne))
        if self.L.or_get_phase_stiffness(self.ptr, int(phase), _dp(out)) != 0:
            raise ValueError("bad phase")
        return out
