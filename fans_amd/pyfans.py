"""pyfans on the GPU path: the Micro Manager API of the reference's pybind11
module (pyfans/micro.cpp:28-93) and a differentiable homogenization layer.

MicroSimulation(sim_id, input_file).solve(macro_data, dt) keeps the reference's
dictionary keys (strains1to3 / strains4to6 in, stresses1to3 / stresses4to6 and
the upper triangle of the tangent as cmat1..cmat7 out).

HomogenizedStress is a torch.autograd.Function: Sigma_bar(K, G, g0) of a linear
isotropic RVE with per-voxel bulk and shear moduli (thermal: conductivity).
Its backward is the hand-written adjoint: for an upstream gradient v of
Sigma_bar, the adjoint state is the solve for macro gradient v (the operator is
self-adjoint), and

    d(v . Sigma_bar)/dK_e = (1/V) int_e eps(v)^T m m^T eps(g0)
    d(v . Sigma_bar)/dG_e = (1/V) int_e eps(v)^T (2 I - 2/3 m m^T) eps(g0)
    d(v . Sigma_bar)/dg0  = Sigma_bar(v)

(m = (1,1,1,0,0,0) in Mandel notation), evaluated per element by the
sensitivity kernel of the library.  The element stiffness is linear in the
moduli and the fluctuation solves the discrete equilibrium, so these are the
exact derivatives of the discrete homogenized stress (up to the solver
tolerance).  No reference counterpart: the reference's per-phase parameters
with uint16 phase ids cannot express per-voxel moduli.

With K and G in GPU memory (torch imported before the library, so that both
share one HIP runtime) the moduli, the forward solution kept for the backward
pass and the per-voxel gradients stay on the device: only the n_str-vectors
g0, Sigma_bar and the upstream gradient cross to the host.  A density may be
stored in FP32; simp_moduli evaluates the moduli in FP64.
"""
from __future__ import annotations

import numpy as np
import torch

from .gpu import GpuSolver
from .problem import MaterialGroup, Problem


class MicroSimulation:
    """pyfans/micro.cpp:28-93 on the GPU path (mechanical, small strain).

    MicroSimulation(sim_id) reads "input.json" from the working directory like
    the reference's constructor (micro.hpp:22, default argument; the pybind11
    module binds the int-only form, micro.cpp:93).  input_file, precision and
    pert_param are keyword extensions; pert_param is the reference's scalar
    strain perturbation of the tangent (micro.hpp:31)."""

    def __init__(self, sim_id: int, input_file="input.json", precision="fp64", pert_param=1e-6):
        self.sim_id = sim_id
        self.problem = Problem.from_json(input_file)
        if self.problem.howmany != 3:
            raise ValueError("MicroSimulation is mechanical (howmany = 3) like pyfans")
        self.solver = GpuSolver(self.problem, precision=precision)
        self.pert_param = pert_param

    def solve(self, macro_data, dt=0.0):
        g0 = np.concatenate([np.asarray(macro_data["strains1to3"], dtype=np.float64),
                             np.asarray(macro_data["strains4to6"], dtype=np.float64)])
        self.solver.set_gradient(g0)
        self.solver.solve()
        s = self.solver.homogenized_stress()
        C = self.solver.homogenized_tangent(self.pert_param)
        iu = np.triu_indices(6)
        tri = C[iu]  # (0,0) (0,1) ... (5,5): 21 values, 3 per cmat entry like micro.cpp:66-79
        out = {"stresses1to3": list(s[:3]), "stresses4to6": list(s[3:])}
        for k in range(7):
            out[f"cmat{k + 1}"] = list(tri[3 * k:3 * k + 3])
        return out


def voxel_problem(dims, howmany=3, L=(1.0, 1.0, 1.0), fe_type="HEX8", tol=1e-10, n_it=2000, measure="Linfinity"):
    """Problem skeleton for per-voxel moduli: one isotropic phase whose table
    values are replaced by GpuSolver.set_voxel_moduli."""
    ms = np.zeros(dims, dtype=np.uint16)
    if howmany == 3:
        grp = MaterialGroup("LinearElasticIsotropic", [0], {"bulk_modulus": [1.0], "shear_modulus": [1.0]})
    else:
        grp = MaterialGroup("LinearThermalIsotropic", [0], {"conductivity": [1.0]})
    return Problem(dims=tuple(dims), howmany=howmany, ms=ms, groups=[grp], L=tuple(L), fe_type=fe_type,
                   method="cg", measure=measure, err_type="absolute", tol=tol, n_it=n_it)


class HomogenizedStress(torch.autograd.Function):
    """Sigma_bar(K, G, g0) with the adjoint backward described above."""

    @staticmethod
    def forward(ctx, K, G, g0, solver):
        mech = solver.p.howmany != 1
        ctx.on_device = K.is_cuda
        if ctx.on_device:
            k = K.detach().to(torch.float64).contiguous()
            g = G.detach().to(torch.float64).contiguous() if mech else None
        else:
            k = K.detach().to("cpu", torch.float64).numpy()
            g = G.detach().to("cpu", torch.float64).numpy() if mech else None
        gv = g0.detach().to("cpu", torch.float64).numpy()
        solver.set_voxel_moduli(k, g)
        solver.reset_state()
        solver.set_gradient(gv)
        solver.solve()
        sig = solver.homogenized_stress()
        ctx.solver = solver
        if ctx.on_device:
            ctx.u = solver.get_u(out=torch.empty(solver.N * solver.h, dtype=torch.float64, device=K.device))
        else:
            ctx.u = solver.get_u().copy()
        ctx.g0 = gv.copy()
        ctx.k, ctx.g = k, g
        ctx.dims = K.shape
        return torch.as_tensor(sig, dtype=torch.float64, device=g0.device)

    @staticmethod
    def backward(ctx, grad_sigma):
        s = ctx.solver
        v = grad_sigma.detach().to("cpu", torch.float64).numpy()
        # the solver may have been used for other moduli since the forward pass
        s.set_voxel_moduli(ctx.k, ctx.g)
        s.reset_state()
        s.set_gradient(v)
        s.solve()                     # adjoint state: the solve for macro gradient v
        dg0 = s.homogenized_stress()  # Sigma_bar(v) = C_bar v = C_bar^T v
        sens = s.sensitivity(ctx.u, ctx.g0)
        V = float(np.prod(s.p.L))
        if ctx.on_device:
            dK = (sens[:, 0] / V).reshape(ctx.dims)
            dG = (sens[:, 1] / V).reshape(ctx.dims) if sens.shape[1] > 1 else None
        else:
            dev = grad_sigma.device
            dK = torch.as_tensor(sens[:, 0] / V, dtype=torch.float64, device=dev).reshape(ctx.dims)
            dG = None
            if sens.shape[1] > 1:
                dG = torch.as_tensor(sens[:, 1] / V, dtype=torch.float64, device=dev).reshape(ctx.dims)
        return dK, dG, torch.as_tensor(dg0, dtype=torch.float64, device=grad_sigma.device), None


def homogenized_stress(solver: GpuSolver, K, G, g0):
    """Sigma_bar for per-voxel moduli K, G (torch tensors [nx, ny, nz]; G is
    ignored for thermal problems) under macro gradient g0 (torch [n_str]);
    differentiable in K, G and g0 through the adjoint solve."""
    return HomogenizedStress.apply(K, G, g0, solver)


def simp_moduli(rho, E_solid, nu_solid, E_void, nu_void, p=3.0):
    """SIMP interpolation E(rho) = E_void + rho^p (E_solid - E_void), likewise
    nu, then bulk and shear moduli; all torch, differentiable.  A density stored
    in FP32 is widened first: the moduli and their gradients are FP64."""
    if rho.dtype != torch.float64:
        rho = rho.to(torch.float64)
    wgt = rho ** p
    E = E_void + wgt * (E_solid - E_void)
    nu = nu_void + wgt * (nu_solid - nu_void)
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    return K, G
