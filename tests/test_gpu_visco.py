"""J2ViscoPlastic_{Linear,NonLinear}IsotropicHardening (J2Plasticity.h) on the
GPU against the oracle.  Their current psi / psi_bar gain gamma at every
get_sigma call (J2Plasticity.h:103-104), so these runs also check that the GPU
evaluates the material exactly as often as the reference's call sequence does:
initial residual, line-search residuals, mixed-BC homogenized stresses and the
postprocess element pass.  Tolerances as for J2PlasticityNew: homogenized
stress 1e-8, CG iterations +-2, committed state 1e-6 of its largest entry."""
import numpy as np
import pytest

import oracle
from fans_amd import gpu
from fans_amd.problem import MaterialGroup, Problem
from tests import cases

pytestmark = pytest.mark.gpu


def relerr(a, b, floor=1e-300):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor))


def visco_problem(ms, nonlinear, Kh=0.0, Hk=0.0, fe="HEX8"):
    props = {"bulk_modulus": [62.5], "shear_modulus": [28.8462], "yield_stress": [0.1],
             "isotropic_hardening_parameter": [Kh], "kinematic_hardening_parameter": [Hk], "viscosity": [1.0],
             "time_step": 0.01}
    name = "J2ViscoPlastic_LinearIsotropicHardening"
    if nonlinear:
        props.update(saturation_stress=[0.15], saturation_exponent=[1000.0])
        name = "J2ViscoPlastic_NonLinearIsotropicHardening"
    groups = [MaterialGroup(name, [0], props),
              MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [222.222], "shear_modulus": [166.6667]})]
    return Problem(dims=ms.shape, howmany=3, ms=ms, groups=groups, fe_type=fe, method="cg", measure="Linfinity",
                   tol=1e-10, n_it=100)


@pytest.mark.parametrize("nonlinear", [False, True])
def test_visco_residual(sphere32_ms, nonlinear):
    p = visco_problem(sphere32_ms, nonlinear, Kh=0.5, Hk=0.3)
    print(f"[gpu-case] visco residual nonlinear={nonlinear}", flush=True)
    rng = np.random.default_rng(4)
    u = rng.standard_normal(p.dims + (3,)) * 0.002
    g0 = np.array([0.006, -0.002, -0.002, 0.002, 0.0, -0.001])
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    g = gpu.GpuSolver(p)
    g.set_gradient(g0)
    assert relerr(g.residual(u), o.residual(u)) < 1e-12
    g.close()


@pytest.mark.parametrize("nonlinear,Kh,Hk", [(False, 0.5, 0.3), (True, 0.0, 0.0), (True, 0.4, 0.2)])
def test_visco_strain_steps(sphere32_ms, nonlinear, Kh, Hk):
    """Strain-driven steps past yield with the CLI's per-step sequence (set
    gradient, solve, postprocess element pass, extrapolate)."""
    p = visco_problem(sphere32_ms, nonlinear, Kh, Hk)
    print(f"[gpu-case] visco strain steps nonlinear={nonlinear} Kh={Kh} Hk={Hk}", flush=True)
    o = oracle.Oracle(p)
    g = gpu.GpuSolver(p)
    for t, e11 in enumerate((0.001, 0.002, 0.003, 0.004)):
        g0 = np.array([e11, 0, 0, 0, 0, 0])
        o.set_gradient(g0)
        it_o = o.solve()
        g.set_gradient(g0)
        it_g = g.solve()
        assert abs(it_g - it_o) <= 2, (t, it_g, it_o)
        fo = o.element_fields()
        fg = g.element_fields(gp=False)
        assert relerr(fg["stress_average"], fo["stress_average"]) < 1e-8, t
        assert relerr(fg["stress"], fo["stress"]) < 1e-6, t
        ep_o, q_o = o.get_state()
        ep_g, q_g = g.get_state()
        assert relerr(ep_g, ep_o, 1e-12) < 1e-6, t
        assert relerr(q_g, q_o, 1e-12) < 1e-6, t
        assert relerr(g.get_field("kinematic_hardening"), o.get_kinematic_state(), 1e-12) < 1e-6, t
        o.extrapolate_displacement()
        g.extrapolate_displacement()
    assert q_o.max() > 0  # the matrix yielded
    g.close()


def test_visco_mixed_bc_steps(sphere32_ms):
    """Uniaxial stress (mixed BC): every line-search residual is preceded by a
    homogenized-stress pass, each one more get_sigma call per point."""
    p = visco_problem(sphere32_ms, True, Kh=0.3, Hk=0.2)
    p.tol = 1e-9
    mbc = cases.uniaxial_mixed(3, 0.006)
    print("[gpu-case] visco mixed BC 3 steps", flush=True)
    o = oracle.Oracle(p)
    g = gpu.GpuSolver(p)
    for t in range(3):
        o.enable_mixed_bc(mbc, t)
        it_o = o.solve()
        g.enable_mixed_bc(mbc, t)
        it_g = g.solve()
        assert abs(it_g - it_o) <= 2, (t, it_g, it_o)
        assert relerr(g.homogenized_stress(), o.homogenized_stress()) < 1e-8, t
        assert relerr(g.macro_gradient(), o.macro_gradient()) < 1e-8, t
        ep_o, q_o = o.get_state()
        ep_g, q_g = g.get_state()
        assert relerr(q_g, q_o, 1e-12) < 1e-6, t
        assert relerr(g.get_field("kinematic_hardening"), o.get_kinematic_state(), 1e-12) < 1e-6, t
        o.extrapolate_displacement()
        g.extrapolate_displacement()
    g.close()
