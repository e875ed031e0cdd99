"""16-bit phase ids (the reference reads the microstructure as unsigned short,
reader.h:53 / reader.cpp:260): grids with more than 256 phases run on the GPU
against the oracle.  The isotropic (modal) matvec and residual read the phase
table beyond the 256 phases they stage in LDS; the per-Gauss-point residual
reads the model kind likewise."""
import numpy as np
import pytest

import oracle
from fans_amd import gpu
from fans_amd.problem import MaterialGroup, Problem

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def many_phase_problem(model, nph=300, dims=(16, 12, 10), seed=4, tol=1e-10):
    rng = np.random.default_rng(seed)
    ms = rng.integers(0, nph, size=dims).astype(np.uint16)
    ms.reshape(-1)[:nph] = np.arange(nph)  # every phase occurs
    K = rng.uniform(2.0, 60.0, nph)
    G = rng.uniform(1.0, 30.0, nph)
    props = {"bulk_modulus": list(K), "shear_modulus": list(G)}
    if model == "PseudoPlasticLinearHardening":
        props.update(yield_stress=list(rng.uniform(0.05, 0.2, nph)), hardening_parameter=list(rng.uniform(0.1, 1, nph)))
    grp = MaterialGroup(model, list(range(nph)), props)
    return Problem(dims=dims, howmany=3, ms=ms, groups=[grp], fe_type="HEX8", method="cg", measure="Linfinity",
                   tol=tol, n_it=500)


@pytest.mark.parametrize("model", ["LinearElasticIsotropic", "PseudoPlasticLinearHardening"])
def test_more_than_256_phases_residual(model):
    p = many_phase_problem(model)
    print(f"[gpu-case] 300 phases {model} residual", flush=True)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(p.dims + (3,)) * 0.01
    g0 = np.array([0.004, -0.001, 0.002, 0.001, -0.002, 0.0005])
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    g = gpu.GpuSolver(p)
    g.set_gradient(g0)
    assert relerr(g.residual(u), o.residual(u)) < 1e-12
    if model == "LinearElasticIsotropic":
        assert relerr(g.residual(u, linear_kd=True), o.residual(u, linear_kd=True)) < 1e-12
    g.close()


def test_more_than_256_phases_solve():
    p = many_phase_problem("LinearElasticIsotropic")
    print("[gpu-case] 300 phases elastic CG, fp64 and mixed", flush=True)
    g0 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.5])
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    it_o = o.solve()
    so = o.homogenized_stress()
    for prec, tol in (("fp64", 1e-10), ("mixed", 1e-8)):
        g = gpu.GpuSolver(p, precision=prec)
        g.set_gradient(g0)
        it_g = g.solve()
        if prec == "fp64":
            assert abs(it_g - it_o) <= 2
        assert relerr(g.homogenized_stress(), so) < tol, prec
        f = g.element_fields(gp=False)
        fo = o.element_fields()
        assert relerr(f["phase_stress_average"], fo["phase_stress_average"]) < 1e-8
        g.close()


def test_phase_id_width_switches_back():
    """A context that saw 16-bit ids takes a byte-sized microstructure again."""
    p = many_phase_problem("LinearElasticIsotropic")
    print("[gpu-case] phase id width 16 -> 8 bit", flush=True)
    g = gpu.GpuSolver(p)
    from dataclasses import replace
    q = replace(p, ms=(p.ms % 200).astype(np.uint16))
    g.load(q)
    o = oracle.Oracle(q)
    g0 = np.array([0.0, 1.0, 0.0, 0.0, 0.0, 0.0])
    o.set_gradient(g0)
    g.set_gradient(g0)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(p.dims + (3,))
    assert relerr(g.residual(u), o.residual(u)) < 1e-12
    g.close()
