"""Differentiation of homogenized responses (per-voxel moduli, adjoint) and the
pyfans Micro Manager API.

CPU: the sensitivity formula itself is pinned with the oracle -- every voxel its
own LinearElasticIsotropic / LinearThermalIsotropic phase, central finite
differences of v . Sigma_bar in one phase's modulus against the Gauss-point
integral eps(v)^T dC/dp eps(g0) / V built from the oracle's own strains.
GPU: per-voxel moduli against the oracle, the library's sensitivity kernel
against the same formula, the torch.autograd gradient of the SIMP objective
(homogenized C_1111 + C_2323 from two load cases) against central finite
differences, and the mixed-precision gradient against the FP64 one.
"""
import numpy as np
import pytest

import oracle
from fans_amd.problem import MaterialGroup, Problem
from fans_amd.workloads import config4_objective, smoothed_density

R2 = np.sqrt(0.5)
FD96_TOL = 2e-5  # central differences at 96^3, step 1e-6, relative to max |dJ/drho|


def per_voxel_problem(K, G=None, fe="HEX8", tol=1e-12, L=(1.0, 1.0, 1.0)):
    dims = K.shape
    N = K.size
    ms = np.arange(N, dtype=np.uint16).reshape(dims)
    if G is None:
        grp = MaterialGroup("LinearThermalIsotropic", list(range(N)), {"conductivity": list(K.reshape(-1))})
        h = 1
    else:
        grp = MaterialGroup("LinearElasticIsotropic", list(range(N)),
                            {"bulk_modulus": list(K.reshape(-1)), "shear_modulus": list(G.reshape(-1))})
        h = 3
    return Problem(dims=dims, howmany=h, ms=ms, groups=[grp], L=L, fe_type=fe, method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=3000)


def oracle_solution(p, g0):
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    o.solve()
    return o


def formula_sensitivity(p, fa, fb):
    """Gauss integrals of eps_a^T dC/dp eps_b per element from GP strains."""
    ea, eb = fa["strain_gp"], fb["strain_gp"]  # [N, ngp, nstr]
    ngp = ea.shape[1]
    hx, hy, hz = (p.L[i] / p.dims[i] for i in range(3))
    w = hx * hy * hz / ngp
    dot = np.einsum("egc,egc->e", ea, eb)
    if p.howmany == 1:
        return (w * dot)[:, None]
    tra, trb = ea[..., :3].sum(-1), eb[..., :3].sum(-1)
    sK = w * np.einsum("eg,eg->e", tra, trb)
    sG = w * (2.0 * dot - (2.0 / 3.0) * np.einsum("eg,eg->e", tra, trb))
    return np.stack([sK, sG], axis=1)


@pytest.mark.parametrize("h,fe", [(1, "HEX8"), (3, "HEX8"), (3, "BBAR")])
def test_sensitivity_formula_matches_fd(h, fe):
    rng = np.random.default_rng(7)
    dims = (4, 4, 6)
    K = rng.uniform(1.0, 5.0, dims)
    G = rng.uniform(0.5, 3.0, dims) if h == 3 else None
    g0 = np.array([1.0, 0.2, -0.1]) if h == 1 else np.array([1.0, 0.1, -0.2, 0.0, 0.3, 0.0])
    v = np.array([0.3, -1.0, 0.5]) if h == 1 else np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0])
    p = per_voxel_problem(K, G, fe=fe)
    fa = oracle_solution(p, g0).element_fields()
    fb = oracle_solution(p, v).element_fields()
    sens = formula_sensitivity(p, fa, fb) / np.prod(p.L)
    for e in (0, 17, 50, 95):
        for q in range(sens.shape[1]):
            hstep = 1e-4
            vals = []
            for sgn in (+1, -1):
                K2 = K.copy()
                G2 = None if G is None else G.copy()
                if q == 0:
                    K2.reshape(-1)[e] += sgn * hstep
                else:
                    G2.reshape(-1)[e] += sgn * hstep
                vals.append(v @ oracle_solution(per_voxel_problem(K2, G2, fe=fe), g0).homogenized_stress())
            fd = (vals[0] - vals[1]) / (2 * hstep)
            assert abs(fd - sens[e, q]) <= 1e-6 * np.max(np.abs(sens[:, q])), (e, q, fd, sens[e, q])
    # and a plausible mistake (dC/dG without the -2/3 m m^T term) is caught
    if h == 3:
        wrong = sens[:, 1] + (2.0 / 3.0) * sens[:, 0]
        assert np.max(np.abs(wrong - sens[:, 1])) > 1e-3 * np.max(np.abs(sens[:, 1]))


# ---------------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("h,fe", [(1, "HEX8"), (3, "HEX8"), (3, "BBAR")])
def test_voxel_moduli_forward_and_sensitivity(h, fe):
    from fans_amd import gpu
    from fans_amd.pyfans import voxel_problem
    print(f"[gpu-case] voxel moduli h={h} {fe}", flush=True)
    rng = np.random.default_rng(3)
    dims = (8, 8, 8)
    K = rng.uniform(1.0, 5.0, dims)
    G = rng.uniform(0.5, 3.0, dims) if h == 3 else None
    g0 = np.array([1.0, 0.2, -0.1]) if h == 1 else np.array([1.0, 0.1, -0.2, 0.0, 0.3, 0.0])
    v = np.array([0.3, -1.0, 0.5]) if h == 1 else np.array([0.0, 0.0, 0.0, 0.2, 0.0, 1.0])
    p = per_voxel_problem(K, G, fe=fe, tol=1e-10)
    o = oracle_solution(p, g0)
    ob = oracle_solution(p, v)
    s = gpu.GpuSolver(voxel_problem(dims, howmany=h, fe_type=fe, tol=1e-10))
    s.set_voxel_moduli(K, G)
    assert np.max(np.abs(s.reference_stiffness() - o.reference_stiffness())) < 1e-14
    s.set_gradient(v)
    s.solve()
    ub = s.get_u().copy()
    s.reset_state()
    s.set_gradient(g0)
    it = s.solve()
    so = o.homogenized_stress()
    assert np.max(np.abs(s.homogenized_stress() - so)) <= 1e-10 * np.max(np.abs(so))
    sens = s.sensitivity(ub, v)
    ref = formula_sensitivity(p, o.element_fields(), ob.element_fields())
    assert np.max(np.abs(sens - ref)) <= 1e-8 * np.max(np.abs(ref))
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n,step,tol", [(16, 1e-6, 2e-4), (96, 1e-6, 2e-5)])
def test_autograd_simp_gradient_fd(n, step, tol):
    """Config-4 objective on the host path (FP64 density in host memory):
    autograd (adjoint) gradient vs central differences with step 1e-6 on 32
    seeded voxels and on the two phase (E, nu) pairs.  At 96^3 a step of 1e-6
    moves J by ~1e-12; the differences still agree with the adjoint to ~1e-6 of
    max |dJ/drho|."""
    import torch
    from fans_amd import gpu
    from fans_amd.pyfans import voxel_problem
    print(f"[gpu-case] autograd simp n={n}", flush=True)
    rho0 = smoothed_density(n, corr=8.0 * n / 96.0, seed=0)
    params = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (1.0, 0.3, 1e-3, 0.3)]
    s = gpu.GpuSolver(voxel_problem((n, n, n), tol=1e-12 if n <= 32 else 1e-11))
    rho = torch.tensor(rho0, dtype=torch.float64, requires_grad=True)
    J = config4_objective(s, rho, params)
    J.backward()
    g = rho.grad.numpy().reshape(-1)
    gp = [p.grad.item() for p in params]
    pick = np.random.default_rng(11).choice(n ** 3, 32, replace=False)
    with torch.no_grad():
        for e in pick:
            vals = []
            for sgn in (+1, -1):
                r = rho0.copy().reshape(-1)
                r[e] += sgn * step
                vals.append(config4_objective(s, torch.tensor(r.reshape(rho0.shape)), params).item())
            fd = (vals[0] - vals[1]) / (2 * step)
            assert abs(fd - g[e]) <= tol * np.max(np.abs(g)), (e, fd, g[e])
        for k in range(4):
            hs = 1e-6
            vals = []
            for sgn in (+1, -1):
                q = [p.detach().clone() for p in params]
                q[k] += sgn * hs
                vals.append(config4_objective(s, torch.tensor(rho0), q).item())
            fd = (vals[0] - vals[1]) / (2 * hs)
            assert abs(fd - gp[k]) <= 1e-5 * max(abs(gp[k]), 1e-3), (k, fd, gp[k])
    s.close()


@pytest.mark.gpu
def test_autograd_device_resident_fp32_density():
    """Config 4 as stated, at 96^3: the density stored in FP32 in GPU memory,
    moduli and gradients FP64 on the device (no per-voxel array crosses to the
    host).  The gradient equals the host path's for the same density in FP64
    and central differences with step 1e-6 on 32 seeded voxels (FP64 density,
    solver converged to L-inf 1e-12)."""
    import torch
    from fans_amd import gpu
    from fans_amd.pyfans import voxel_problem
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    n = 96
    print(f"[gpu-case] autograd device-resident n={n}", flush=True)
    dev = torch.device("cuda")
    rho0 = smoothed_density(n, corr=8.0, seed=0)
    params = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (1.0, 0.3, 1e-3, 0.3)]
    s = gpu.GpuSolver(voxel_problem((n, n, n), tol=1e-12))
    rho = torch.tensor(rho0, dtype=torch.float32, device=dev, requires_grad=True)
    J = config4_objective(s, rho, params)
    J.backward()
    assert rho.grad.is_cuda and rho.grad.dtype == torch.float32
    g = rho.grad.double().cpu().numpy().reshape(-1)
    gmax = np.max(np.abs(g))
    gp = [p.grad.item() for p in params]
    # the host path on the same (FP32-valued) density held in FP64
    base = rho.detach().double()
    rho_h = torch.tensor(base.cpu().numpy(), requires_grad=True)
    params_h = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (1.0, 0.3, 1e-3, 0.3)]
    Jh = config4_objective(s, rho_h, params_h)
    Jh.backward()
    assert abs(J.item() - Jh.item()) <= 1e-12 * abs(Jh.item())
    assert np.max(np.abs(g - rho_h.grad.numpy().reshape(-1))) <= 1e-6 * gmax
    for k in range(4):
        assert abs(gp[k] - params_h[k].grad.item()) <= 1e-10 * max(abs(gp[k]), 1e-3)
    pick = np.random.default_rng(11).choice(n ** 3, 32, replace=False)
    step = 1e-6
    with torch.no_grad():
        for e in pick:
            vals = []
            for sgn in (+1, -1):
                r = base.clone().reshape(-1)
                r[e] += sgn * step
                vals.append(config4_objective(s, r.reshape(base.shape), params).item())
            fd = (vals[0] - vals[1]) / (2 * step)
            assert abs(fd - g[e]) <= FD96_TOL * gmax, (e, fd, g[e])
    s.close()


@pytest.mark.gpu
def test_autograd_mixed_matches_fp64():
    """The mixed FP32/FP64 form gives the FP64 adjoint gradient to ~1e-6."""
    import torch
    from fans_amd import gpu
    from fans_amd.pyfans import voxel_problem
    print("[gpu-case] autograd mixed vs fp64 32^3", flush=True)
    rho0 = smoothed_density(32, corr=8.0 / 3.0, seed=1)
    params = [torch.tensor(v, dtype=torch.float64) for v in (1.0, 0.3, 1e-3, 0.3)]
    grads = []
    for prec in ("fp64", "mixed"):
        s = gpu.GpuSolver(voxel_problem((32, 32, 32), tol=1e-10), precision=prec)
        rho = torch.tensor(rho0, requires_grad=True)
        config4_objective(s, rho, params).backward()
        grads.append(rho.grad.numpy().copy())
        s.close()
    assert np.max(np.abs(grads[1] - grads[0])) <= 1e-6 * np.max(np.abs(grads[0]))


def _micro_input(tmp_path, ms):
    """test_LinearElastic.json of the reference on sphere32, written as
    input.json next to its microstructure file."""
    import json
    import os

    from fans_amd import hdf5io
    from tests.conftest import GOLDEN
    dset = "/sphere/32x32x32/ms"
    hdf5io.write_microstructure(str(tmp_path / "sphere32.h5"), dset, ms)
    j = json.load(open(os.path.join(GOLDEN, "test_LinearElastic.json")))
    j["microstructure"]["filepath"] = str(tmp_path / "sphere32.h5")
    j["microstructure"]["datasetname"] = dset
    json.dump(j, open(tmp_path / "input.json", "w"))
    return j


def test_micro_simulation_constructor_shape():
    """MicroSimulation(sim_id) with the reference's default input file name."""
    import inspect

    from fans_amd.pyfans import MicroSimulation
    sig = inspect.signature(MicroSimulation.__init__)
    assert list(sig.parameters)[:3] == ["self", "sim_id", "input_file"]
    assert sig.parameters["input_file"].default == "input.json"


@pytest.mark.gpu
def test_micro_simulation_default_input(tmp_path, monkeypatch, sphere32_ms):
    """MicroSimulation(0) reads ./input.json; solve() returns the homogenized
    stress and the upper triangle of the tangent in the reference's cmat1..7
    layout (micro.cpp:66-79), against the oracle."""
    from fans_amd.pyfans import MicroSimulation
    _micro_input(tmp_path, sphere32_ms)
    monkeypatch.chdir(tmp_path)
    print("[gpu-case] MicroSimulation(0) on ./input.json (sphere32 elastic)", flush=True)
    sim = MicroSimulation(0)
    g0 = np.array([1e-3, -2e-4, 5e-4, 1e-4, 0.0, -3e-4])
    out = sim.solve({"strains1to3": g0[:3], "strains4to6": g0[3:]}, 0.0)
    o = oracle.Oracle(sim.problem)
    o.set_gradient(g0)
    o.solve()
    so = o.homogenized_stress()
    Co = o.homogenized_tangent(1e-6)
    s = np.array(out["stresses1to3"] + out["stresses4to6"])
    assert np.max(np.abs(s - so)) <= 1e-10 * np.max(np.abs(so))
    cm = np.concatenate([out[f"cmat{k}"] for k in range(1, 8)])
    ref = np.array([Co[0, 0], Co[0, 1], Co[0, 2], Co[0, 3], Co[0, 4], Co[0, 5], Co[1, 1], Co[1, 2], Co[1, 3],
                    Co[1, 4], Co[1, 5], Co[2, 2], Co[2, 3], Co[2, 4], Co[2, 5], Co[3, 3], Co[3, 4], Co[3, 5],
                    Co[4, 4], Co[4, 5], Co[5, 5]])
    assert np.max(np.abs(cm - ref)) <= 1e-6 * np.max(np.abs(ref))
    sim.solver.close()
