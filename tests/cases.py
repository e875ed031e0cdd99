"""Problem builders shared by the parity tests (seeded, small)."""
import json
import os

import numpy as np

from fans_amd.microstructure import random_spheres
from fans_amd.problem import MaterialGroup, Problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_input(name, ms):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        j = json.load(f)
    return Problem.from_json(j, ms=ms)


def thermal(n=16, seed=0, fe="HEX8", method="cg", tol=1e-10, n_it=300, measure="Linfinity", k=(1.0, 10.0),
            count=4, rmin=2.0, rmax=4.0, L=(1.0, 1.0, 1.0), dims=None):
    dims = dims or (n, n, n)
    ms = random_spheres(dims, count, rmin, rmax, seed=seed, n_inclusion_phases=len(k) - 1)
    groups = [MaterialGroup("LinearThermalIsotropic", list(range(len(k))), {"conductivity": list(k)})]
    return Problem(dims=dims, howmany=1, ms=ms, groups=groups, L=L, fe_type=fe, method=method, measure=measure,
                   tol=tol, n_it=n_it)


def elastic(n=16, seed=0, fe="HEX8", method="cg", tol=1e-10, n_it=300, measure="Linfinity",
            K=(3.3333333333333335, 41.666666666666664), G=(1.1111111111111112, 28.688524590163937),
            count=4, rmin=2.0, rmax=4.0, L=(1.0, 1.0, 1.0), dims=None):
    dims = dims or (n, n, n)
    ms = random_spheres(dims, count, rmin, rmax, seed=seed, n_inclusion_phases=len(K) - 1)
    groups = [MaterialGroup("LinearElasticIsotropic", list(range(len(K))),
                            {"bulk_modulus": list(K), "shear_modulus": list(G)})]
    return Problem(dims=dims, howmany=3, ms=ms, groups=groups, L=L, fe_type=fe, method=method, measure=measure,
                   tol=tol, n_it=n_it)


def j2(n=16, seed=0, fe="HEX8", tol=1e-8, n_it=300, count=4, rmin=2.0, rmax=4.0, dims=None):
    """Config-3 materials on a small grid: J2 matrix (E=70, nu=0.33) + elastic spheres (E=400, nu=0.2)."""
    dims = dims or (n, n, n)
    ms = random_spheres(dims, count, rmin, rmax, seed=seed)
    Km, Gm = 70.0 / (3 * (1 - 2 * 0.33)), 70.0 / (2 * 1.33)
    Ki, Gi = 400.0 / (3 * (1 - 0.4)), 400.0 / 2.4
    groups = [MaterialGroup("J2PlasticityNew_LinearIsotropicHardening", [0],
                            {"bulk_modulus": [Km], "shear_modulus": [Gm], "yield_stress": [0.25],
                             "isotropic_hardening_parameter": [1.0]}),
              MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [Ki], "shear_modulus": [Gi]})]
    return Problem(dims=dims, howmany=3, ms=ms, groups=groups, fe_type=fe, method="cg", tol=tol, n_it=n_it)


def uniaxial_mixed(n_steps, emax):
    """Strain-controlled e_11 ramp with zero stress on 22/33 and zero shear strain (config 3 control)."""
    steps = np.linspace(emax / n_steps, emax, n_steps)
    return dict(strain_indices=[0, 3, 4, 5], stress_indices=[1, 2],
                strain_path=np.stack([steps, 0 * steps, 0 * steps, 0 * steps], 1),
                stress_path=np.zeros((n_steps, 2)), n_steps=n_steps)


def _spd(n, rng, shift):
    a = rng.standard_normal((n, n))
    return a @ a.T + shift * np.eye(n)


def general(model, n=12, seed=0, fe="HEX8", tol=1e-10, n_it=300, dims=None, count=3, rmin=2.0, rmax=3.0):
    """Two phases of one of the per-Gauss-point models (triclinic thermal or
    elastic with random SPD tensors, PseudoPlastic or J2 matrix with an elastic
    inclusion)."""
    dims = dims or (n, n, n)
    rng = np.random.default_rng(100 + seed)
    ms = random_spheres(dims, count, rmin, rmax, seed=seed)
    h = 1 if model == "LinearThermalTriclinic" else 3
    if model == "LinearThermalTriclinic":
        Ks = [_spd(3, rng, 1.0), _spd(3, rng, 5.0)]
        keys = [(i, j) for i in range(3) for j in range(i, 3)]
        props = {f"K_{i + 1}{j + 1}": [float(K[i, j]) for K in Ks] for i, j in keys}
        groups = [MaterialGroup(model, [0, 1], props)]
    elif model == "LinearElasticTriclinic":
        Cs = [_spd(6, rng, 3.0), _spd(6, rng, 20.0)]
        keys = [(i, j) for i in range(6) for j in range(i, 6)]
        props = {f"C_{i + 1}{j + 1}": [float(C[i, j]) for C in Cs] for i, j in keys}
        groups = [MaterialGroup(model, [0, 1], props)]
    elif model == "PseudoPlasticNonLinearHardening":  # test_PseudoPlastic.json's materials
        groups = [MaterialGroup(model, [0], {"bulk_modulus": [62.5], "shear_modulus": [28.8462],
                                             "yield_stress": [0.1], "hardening_exponent": [0.2], "eps_0": [0.01]}),
                  MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [222.222],
                                                                "shear_modulus": [166.6667]})]
    elif model == "PseudoPlasticLinearHardening":
        groups = [MaterialGroup(model, [0], {"bulk_modulus": [62.5], "shear_modulus": [28.8462],
                                             "yield_stress": [0.1], "hardening_parameter": [0.5]}),
                  MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [222.222],
                                                                "shear_modulus": [166.6667]})]
    else:
        return j2(n=n, seed=seed, fe=fe, tol=tol, n_it=n_it, count=count, rmin=rmin, rmax=rmax)
    return Problem(dims=dims, howmany=h, ms=ms, groups=groups, fe_type=fe, method="cg", tol=tol, n_it=n_it)
