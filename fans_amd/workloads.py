"""Synthetic inputs of the benchmark configurations (seeded, deterministic).

Every microstructure comes from `random_spheres` (numpy PCG64 stream, periodic
non-overlapping placement, count + radius law honoured; the achieved volume
fraction is whatever that law gives and is reported by `inclusion_fraction`).
Loads are Mandel unit vectors (the reference's stress/strain ordering,
matmodel.h:284-304), the order get_homogenized_tangent uses
(solver.h:759-763).

    config1_problem()   thermal 32^3, 12 spheres r~U[3,6], k = 1 / 10, three
                        load cases e_x, e_y, e_z, CG L-inf 1e-10, n_it 500
    config2_problem()   elastic 128^3, 40 spheres r~U[6,14], matrix E=3
                        nu=0.35, inclusions E=70 nu=0.22, six Mandel unit
                        strains, CG L-inf 1e-10
    config3_problem()   J2 matrix (E=70 nu=0.33 yield 0.25 H=1) + elastic
                        spheres (E=400 nu=0.2), 150 spheres r~U[8,20] at 256^3,
                        with `config3_ramp()` (uniaxial strain to 2 % in 20
                        steps, zero lateral stress), CG L-inf 1e-8 per step
    headline_problem()  256^3 elastic: config 3's sphere law with config 2's
                        materials, Mandel e_11, CG L-inf 1e-10
    smoothed_density() / config4_objective()
                        config 4: SIMP density on a smoothed Gaussian
                        field, objective C_1111 + C_2323 (pyfans autograd)
(config 5 lives in batch.config5_rves.)
"""
from __future__ import annotations

import numpy as np

from .microstructure import random_spheres
from .problem import LoadCase, MaterialGroup, Problem, isotropic_KG


def mandel_unit(n_str, k):
    e = np.zeros(n_str)
    e[k] = 1.0
    return e


def inclusion_fraction(ms):
    return float(np.mean(np.asarray(ms) != 0))


def _scaled_spheres(n, count_at_256, rmin_256, rmax_256, seed):
    s = n / 256.0
    return random_spheres(n, max(1, int(round(count_at_256 * s ** 3))), rmin_256 * s, rmax_256 * s, seed=seed)


def config1_problem(seed=0, n=32):
    ms = random_spheres(n, 12, 3.0, 6.0, seed=seed)
    groups = [MaterialGroup("LinearThermalIsotropic", [0, 1], {"conductivity": [1.0, 10.0]})]
    lcs = [LoadCase(mixed=False, g0_path=mandel_unit(3, k)[None, :], n_steps=1) for k in range(3)]
    return Problem(dims=(n, n, n), howmany=1, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=1e-10, n_it=500, load_cases=lcs,
                   problem_type="thermal")


def config2_problem(seed=0, n=128, tol=1e-10):
    s = n / 128.0
    ms = random_spheres(n, max(1, int(round(40 * s ** 3))), 6.0 * s, 14.0 * s, seed=seed)
    Km, Gm = isotropic_KG(3.0, 0.35)
    Ki, Gi = isotropic_KG(70.0, 0.22)
    groups = [MaterialGroup("LinearElasticIsotropic", [0, 1], {"bulk_modulus": [Km, Ki], "shear_modulus": [Gm, Gi]})]
    lcs = [LoadCase(mixed=False, g0_path=mandel_unit(6, k)[None, :], n_steps=1) for k in range(6)]
    return Problem(dims=(n, n, n), howmany=3, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=2000, load_cases=lcs)


def headline_problem(n=256, seed=0, tol=1e-10):
    ms = _scaled_spheres(n, 150, 8.0, 20.0, seed)
    Km, Gm = isotropic_KG(3.0, 0.35)
    Ki, Gi = isotropic_KG(70.0, 0.22)
    groups = [MaterialGroup("LinearElasticIsotropic", [0, 1], {"bulk_modulus": [Km, Ki], "shear_modulus": [Gm, Gi]})]
    lcs = [LoadCase(mixed=False, g0_path=mandel_unit(6, 0)[None, :], n_steps=1)]
    return Problem(dims=(n, n, n), howmany=3, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=2000, load_cases=lcs)


def config3_problem(n=256, seed=0, tol=1e-8):
    ms = _scaled_spheres(n, 150, 8.0, 20.0, seed)
    Km, Gm = isotropic_KG(70.0, 0.33)
    Ki, Gi = isotropic_KG(400.0, 0.2)
    groups = [MaterialGroup("J2PlasticityNew_LinearIsotropicHardening", [0],
                            {"bulk_modulus": [Km], "shear_modulus": [Gm], "yield_stress": [0.25],
                             "isotropic_hardening_parameter": [1.0]}),
              MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [Ki], "shear_modulus": [Gi]})]
    return Problem(dims=(n, n, n), howmany=3, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=2000)


def config3_ramp(n_steps=20, emax=0.02):
    """Strain-controlled e_11 ramp, zero stress on 22 and 33, zero shear strain
    (mixed BC in the reference's format, mixedBCs.h:51-133)."""
    strain = np.linspace(emax / n_steps, emax, n_steps)
    return dict(strain_indices=[0, 3, 4, 5], stress_indices=[1, 2],
                strain_path=np.stack([strain, 0 * strain, 0 * strain, 0 * strain], 1),
                stress_path=np.zeros((n_steps, 2)), n_steps=n_steps)


def smoothed_density(n, corr=8.0, seed=0):
    """Configuration 4's density: a seeded Gaussian random field smoothed with
    a periodic Gaussian filter of correlation length `corr` voxels, through a
    sigmoid, rho in (0, 1)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    filt = np.exp(-0.5 * (2 * np.pi * corr) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    f = np.real(np.fft.ifftn(np.fft.fftn(x) * filt))
    f /= f.std()
    return 1.0 / (1.0 + np.exp(-3.0 * f))


def config4_objective(solver, rho, params):
    """C_1111 + C_2323 of the homogenized tangent (Mandel entry (5,5) / 2) from
    two load cases, SIMP moduli of `rho` with params = (E_solid, nu_solid,
    E_void, nu_void); differentiable through pyfans.homogenized_stress."""
    import torch

    from .pyfans import homogenized_stress, simp_moduli
    K, G = simp_moduli(rho, *params)
    e1 = torch.tensor([1.0, 0, 0, 0, 0, 0], dtype=torch.float64)
    e6 = torch.tensor([0, 0, 0, 0, 0, 1.0], dtype=torch.float64)
    return homogenized_stress(solver, K, G, e1)[0] + 0.5 * homogenized_stress(solver, K, G, e6)[5]
