"""Pins of the CPU oracle against independent references (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a transposed index, a wrong sign or scale) fails it:
  * FFT            -- numpy.fft and committed golden vectors
  * B, Ke          -- numerically differentiated shape functions, 3x3x3 Gauss
  * residual       -- brute-force sparse assembly (tests/fe_numpy.py)
  * Green operator -- G(xi) K_hat(xi) = I from impulse responses; one-step solve
  * CG / FP        -- scipy direct solve; laminate closed forms
  * J2             -- uniaxial-strain closed form of linear-hardening plasticity
  * mixed BC + LS  -- golden macro strains stored in the reference's test_MixedBCs.json
"""
import json
import os

import numpy as np
import pytest

import oracle
from fans_amd.microstructure import laminate, random_spheres
from fans_amd.problem import MaterialGroup, Problem
from tests import cases, fe_numpy
from tests.conftest import GOLDEN


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


# ----------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 6, 8, 1), (5, 3, 10, 3), (12, 10, 6, 3), (15, 9, 20, 1), (7, 4, 4, 1)])
def test_fft_against_numpy(shape):
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape)
    X = oracle.fft_r2c(x)
    assert rel(X, np.fft.rfftn(x, axes=(0, 1, 2))) < 1e-13
    y = oracle.fft_c2r(X, shape[2])
    assert rel(y, x * np.prod(shape[:3])) < 1e-13


def test_fft_golden_vectors():
    g = np.load(os.path.join(GOLDEN, "fft_golden.npz"))
    for key in [k for k in g.files if k.startswith("x_")]:
        tag = key[2:]
        X = oracle.fft_r2c(g[key])
        assert rel(X, g["X_" + tag]) < 1e-13, tag


# ----------------------------------------------------------------------------- element
@pytest.mark.parametrize("h,fe", [(1, "HEX8"), (3, "HEX8"), (1, "HEX8R"), (3, "HEX8R")])
def test_B_matrices(h, fe):
    L = (1.0, 0.8, 1.3)
    p = cases.thermal(n=4, fe=fe, L=L) if h == 1 else cases.elastic(n=4, fe=fe, L=L)
    o = oracle.Oracle(p)
    B = o.B()
    le = np.array(L) / 4
    xp, xm = 0.5 + np.sqrt(3) / 6, 0.5 - np.sqrt(3) / 6
    pts = [(0.5, 0.5, 0.5)] if fe == "HEX8R" else \
        [(xm if i == 0 else xp, xm if j == 0 else xp, xm if k == 0 else xp) for k in (0, 1) for j in (0, 1) for i in (0, 1)]
    for gp, xi in enumerate(pts):
        assert np.allclose(B[gp], fe_numpy.bmat(xi, le, h), atol=1e-8)
        # partition of unity: constant fields produce no strain
        const = np.tile(np.arange(1, h + 1), 8)
        assert np.max(np.abs(B[gp] @ const)) < 1e-12
        # exact gradient of a linear field
        Gr = np.arange(1, 3 * h + 1).reshape(h, 3) * 0.37
        X = np.array([[(a & 1) * le[0], ((a >> 1) & 1) * le[1], ((a >> 2) & 1) * le[2]] for a in range(8)])
        ue = (X @ Gr.T).reshape(-1)
        eps = B[gp] @ ue
        if h == 1:
            expect = Gr[0]
        else:
            sym = 0.5 * (Gr + Gr.T)
            expect = np.array([sym[0, 0], sym[1, 1], sym[2, 2], np.sqrt(2) * sym[0, 1], np.sqrt(2) * sym[0, 2],
                               np.sqrt(2) * sym[1, 2]])
        assert rel(eps, expect) < 1e-12


@pytest.mark.parametrize("h,fe", [(1, "HEX8"), (3, "HEX8"), (3, "HEX8R")])
def test_element_stiffness(h, fe):
    L = (1.0, 0.9, 1.1)
    p = cases.thermal(n=4, fe=fe, L=L) if h == 1 else cases.elastic(n=4, fe=fe, L=L)
    o = oracle.Oracle(p)
    le = np.array(L) / 4
    params = p.groups[0].param_matrix()
    for ph in range(2):
        Ke = o.phase_stiffness(ph)
        assert np.allclose(Ke, Ke.T, atol=1e-12 * np.abs(Ke).max())
        ev = np.linalg.eigvalsh(Ke)
        assert ev.min() > -1e-10 * ev.max()
        trans = [np.tile(np.eye(h)[c], 8) for c in range(h)]
        for t in trans:
            assert np.max(np.abs(Ke @ t)) < 1e-10 * np.abs(Ke).max()
        if fe == "HEX8":  # 2x2x2 Gauss integrates the bilinear integrand exactly
            Kref, _ = fe_numpy.element_matrices(fe_numpy.iso_C(h, params[ph]), le, h, order=3)
            assert rel(Ke, Kref) < 1e-8


@pytest.mark.parametrize("maker,fe", [(cases.thermal, "HEX8"), (cases.elastic, "HEX8")])
def test_residual_against_sparse_assembly(maker, fe):
    p = maker(n=5, fe=fe, dims=(4, 5, 6), count=2, rmin=1.0, rmax=2.0)
    o = oracle.Oracle(p)
    h = p.howmany
    le = np.array(p.L) / np.array(p.dims)
    params = p.groups[0].param_matrix()
    Cs = {ph: fe_numpy.iso_C(h, params[ph]) for ph in range(len(params))}
    K, F = fe_numpy.assemble(p.ms, Cs, le, h)
    rng = np.random.default_rng(7)
    u = rng.standard_normal(p.dims + (h,))
    g0 = rng.standard_normal(p.n_str)
    o.set_gradient(g0)
    r = o.residual(u)
    r_ref = (K @ u.reshape(-1) + F @ g0).reshape(u.shape)
    assert rel(r, r_ref) < 1e-8
    assert rel(o.residual(u, linear_kd=True), (K @ u.reshape(-1)).reshape(u.shape)) < 1e-8


# ----------------------------------------------------------------------------- Green operator
@pytest.mark.parametrize("h", [1, 3])
def test_green_inverts_reference_operator(h):
    """Homogeneous medium equal to the reference material: G(xi) K_hat(xi) = I / N."""
    dims = (6, 4, 8)
    ms = np.zeros(dims, dtype=np.uint16)
    if h == 1:
        groups = [MaterialGroup("LinearThermalIsotropic", [0], {"conductivity": [2.5]})]
    else:
        groups = [MaterialGroup("LinearElasticIsotropic", [0], {"bulk_modulus": [5.0], "shear_modulus": [2.0]})]
    p = Problem(dims=dims, howmany=h, ms=ms, groups=groups, L=(1.0, 0.7, 1.2))
    o = oracle.Oracle(p)
    N = int(np.prod(dims))
    Khat = np.zeros((dims[0], dims[1], dims[2] // 2 + 1, h, h), dtype=complex)
    for c in range(h):
        e = np.zeros(dims + (h,))
        e[0, 0, 0, c] = 1.0
        Khat[..., :, c] = np.fft.rfftn(o.residual(e, linear_kd=True), axes=(0, 1, 2))
    G = o.green()
    iu = np.triu_indices(h)
    Gm = np.zeros(G.shape[:3] + (h, h))
    Gm[..., iu[0], iu[1]] = G
    Gm[..., iu[1], iu[0]] = G
    prod = np.einsum("...ij,...jk->...ik", Gm, Khat.real) * N
    eye = np.broadcast_to(np.eye(h), prod.shape).copy()
    eye[0, 0, 0] = 0.0
    assert np.max(np.abs(prod - eye)) < 1e-10
    assert np.max(np.abs(G[0, 0, 0])) == 0.0


@pytest.mark.parametrize("h", [1, 3])
def test_green_mirror_symmetry(h):
    """The GPU stores the Green operator for kx <= nx/2, ky <= ny/2 only when the
    reference element is mirror symmetric.  Pin the property it relies on with
    the oracle's directly computed operator (solver.h:144-204): for an
    isotropic reference, G(-kx) = Rx G(kx) Rx and G(-ky) = Ry G(ky) Ry, where
    Rx flips the xy and xz entries and Ry the xy and yz entries."""
    dims = (8, 6, 10)
    ms = (np.arange(np.prod(dims)).reshape(dims) % 3 == 0).astype(np.uint16)
    if h == 1:
        groups = [MaterialGroup("LinearThermalIsotropic", [0, 1], {"conductivity": [1.0, 7.0]})]
    else:
        groups = [MaterialGroup("LinearElasticIsotropic", [0, 1],
                                {"bulk_modulus": [3.0, 40.0], "shear_modulus": [1.0, 25.0]})]
    p = Problem(dims=dims, howmany=h, ms=ms, groups=groups, L=(1.0, 0.8, 1.3))
    G = oracle.Oracle(p).green()
    nx, ny = dims[0], dims[1]
    mx = np.roll(G[::-1], 1, axis=0)          # G(-kx): index (nx - kx) % nx
    my = np.roll(G[:, ::-1], 1, axis=1)
    if h == 1:
        sx = sy = np.ones(1)
    else:  # packed 00 01 02 11 12 22
        sx = np.array([1, -1, -1, 1, 1, 1.0])
        sy = np.array([1, -1, 1, 1, -1, 1.0])
    scale = np.max(np.abs(G))
    assert np.max(np.abs(mx - G * sx)) < 1e-13 * scale
    assert np.max(np.abs(my - G * sy)) < 1e-13 * scale
    # and a plausible mistake (unflipped xy entry) is caught
    if h == 3:
        assert np.max(np.abs(mx - G)) > 1e-3 * scale


def test_homogeneous_rve_one_step():
    """Homogeneous RVE: the Green operator solves the problem exactly; u stays 0 and
    the homogenized stress is C g0 (Q-free closed form)."""
    dims = (8, 8, 8)
    p = Problem(dims=dims, howmany=3, ms=np.zeros(dims, np.uint16),
                groups=[MaterialGroup("LinearElasticIsotropic", [0], {"bulk_modulus": [5.0], "shear_modulus": [2.0]})],
                tol=1e-12)
    o = oracle.Oracle(p)
    g0 = np.array([0.01, -0.02, 0.005, 0.003, 0.0, -0.004])
    o.set_gradient(g0)
    assert o.solve() == 0
    C = fe_numpy.iso_C(3, (5.0, 2.0))
    assert rel(o.homogenized_stress(), C @ g0) < 1e-13


# ----------------------------------------------------------------------------- solvers
@pytest.mark.parametrize("maker,method", [(cases.thermal, "cg"), (cases.elastic, "cg"), (cases.elastic, "fp")])
def test_solve_against_direct(maker, method):
    p = maker(n=6, dims=(6, 5, 4), count=2, rmin=1.0, rmax=2.0, method=method, tol=1e-12,
              n_it=3000 if method == "fp" else 500)
    o = oracle.Oracle(p)
    h = p.howmany
    g0 = np.linspace(0.3, -0.2, p.n_str) * 0.01
    o.set_gradient(g0)
    o.solve()
    le = np.array(p.L) / np.array(p.dims)
    params = p.groups[0].param_matrix()
    Cs = {ph: fe_numpy.iso_C(h, params[ph]) for ph in range(len(params))}
    u_ref, sig_ref = fe_numpy.solve_periodic(p.ms, Cs, le, h, g0)
    u = o.get_u()
    u = u - u.reshape(-1, h).mean(0)
    assert rel(u, u_ref) < 1e-7
    assert rel(o.homogenized_stress(), sig_ref) < 1e-9


def test_thermal_laminate_closed_form():
    ms = laminate((8, 6, 4), axis=0, widths=(3, 5), phases=(0, 1))
    k = (1.0, 7.0)
    p = Problem(dims=(8, 6, 4), howmany=1, ms=ms,
                groups=[MaterialGroup("LinearThermalIsotropic", [0, 1], {"conductivity": list(k)})], tol=1e-13)
    o = oracle.Oracle(p)
    f = np.array([3, 5]) / 8
    kh = 1.0 / (f[0] / k[0] + f[1] / k[1])
    ka = f[0] * k[0] + f[1] * k[1]
    o.set_gradient([1.0, 0.0, 0.0])
    o.solve()
    assert abs(o.homogenized_stress()[0] - kh) < 1e-10 * kh
    o.set_gradient([0.0, 1.0, 0.0])
    o.solve()
    assert abs(o.homogenized_stress()[1] - ka) < 1e-10 * ka


def elastic_laminate(f, lam, mu, eps):
    """Closed-form laminate (normal x) for isotropic layers, Mandel in/out."""
    f, lam, mu = map(np.asarray, (f, lam, mu))
    e = np.asarray(eps, float)
    a = 1.0 / (lam + 2 * mu)
    s11 = (e[0] + np.sum(f * lam * a) * (e[1] + e[2])) / np.sum(f * a)
    s3 = e[3] / np.sum(f / (2 * mu))
    s4 = e[4] / np.sum(f / (2 * mu))
    e11 = (s11 - lam * (e[1] + e[2])) * a
    tr = e11 + e[1] + e[2]
    s22 = lam * tr + 2 * mu * e[1]
    s33 = lam * tr + 2 * mu * e[2]
    s5 = 2 * mu * e[5]
    return np.array([s11, np.sum(f * s22), np.sum(f * s33), s3, s4, np.sum(f * s5)])


@pytest.mark.parametrize("comp", range(6))
def test_elastic_laminate_closed_form(comp):
    dims = (8, 4, 4)
    ms = laminate(dims, axis=0, widths=(5, 3), phases=(0, 1))
    K, G = (2.0, 30.0), (1.0, 12.0)
    p = Problem(dims=dims, howmany=3, ms=ms,
                groups=[MaterialGroup("LinearElasticIsotropic", [0, 1], {"bulk_modulus": list(K),
                                                                          "shear_modulus": list(G)})], tol=1e-13,
                n_it=500)
    o = oracle.Oracle(p)
    e = np.zeros(6)
    e[comp] = 1.0
    o.set_gradient(e)
    o.solve()
    lam = [K[i] - 2 / 3 * G[i] for i in range(2)]
    expect = elastic_laminate([5 / 8, 3 / 8], lam, G, e)
    assert rel(o.homogenized_stress(), expect) < 1e-10


def test_fp_equals_cg_solution():
    p = cases.elastic(n=8, seed=3, count=2, rmin=1.5, rmax=2.5, tol=1e-12)
    g0 = np.array([0.0, 0.0, 0.0, 0.01, 0.0, 0.0])
    o1 = oracle.Oracle(p)
    o1.set_gradient(g0)
    o1.solve()
    p.method, p.n_it = "fp", 5000
    o2 = oracle.Oracle(p)
    o2.set_gradient(g0)
    o2.solve()
    assert rel(o2.homogenized_stress(), o1.homogenized_stress()) < 1e-9


# ----------------------------------------------------------------------------- materials
def test_j2_material_point_properties():
    prm = np.array([60.0, 25.0, 0.2, 3.0])  # K G sigma_y H
    rng = np.random.default_rng(2)
    for _ in range(20):
        e = rng.standard_normal(6) * 0.01
        s, ep, q = oracle.material_point("J2PlasticityNew_LinearIsotropicHardening", prm, e)
        dev = s.copy()
        dev[:3] -= s[:3].mean()
        if q > 0:  # on the updated yield surface, plastic flow is deviatoric
            assert abs(np.linalg.norm(dev) - np.sqrt(2 / 3) * (prm[2] + prm[3] * q)) < 1e-12
            assert abs(ep[:3].sum()) < 1e-15
        else:
            lam = prm[0] - 2 / 3 * prm[1]
            C = np.zeros((6, 6))
            C[:3, :3] = lam
            C += 2 * prm[1] * np.eye(6)
            assert rel(s, C @ e) < 1e-13
        assert abs(s[:3].mean() - prm[0] * e[:3].sum()) < 1e-12  # pressure = K tr(eps)


def test_j2_uniaxial_strain_closed_form():
    """Homogeneous J2 RVE under a monotonic uniaxial strain path, stepped through
    the solver (state committed per step).  Closed form of linear-hardening J2
    plasticity under proportional loading: with the deviatoric norm
    |s| = 2 G eps sqrt(2/3) - 2 G gamma on the yield surface
    |s| = sqrt(2/3) (sigma_y + H sqrt(2/3) gamma):
      gamma = sqrt(2/3) (2 G eps - sigma_y) / (2 G + 2 H / 3),
      sigma_11 = K eps + sqrt(2/3) |s|,  sigma_22 = K eps - |s| / sqrt(6).
    (Under mixed stress control a homogeneous RVE starts with a zero residual, so
    the reference's CG performs no iteration and its controller applies a single
    elastic-predictor update; the strain path avoids that.)"""
    dims = (4, 4, 4)
    E, nu, sy, H = 70.0, 0.33, 0.25, 1.0
    Kb, Gs = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    p = Problem(dims=dims, howmany=3, ms=np.zeros(dims, np.uint16),
                groups=[MaterialGroup("J2PlasticityNew_LinearIsotropicHardening", [0],
                                      {"bulk_modulus": [Kb], "shear_modulus": [Gs], "yield_stress": [sy],
                                       "isotropic_hardening_parameter": [H]})], tol=1e-12, n_it=200)
    o = oracle.Oracle(p)
    for eps in np.linspace(0.0005, 0.01, 12):
        o.set_gradient([eps, 0, 0, 0, 0, 0])
        o.solve()
        s = o.homogenized_stress()
        gamma = max(0.0, np.sqrt(2 / 3) * (2 * Gs * eps - sy) / (2 * Gs + 2 * H / 3))
        sn = 2 * Gs * eps * np.sqrt(2 / 3) - 2 * Gs * gamma
        expect = np.array([Kb * eps + np.sqrt(2 / 3) * sn, Kb * eps - sn / np.sqrt(6), Kb * eps - sn / np.sqrt(6),
                           0, 0, 0])
        assert np.max(np.abs(s - expect)) < 1e-12 * np.abs(expect).max(), eps
        ep, q = o.get_state()
        assert abs(q.max() - np.sqrt(2 / 3) * gamma) < 1e-14


# ----------------------------------------------------------------------------- reference goldens
def test_mixed_bc_golden_from_reference_input(sphere32_ms):
    """test_MixedBCs.json load case 2: hydrostatic stress -0.05/-0.1 on the sphere32
    RVE.  The reference stored the resulting macro strains as load case 3
    (test_MixedBCs.json:74-75); the PseudoPlastic matrix stays elastic there, and the
    reference ran the nonlinear CG with the secant line search and the mixed-BC
    controller to produce them."""
    with open(os.path.join(GOLDEN, "test_MixedBCs.json")) as f:
        j = json.load(f)
    p = Problem.from_json(j, ms=sphere32_ms)
    lc = p.load_cases[1]
    golden = p.load_cases[2].g0_path
    o = oracle.Oracle(p)
    for t in range(lc.n_steps):
        o.enable_mixed_bc(lc.mbc, t)
        o.solve()
        g = o.macro_gradient()
        assert np.max(np.abs(g[:3] - golden[t][:3])) < 1e-8 * np.abs(golden[t][:3]).max(), (t, g, golden[t])
        assert np.max(np.abs(g[3:])) < 1e-12
        s = o.homogenized_stress()
        assert np.max(np.abs(s - np.array(lc.mbc["stress_path"][t]))) < 1e-9
        o.extrapolate_displacement()


def test_reference_pytest_invariants(sphere32_ms):
    """The invariants of the reference's test/pytest suite on the thermal/elastic inputs."""
    for name in ("test_LinearThermal", "test_LinearElastic"):
        p = cases.reference_input(name, sphere32_ms)
        o = oracle.Oracle(p)
        g0 = p.load_cases[0].g0_path[0]
        o.set_gradient(g0)
        o.solve()
        u = o.get_u()
        # test_displacement_averaging.py:78-85 (zero-mean fluctuation)
        assert np.allclose(u.reshape(-1, p.howmany).mean(0), 0, rtol=1e-5, atol=1e-8)
        ef = o.element_fields()
        # test_loading_to_strain_average.py:35-84
        assert np.allclose(ef["strain_average"], g0)
        # test_strain_stress_averaging.py:91-96
        assert np.allclose(ef["stress"].mean(0), ef["stress_average"], rtol=1e-5, atol=1e-8)
        C = o.homogenized_tangent()
        # test_homogenized_tangent_spd.py:79-95
        assert np.allclose(C, C.T, rtol=1e-5, atol=1e-8)
        assert np.all(np.linalg.eigvalsh(C) > 0)
        # test_homogenization_consistency.py:91-99
        assert np.allclose(ef["stress_average"], C @ ef["strain_average"], rtol=1e-4, atol=1e-1)
