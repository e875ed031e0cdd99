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


def test_bbar_B_matrices():
    """B-bar (matmodel.h:113-140) against the HEX8 and HEX8R B matrices, which
    the test above pins to differentiated shape functions: shear rows and the
    deviatoric part of the normal rows are HEX8's, the trace row is the centre
    (HEX8R) one at every Gauss point, and that equals the Gauss mean of the
    HEX8 trace rows.  Thermal B-bar is HEX8."""
    L = (1.0, 0.8, 1.3)
    small = dict(n=4, L=L, count=1, rmin=1.0, rmax=1.0)
    Bh = oracle.Oracle(cases.elastic(fe="HEX8", **small)).B()
    Bc = oracle.Oracle(cases.elastic(fe="HEX8R", **small)).B()[0]
    Bb = oracle.Oracle(cases.elastic(fe="BBAR", **small)).B()
    assert Bb.shape == Bh.shape == (8, 6, 24)
    assert np.array_equal(Bb[:, 3:], Bh[:, 3:])
    tr_h, tr_b, tr_c = Bh[:, :3].sum(1), Bb[:, :3].sum(1), Bc[:3].sum(0)
    for gp in range(8):
        assert np.max(np.abs(tr_b[gp] - tr_c)) < 1e-12
        assert np.max(np.abs((Bb[gp, :3] - tr_b[gp] / 3) - (Bh[gp, :3] - tr_h[gp] / 3))) < 1e-12
    assert np.max(np.abs(tr_h.mean(0) - tr_c)) < 1e-12
    Th = oracle.Oracle(cases.thermal(fe="HEX8", **small)).B()
    Tb = oracle.Oracle(cases.thermal(fe="BBAR", **small)).B()
    assert np.array_equal(Th, Tb)


def test_bbar_stiffness_energy_identity():
    """For sigma = K tr(eps) m + 2 G dev(eps), B-bar removes the bulk energy of
    the trace's deviation from its element mean:
        Ke_HEX8 - Ke_BBAR = K w sum_gp (b_gp - b_mean)^T (b_gp - b_mean),
    b_gp = trace row of the HEX8 B at Gauss point gp, w = v_e / 8."""
    L = (1.0, 0.9, 1.1)
    small = dict(n=4, L=L, count=1, rmin=1.0, rmax=1.0)
    ph = cases.elastic(fe="HEX8", **small)
    oh = oracle.Oracle(ph)
    ob = oracle.Oracle(cases.elastic(fe="BBAR", **small))
    b = oh.B()[:, :3].sum(1)
    db = b - b.mean(0)
    w = np.prod(np.array(L) / 4) / 8
    params = ph.groups[0].param_matrix()
    for phase in range(2):
        K = params[phase][0]
        expect = K * w * db.T @ db
        diff = oh.phase_stiffness(phase) - ob.phase_stiffness(phase)
        assert rel(diff, expect) < 1e-12
        # and the identity is not trivially satisfied
        assert np.abs(expect).max() > 1e-3 * np.abs(oh.phase_stiffness(phase)).max()
    kh, kb = oh.ref_element_stiffness(), ob.ref_element_stiffness()
    assert np.max(np.abs(kh - kb)) > 0


@pytest.mark.parametrize("comp", [0, 3])
def test_bbar_laminate_closed_form(comp):
    """Laminate with interfaces on voxel faces: the exact solution has a
    constant strain per element, so B-bar changes nothing and reproduces the
    closed form."""
    dims = (8, 4, 4)
    ms = laminate(dims, axis=0, widths=(5, 3), phases=(0, 1))
    K, G = (2.0, 30.0), (1.0, 12.0)
    p = Problem(dims=dims, howmany=3, ms=ms, fe_type="BBAR",
                groups=[MaterialGroup("LinearElasticIsotropic", [0, 1], {"bulk_modulus": list(K),
                                                                          "shear_modulus": list(G)})], tol=1e-13,
                n_it=500)
    o = oracle.Oracle(p)
    e = np.zeros(6)
    e[comp] = 1.0
    o.set_gradient(e)
    o.solve()
    lam = [K[i] - 2 / 3 * G[i] for i in range(2)]
    assert rel(o.homogenized_stress(), elastic_laminate([5 / 8, 3 / 8], lam, G, e)) < 1e-10


def test_bbar_relieves_volumetric_locking():
    """Ke_BBAR <= Ke_HEX8 in the energy sense, so the homogenized shear
    stiffness with B-bar lies below HEX8's.  With a nearly incompressible
    matrix (nu = 0.4999) around a stiff inclusion the gap is several percent:
    the volumetric locking of full integration that B-bar relieves."""
    E, nu = 1.0, 0.4999
    K, G = E / (3 * (1 - 2 * nu)), E / (2 * (1 + nu))
    res = {}
    for fe in ("HEX8", "BBAR"):
        p = cases.elastic(n=8, fe=fe, K=(K, 50 * K), G=(G, 50 * G), count=1, rmin=2.5, rmax=2.5, seed=1,
                          tol=1e-12, n_it=3000)
        o = oracle.Oracle(p)
        o.set_gradient([0.0, 0.0, 0.0, 1.0, 0.0, 0.0])
        o.solve()
        res[fe] = o.homogenized_stress()[3]
    assert 0.02 < (res["HEX8"] - res["BBAR"]) / res["HEX8"] < 0.2


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


# ----------------------------------------------------------------------------- power-law pseudo-plasticity
PNL = dict(K=62.5, G=28.8462, sy=0.1, n=0.2, e0=0.01)  # test_PseudoPlastic.json's matrix


def _pnl_params(K=PNL["K"], G=PNL["G"], sy=PNL["sy"], n=PNL["n"], e0=PNL["e0"]):
    return np.array([K, G, sy, n, e0])


def test_pseudo_nonlinear_power_law():
    """PseudoPlasticNonLinearHardening (PseudoPlastic.h:125-168) as a deformation
    theory: beyond its threshold the von Mises stress sqrt(3/2)|dev sigma|
    equals sigma_y (eps_eq / eps_0)^n with eps_eq = sqrt(2/3)|dev eps|, the
    pressure stays K tr(eps), and the stress deviator is parallel to the strain
    deviator; below it the law is LinearElasticIsotropic with the same K, G."""
    rng = np.random.default_rng(3)
    K, G, sy, n, e0 = (PNL[k] for k in ("K", "G", "sy", "n", "e0"))
    crit = e0 * (sy / (3 * G * e0)) ** (1 / (1 - n))
    for scale in (0.2 * crit, 0.9 * crit, 1.5 * crit, 10 * crit, 100 * crit):
        d = rng.standard_normal(6)
        d[:3] -= d[:3].mean()
        eq_dir = np.sqrt(2 / 3) * np.linalg.norm(d)
        e = d * (scale / eq_dir) + 1e-3 * np.array([1, 1, 1, 0, 0, 0]) * rng.standard_normal()
        s, _, _ = oracle.material_point("PseudoPlasticNonLinearHardening", _pnl_params(), e)
        tr = e[:3].sum()
        de = e - np.r_[np.full(3, tr / 3), np.zeros(3)]
        ds = s - np.r_[np.full(3, s[:3].mean()), np.zeros(3)]
        eps_eq = np.sqrt(2 / 3) * np.linalg.norm(de)
        assert abs(s[:3].mean() - K * tr) < 1e-13 * max(abs(K * tr), 1.0)
        if eps_eq <= crit:
            se, _, _ = oracle.material_point("LinearElasticIsotropic", np.array([K, G]), e)
            assert np.max(np.abs(s - se)) < 1e-13 * np.abs(se).max()
        else:
            assert abs(np.sqrt(1.5) * np.linalg.norm(ds) - sy * (eps_eq / e0) ** n) < 1e-12 * sy * (eps_eq / e0) ** n
            assert np.max(np.abs(np.cross(ds[:3], de[:3]))) < 1e-12 * np.linalg.norm(ds) * np.linalg.norm(de)
            assert np.max(np.abs(ds * np.linalg.norm(de) - de * np.linalg.norm(ds))) < 1e-12 * np.linalg.norm(ds) * np.linalg.norm(de)


def test_pseudo_nonlinear_continuous_at_threshold():
    """The threshold eps_crit = eps_0 (sigma_y / (3 G eps_0))^(1/(1-n))
    (PseudoPlastic.h:137-140) is where the elastic 3 G eps_eq meets the power
    law: the stress is continuous across it for several exponents."""
    d = np.array([0.7, -0.2, -0.5, 0.3, -0.1, 0.25])
    for n in (0.1, 0.2, 0.5):
        K, G, sy, e0 = PNL["K"], PNL["G"], PNL["sy"], PNL["e0"]
        crit = e0 * (sy / (3 * G * e0)) ** (1 / (1 - n))
        u = d / (np.sqrt(2 / 3) * np.linalg.norm(d))
        lo, _, _ = oracle.material_point("PseudoPlasticNonLinearHardening", _pnl_params(n=n), u * crit * (1 - 1e-12))
        hi, _, _ = oracle.material_point("PseudoPlasticNonLinearHardening", _pnl_params(n=n), u * crit * (1 + 1e-12))
        assert np.max(np.abs(hi - lo)) < 1e-9 * np.abs(lo).max(), n


def test_pseudo_nonlinear_plastic_flag():
    """PseudoPlastic::postprocess flags (PseudoPlastic.h:56-64, :153-167) on a
    two-phase PseudoPlasticNonLinearHardening group with a stiff second phase:
    phase 1 carries local index 1, phase 0 local index 0; a plastic Gauss point
    adds the group's phase count 2, so elastic / plastic points give 0 / 2 in
    phase 0 and 1 / 3 in phase 1, element means in between."""
    from fans_amd.problem import MaterialGroup, Problem
    n = 8
    ms = np.zeros((n, n, n), dtype=np.uint8)
    ms[2:5, 2:5, 2:5] = 1
    grp = MaterialGroup("PseudoPlasticNonLinearHardening", [0, 1],
                        {"bulk_modulus": [62.5, 222.0], "shear_modulus": [28.8462, 166.0],
                         "yield_stress": [0.1, 50.0], "hardening_exponent": [0.2, 0.2], "eps_0": [0.01, 0.01]})
    p = Problem(dims=(n, n, n), howmany=3, ms=ms, groups=[grp], fe_type="HEX8", method="cg", tol=1e-10, n_it=300)
    o = oracle.Oracle(p)
    for g, expect0, expect1 in (([1e-5, 0, 0, 0, 0, 0], 0.0, 1.0), ([0.02, -0.01, -0.01, 0, 0, 0], 2.0, 1.0)):
        o.set_gradient(np.array(g))
        o.solve()
        f = o.plastic_flag().reshape(n, n, n)
        # element (x, y, z) takes the phase of its node (x, y, z); its Gauss
        # points give 0 / 2 (phase 0) or 1 / 3 (phase 1), means in eighths
        f8 = f * 8
        assert np.allclose(f8, np.round(f8), atol=1e-12)
        assert np.all((f[ms == 0] >= 0) & (f[ms == 0] <= 2)) and np.all((f[ms == 1] >= 1) & (f[ms == 1] <= 3))
        assert np.abs(np.median(f[ms == 0]) - expect0) < 1e-12, g
        assert np.abs(f[3, 3, 3] - expect1) < 1e-12, g


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


# ----------------------------------------------------------------------------- finite strain
LARGE_KG = {"bulk_modulus": [62.5, 222.222], "shear_modulus": [28.8462, 166.6667]}


def large_problem(model, dims=(6, 4, 8), ms=None, tol=1e-10, n_it=300, fe="HEX8", L=(1.0, 1.0, 1.0)):
    if ms is None:
        ms = np.zeros(dims, dtype=np.uint16)
        ms[1:3, 1:3, 2:5] = 1
    return Problem(dims=tuple(ms.shape), howmany=3, ms=ms, groups=[MaterialGroup(model, [0, 1], dict(LARGE_KG))],
                   strain_type="large", fe_type=fe, tol=tol, n_it=n_it, L=L)


def hyperelastic_energy(model, K, G, F):
    """Stored energies whose F-gradients are the models' P: Saint-Venant
    Kirchhoff W = lam/2 tr(E)^2 + mu E:E; compressible Neo-Hooke
    W = mu/2 (tr C - 3) - mu ln J + lam/2 (ln J)^2."""
    lam, mu = K - 2 * G / 3, G
    C = F.T @ F
    if model == "SaintVenantKirchhoff":
        E = 0.5 * (C - np.eye(3))
        return 0.5 * lam * np.trace(E) ** 2 + mu * np.sum(E * E)
    lj = np.log(np.linalg.det(F))
    return 0.5 * mu * (np.trace(C) - 3) - mu * lj + 0.5 * lam * lj ** 2


def test_large_strain_B_matrix():
    """LargeStrainMechModel::Compute_B: B_F u_e of a linear displacement field
    u = H X is H row major (F_iJ - delta_iJ = du_i/dX_J) at every Gauss point;
    a transposed layout fails for the non-symmetric H used here."""
    L = (1.0, 0.8, 1.3)
    o = oracle.Oracle(large_problem("SaintVenantKirchhoff", L=L))
    B = o.B()
    assert B.shape == (8, 9, 24)
    le = np.array(L) / np.array((6, 4, 8))
    H = np.array([[0.1, 0.7, -0.3], [0.2, -0.4, 0.9], [-0.6, 0.05, 0.3]])
    X = np.array([[(a & 1) * le[0], ((a >> 1) & 1) * le[1], ((a >> 2) & 1) * le[2]] for a in range(8)])
    ue = (X @ H.T).reshape(-1)
    for gp in range(8):
        assert rel(B[gp] @ ue, H.reshape(-1)) < 1e-12
        assert np.max(np.abs(B[gp] @ np.tile([1.0, 2.0, 3.0], 8))) < 1e-12


@pytest.mark.parametrize("model", ["SaintVenantKirchhoff", "CompressibleNeoHookean"])
def test_hyperelastic_points(model):
    """P(F) = dW/dF (central differences of the stored energy), objectivity
    P(QF) = Q P(F), and the small-strain limit P(I + eps H) / eps -> the
    linear elastic stress of sym(H) with the same K, G."""
    K, G = 62.5, 28.8462
    prm = np.array([K, G])
    rng = np.random.default_rng(4)
    for _ in range(5):
        F = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        P, _, _ = oracle.material_point(model, prm, F.reshape(-1))
        P = P.reshape(3, 3)
        h = 1e-6
        dW = np.zeros((3, 3))
        for i in range(3):
            for j in range(3):
                Fp, Fm = F.copy(), F.copy()
                Fp[i, j] += h
                Fm[i, j] -= h
                dW[i, j] = (hyperelastic_energy(model, K, G, Fp) - hyperelastic_energy(model, K, G, Fm)) / (2 * h)
        assert rel(P, dW) < 1e-7
        a = rng.standard_normal(3)
        a /= np.linalg.norm(a)
        th = 0.7
        A = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        Q = np.eye(3) + np.sin(th) * A + (1 - np.cos(th)) * A @ A
        PQ, _, _ = oracle.material_point(model, prm, (Q @ F).reshape(-1))
        assert rel(PQ.reshape(3, 3), Q @ P) < 1e-12
    H = rng.standard_normal((3, 3))
    eps = 1e-7
    P, _, _ = oracle.material_point(model, prm, (np.eye(3) + eps * H).reshape(-1))
    S = 0.5 * (H + H.T)
    r2 = np.sqrt(2)
    e6 = np.array([S[0, 0], S[1, 1], S[2, 2], r2 * S[0, 1], r2 * S[0, 2], r2 * S[1, 2]])
    s6, _, _ = oracle.material_point("LinearElasticIsotropic", prm, e6)
    lin = np.array([[s6[0], s6[3] / r2, s6[4] / r2], [s6[3] / r2, s6[1], s6[5] / r2], [s6[4] / r2, s6[5] / r2, s6[2]]])
    assert rel(P.reshape(3, 3) / eps, lin) < 1e-5


@pytest.mark.parametrize("model", ["SaintVenantKirchhoff", "CompressibleNeoHookean"])
def test_large_strain_reference_stiffness(model):
    """get_reference_stiffness of the finite strain models: dP/dF at F = I
    (central differences of P) averaged over the phases, with the reference's
    sum over P <= Q only (LargeStrainMechModel.h:142-170), which halves the
    columns of the off-diagonal F_kL."""
    o = oracle.Oracle(large_problem(model))
    kref = o.reference_stiffness()
    exact = np.zeros((9, 9))
    for K, G in zip(LARGE_KG["bulk_modulus"], LARGE_KG["shear_modulus"]):
        for c in range(9):
            h = 1e-6
            Fp, Fm = np.eye(3).reshape(-1).copy(), np.eye(3).reshape(-1).copy()
            Fp[c] += h
            Fm[c] -= h
            Pp, _, _ = oracle.material_point(model, [K, G], Fp)
            Pm, _, _ = oracle.material_point(model, [K, G], Fm)
            exact[:, c] += (Pp - Pm) / (2 * h) / 2
    half = np.array([1.0 if c in (0, 4, 8) else 0.5 for c in range(9)])
    assert rel(kref, exact * half[None, :]) < 1e-8
    assert np.allclose(kref, kref.T)


@pytest.mark.parametrize("model", ["SaintVenantKirchhoff", "CompressibleNeoHookean"])
def test_large_strain_homogeneous(model):
    """Homogeneous RVE: the fluctuation stays zero, no iteration is needed and
    P_bar = P(F_bar)."""
    ms = np.zeros((6, 4, 8), dtype=np.uint16)
    p = Problem(dims=ms.shape, howmany=3, ms=ms, strain_type="large", tol=1e-10, n_it=100,
                groups=[MaterialGroup(model, [0], {"bulk_modulus": [62.5], "shear_modulus": [28.8462]})])
    o = oracle.Oracle(p)
    F = np.array([1.1, 0.05, 0.0, -0.02, 0.95, 0.03, 0.01, 0.0, 1.2])
    o.set_gradient(F)
    assert o.solve() == 0
    P, _, _ = oracle.material_point(model, [62.5, 28.8462], F)
    assert rel(o.homogenized_stress(), P) < 1e-13


def test_large_strain_mixed_bc_golden(sphere32_ms):
    """test_MixedBCs_LargeStrain.json (CompressibleNeoHookean on sphere32,
    nonlinear CG, line search, mixed BCs starting from F = I): load case 2
    prescribes P_bar = (0, ..., 74.757449712464) on every component, the
    homogenized stress that the reference printed for load case 1's last step;
    load case 3 of that file stores the deformation gradient the reference
    reached, diag(0.832542244829, 0.832542244829, 2.000000000001)."""
    with open(os.path.join(GOLDEN, "test_MixedBCs_LargeStrain.json")) as f:
        j = json.load(f)
    p = Problem.from_json(j, ms=sphere32_ms)
    assert p.n_str == 9
    lc = p.load_cases[1]
    golden = p.load_cases[2].g0_path[0]
    o = oracle.Oracle(p)
    o.enable_mixed_bc(lc.mbc, 0)
    o.solve()
    assert np.max(np.abs(o.macro_gradient() - golden)) < 1e-10
    assert np.max(np.abs(o.homogenized_stress() - np.array(lc.mbc["stress_path"][0]))) < 1e-9


# ----------------------------------------------------------------------------- error measures
@pytest.mark.parametrize("measure,order", [("L1", 1), ("L2", 2), ("Linfinity", np.inf)])
@pytest.mark.parametrize("maker", [cases.thermal, cases.elastic])
def test_error_norms_against_numpy(maker, measure, order):
    """compute_error (solver.h:414-430): err_all[0] is the L1 / L2 / L-inf norm
    of the first residual over the unpadded field, checked with numpy's vector
    norms of the oracle's own residual at u = 0 (a dropped square root, a
    missing absolute value or a max in place of a sum fails here)."""
    p = maker(n=6, dims=(6, 5, 4), count=2, rmin=1.0, rmax=2.0, measure=measure, n_it=0)
    o = oracle.Oracle(p)
    g0 = np.linspace(0.3, -0.2, p.n_str) * 0.01
    o.set_gradient(g0)
    assert o.solve() == 0
    r0 = o.residual(np.zeros(p.dims + (p.howmany,)))
    ref = np.linalg.norm(r0.reshape(-1), ord=order)
    assert abs(o.error_history()[0] - ref) < 1e-13 * ref


def test_relative_error_stop_rule():
    """Relative error (solver.h:433-450): err / err_0 after the first iteration
    and the constant 100 at iteration 0.  A solve stops at the first iteration
    whose ratio meets the tolerance; a tolerance above 100 stops before any
    iteration, one below 100 does not."""
    p = cases.elastic(n=8, seed=2, count=2, rmin=1.5, rmax=2.5, measure="L2", tol=1e-6, n_it=300)
    p.err_type = "relative"
    g0 = np.array([0.01, -0.02, 0.005, 0.01, 0.0, -0.005])
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    it = o.solve()
    h = o.error_history()
    ratio = h / h[0]
    first = int(np.argmax(ratio[1:] <= 1e-6)) + 1
    assert it == first and len(h) == it + 1
    for tol, expect_zero in ((150.0, True), (50.0, False)):
        p.tol = tol
        o = oracle.Oracle(p)
        o.set_gradient(g0)
        assert (o.solve() == 0) == expect_zero


# ----------------------------------------------------------------------------- triclinic models
def _triclinic_props(C, sym):
    """material_properties keys of LinearElasticTriclinic / LinearThermalTriclinic
    (LinearElastic.h:84-90, LinearThermal.h:59-62): one key per upper-triangle
    entry, named by its 1-based (row, column)."""
    n = C[0].shape[0]
    return {f"{sym}_{i + 1}{j + 1}": [float(c[i, j]) for c in C] for i in range(n) for j in range(i, n)}


@pytest.mark.parametrize("model,h", [("LinearElasticTriclinic", 3), ("LinearThermalTriclinic", 1)])
def test_triclinic_homogeneous_and_reference(model, h):
    """A homogeneous triclinic RVE: the solve needs no iteration and returns
    Sigma = C g0 with C the symmetric matrix whose (i, j) entry is the key
    C_ij (K_ij); the reference stiffness of a two-phase group is the plain mean
    of the phase matrices (LinearElastic.h:146-154, LinearThermal.h:101)."""
    rng = np.random.default_rng(21)
    n = 6 if h == 3 else 3
    A = rng.standard_normal((n, n))
    C0 = A @ A.T + n * np.eye(n)
    B = rng.standard_normal((n, n))
    C1 = B @ B.T + 2 * n * np.eye(n)
    sym = "C" if h == 3 else "K"
    dims = (6, 4, 8)
    p = Problem(dims=dims, howmany=h, ms=np.zeros(dims, np.uint16),
                groups=[MaterialGroup(model, [0, 1], _triclinic_props([C0, C1], sym))], tol=1e-12)
    o = oracle.Oracle(p)
    g0 = rng.standard_normal(n) * 0.01
    o.set_gradient(g0)
    assert o.solve() == 0
    assert rel(o.homogenized_stress(), C0 @ g0) < 1e-13
    assert rel(o.reference_stiffness(), 0.5 * (C0 + C1)) < 1e-15
    # a transposed key order (column-wise fill) would be caught: C0 is not symmetric under it
    assert np.max(np.abs(C0 - C0.T)) < 1e-12


@pytest.mark.parametrize("model,h", [("LinearElasticTriclinic", 3), ("LinearThermalTriclinic", 1)])
def test_triclinic_residual_against_sparse_assembly(model, h):
    """Two anisotropic phases: the oracle's residual and linear matvec against the
    independent sparse assembly (numerical shape-function derivatives, 3x3x3
    Gauss rule) with the same full matrices."""
    rng = np.random.default_rng(22)
    n = 6 if h == 3 else 3
    Cs = []
    for shift in (1.0, 5.0):
        A = rng.standard_normal((n, n))
        Cs.append(A @ A.T + shift * np.eye(n))
    dims = (4, 5, 6)
    from fans_amd.microstructure import random_spheres
    ms = random_spheres(dims, 2, 1.0, 2.0, seed=4)
    p = Problem(dims=dims, howmany=h, ms=ms,
                groups=[MaterialGroup(model, [0, 1], _triclinic_props(Cs, "C" if h == 3 else "K"))])
    o = oracle.Oracle(p)
    le = np.array(p.L) / np.array(dims)
    K, F = fe_numpy.assemble(ms, {0: Cs[0], 1: Cs[1]}, le, h)
    u = rng.standard_normal(dims + (h,))
    g0 = rng.standard_normal(n)
    o.set_gradient(g0)
    assert rel(o.residual(u), (K @ u.reshape(-1) + F @ g0).reshape(u.shape)) < 1e-8
    assert rel(o.residual(u, linear_kd=True), (K @ u.reshape(-1)).reshape(u.shape)) < 1e-8


# ----------------------------------------------------------------------------- linear-hardening pseudo-plasticity
def test_pseudo_linear_hardening_beyond_yield():
    """PseudoPlasticLinearHardening (PseudoPlastic.h:78-116) as deformation
    plasticity with linear hardening H: with eps_eq = sqrt(2/3)|dev eps| and
    sigma_eq = sqrt(3/2)|dev sigma|, the law is elastic (sigma_eq = 3 G eps_eq)
    up to sigma_y and beyond it follows the series combination of the elastic
    3 G and the hardening H, sigma_eq = sigma_y + 3GH/(3G+H) (eps_eq - sigma_y/(3G));
    the pressure stays K tr(eps) and dev sigma is parallel to dev eps."""
    K, G, sy, H = 62.5, 28.8462, 0.1, 0.5
    prm = np.array([K, G, sy, H])
    eq_y = sy / (3 * G)
    slope = 3 * G * H / (3 * G + H)
    rng = np.random.default_rng(8)
    for scale in (0.3, 0.99, 1.01, 2.0, 10.0, 200.0):
        d = rng.standard_normal(6)
        d[:3] -= d[:3].mean()
        d *= scale * eq_y / (np.sqrt(2 / 3) * np.linalg.norm(d))
        e = d + np.r_[np.full(3, 1e-3 * rng.standard_normal()), np.zeros(3)]
        s, _, _ = oracle.material_point("PseudoPlasticLinearHardening", prm, e)
        tr = e[:3].sum()
        de = e - np.r_[np.full(3, tr / 3), np.zeros(3)]
        ds = s - np.r_[np.full(3, s[:3].mean()), np.zeros(3)]
        eps_eq = np.sqrt(2 / 3) * np.linalg.norm(de)
        sig_eq = np.sqrt(1.5) * np.linalg.norm(ds)
        expect = 3 * G * eps_eq if eps_eq <= eq_y else sy + slope * (eps_eq - eq_y)
        assert abs(sig_eq - expect) < 1e-12 * expect, scale
        assert abs(s[:3].mean() - K * tr) < 1e-13 * max(abs(K * tr), 1.0)
        assert np.max(np.abs(ds * np.linalg.norm(de) - de * np.linalg.norm(ds))) < \
            1e-12 * np.linalg.norm(ds) * np.linalg.norm(de)


# ----------------------------------------------------------------------------- J2ViscoPlastic
VISCO = dict(K=62.5, G=28.8462, sy=0.1, Kh=0.0, Hk=0.0, eta=1.0, dt=0.01, sinf=0.15, delta=1000.0)  # test_J2Plasticity.json


def _visco_params(nonlinear, **kw):
    v = dict(VISCO, **kw)
    p = [v["K"], v["G"], v["sy"], v["Kh"], v["Hk"], v["eta"], v["dt"]]
    return np.array(p + ([v["sinf"], v["delta"]] if nonlinear else []))


def test_visco_linear_closed_form():
    """Linear law, one call from a virgin state: the trial deviator n = dev(2G e)
    / |.|, gamma = f / (2G + 2/3 (K_iso + H_kin) + eta / dt) with f = |dev| -
    sqrt(2/3) sigma_y, so the returned deviatoric norm exceeds the yield radius
    by exactly gamma (eta / dt + 2/3 (K_iso + H_kin)); flow is deviatoric, the
    pressure is K tr(eps), psi = sqrt(2/3) gamma and psi_bar = -gamma n."""
    rng = np.random.default_rng(12)
    for Kh, Hk in ((0.0, 0.0), (1.5, 0.7)):
        prm = _visco_params(False, Kh=Kh, Hk=Hk)
        K, G, sy, eta, dt = prm[0], prm[1], prm[2], prm[5], prm[6]
        for _ in range(20):
            e = rng.standard_normal(6) * 0.004
            s, ep, psi, pb = oracle.visco_point(10, prm, e)
            dev_e = e.copy()
            dev_e[:3] -= e[:3].mean()
            nt = np.linalg.norm(2 * G * dev_e)
            f = nt - np.sqrt(2 / 3) * sy
            gamma = max(0.0, f / (2 * G + 2 / 3 * (Kh + Hk) + eta / dt))
            dev = s.copy()
            dev[:3] -= s[:3].mean()
            assert abs(np.linalg.norm(dev) - (nt - 2 * G * gamma)) < 1e-12 * nt
            if gamma > 0:
                assert abs(np.linalg.norm(dev) - np.sqrt(2 / 3) * sy - gamma * (eta / dt + 2 / 3 * (Kh + Hk))) < 1e-12
            assert abs(psi - np.sqrt(2 / 3) * gamma) < 1e-15
            assert np.max(np.abs(pb + ep)) < 1e-15  # psi_bar = -gamma n, ep = gamma n from a virgin state
            assert abs(ep[:3].sum()) < 1e-15
            assert abs(s[:3].mean() - K * e[:3].sum()) < 1e-12


def test_visco_nonlinear_iteration():
    """Nonlinear (saturation) law, gamma of one call from a virgin state against
    the root of g(x) = f - x den - sqrt(2/3) (s_inf - s_y) (1 - e^(-d sqrt(2/3) x))
    found independently by bisection.  The reference's Newton loop uses the slope
    -den alone (its (2 / 3) * ... term is integer division, J2Plasticity.h:217)
    and stops at the first increment that is not above 1e-10 -- signed, not
    absolute -- so from gamma = 0 it overshoots the root once and stops:
    gamma = x1 + g(x1) / den with x1 = f / den, within O(r^2) of the root for
    r = sqrt(2/3)^2 (s_inf - s_y) d / den.  With the reference's exponent
    (1000, r = 0.2) it stays visibly off the root; with a small one (5) it is
    within r^2."""
    from scipy.optimize import brentq
    e = np.array([0.004, -0.002, -0.002, 0.0008, 0.0, 0.0])  # d sqrt(2/3) gamma ~ 1 at d = 1000
    for dl, close in ((5.0, True), (1000.0, False)):
        prm = _visco_params(True, delta=dl)
        K, G, sy, Kh, Hk, eta, dt, sinf, _ = prm
        den = 2 * G + 2 / 3 * (Kh + Hk) + eta / dt
        c = np.sqrt(2 / 3)
        dev_e = e.copy()
        dev_e[:3] -= e[:3].mean()
        f = np.linalg.norm(2 * G * dev_e) - c * sy
        g = lambda x: f - x * den - c * (sinf - sy) * (1.0 - np.exp(-dl * c * x))
        root = brentq(g, 0.0, f / den, xtol=1e-18)
        _, _, psi, _ = oracle.visco_point(11, prm, e)
        gamma = psi / c
        r = c * c * (sinf - sy) * dl / den
        assert gamma < f / den and abs(gamma - root) <= 2 * r * r * root
        if close:
            assert abs(gamma - root) < 1e-5 * root
        else:
            assert abs(gamma - root) > 1e-4 * root  # a converged Newton solve would sit on the root


def test_visco_accumulates_per_call():
    """psi and psi_bar gain gamma at every get_sigma call within a step
    (J2Plasticity.h:103-104) while the plastic strain is overwritten: two calls
    at the same strain double the increments and leave ep unchanged."""
    prm = _visco_params(False, Kh=0.5, Hk=0.2)
    e = np.array([0.01, -0.004, -0.004, 0.003, 0.001, 0.0])
    s1, ep1, psi1, pb1 = oracle.visco_point(10, prm, e)
    s2, ep2, psi2, pb2 = oracle.visco_point(10, prm, e, psi=psi1, pb=pb1)
    assert psi1 > 0
    assert np.array_equal(s1, s2) and np.array_equal(ep1, ep2)
    assert abs(psi2 - 2 * psi1) < 1e-16 and np.max(np.abs(pb2 - 2 * pb1)) < 1e-16


def test_visco_homogeneous_rve_steps():
    """Homogeneous RVE under a strain path: the residual vanishes at the start,
    so each solve makes exactly one get_sigma call per Gauss point (the initial
    residual) before committing, and the homogenized stress after it one more;
    the state then follows a single-point integration with those calls: the
    current psi / psi_bar carry every call's increment and become the committed
    ones at the next solve's commit."""
    dims = (4, 4, 4)
    for kind, name in ((10, "J2ViscoPlastic_LinearIsotropicHardening"),
                       (11, "J2ViscoPlastic_NonLinearIsotropicHardening")):
        prm = _visco_params(kind == 11, Kh=0.4, Hk=0.3)
        keys = ("bulk_modulus", "shear_modulus", "yield_stress", "isotropic_hardening_parameter",
                "kinematic_hardening_parameter", "viscosity", "time_step", "saturation_stress",
                "saturation_exponent")
        props = {k: ([v] if k != "time_step" else v) for k, v in zip(keys, prm)}
        p = Problem(dims=dims, howmany=3, ms=np.zeros(dims, np.uint16), groups=[MaterialGroup(name, [0], props)],
                    tol=1e-12, n_it=100)
        o = oracle.Oracle(p)
        ep_t, psi_t, pb_t = np.zeros(6), 0.0, np.zeros(6)  # committed
        psi_c, pb_c = 0.0, np.zeros(6)                       # current
        yielded = 0
        for eps in np.linspace(0.001, 0.01, 8):
            g0 = np.array([eps, -0.3 * eps, -0.3 * eps, 0.2 * eps, 0, 0])
            o.set_gradient(g0)
            assert o.solve() == 0
            # the solve's call, then updateInternalVariables
            _, ep, psi_c, pb_c = oracle.visco_point(kind, prm, g0, ep_t, psi_t, pb_t, psi_c, pb_c)
            ep_t, psi_t, pb_t = ep, psi_c, pb_c.copy()
            epo, qo = o.get_state(committed=True)
            assert np.max(np.abs(epo - ep_t)) < 1e-15 and np.max(np.abs(qo - psi_t)) < 1e-15
            assert np.max(np.abs(o.get_kinematic_state(committed=True) - pb_t)) < 1e-15
            # the homogenized stress: one more call per point after the commit
            so = o.homogenized_stress()
            s2, _, psi_c, pb_c = oracle.visco_point(kind, prm, g0, ep_t, psi_t, pb_t, psi_c, pb_c)
            assert np.max(np.abs(so - s2)) < 1e-13 * np.max(np.abs(s2))
            yielded += psi_t > 0
        assert yielded >= 4
        eq, qc = o.get_state(committed=False)
        assert np.max(np.abs(qc - psi_c)) < 1e-15
        assert np.max(np.abs(o.get_kinematic_state(committed=False) - pb_c)) < 1e-15
