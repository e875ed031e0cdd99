"""GPU parity at the sizes the benchmark configurations run.

Every kernel variant the headline and the configurations launch is compared
with the oracle (or numpy) at its own shape: the compile-time z passes for
n_z = 64 / 128 / 256, the register four-step y and x+Green passes at lengths
128 and 256, the generic Stockham lengths 48 / 80 / 96 of the batch, the
stencil tiles with their partial last tile at 256 (= 17 * 15 + 1), and the FP32
twins of the mixed form's inner solve.

Tolerances: FP64 operators 1e-12 relative to the field's max (summation order
only); FP32 operators 1e-5 (single-precision rounding of O(10) operations per
output, well below any indexing or layout error, which is O(1)); FP64 solves
1e-10 on homogenized quantities with CG iterations within +-2; the mixed form
1e-8 against FP64.
"""
import numpy as np
import pytest

import oracle
from fans_amd import gpu, workloads
from fans_amd.batch import config5_rves
from tests import cases

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.parametrize("shape", [(8, 64, 64, 3), (4, 128, 128, 3), (4, 256, 256, 3), (2, 256, 256, 1),
                                   (256, 4, 64, 3), (128, 8, 256, 1), (6, 96, 80, 3), (4, 48, 96, 1),
                                   (4, 80, 48, 3)])
def test_fft_lengths(shape):
    """Standalone r2c / c2r against numpy for every power-of-two length up to
    256 along z and y (x through the strided register kernel) and the batch's
    radix-3/5 lengths (along y 48, 80 and 96 through the mixed-radix register
    kernels)."""
    say(f"fft shape={shape}")
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape)
    X = gpu.fft_r2c(x)
    ref = np.fft.rfftn(x, axes=(0, 1, 2))
    assert relerr(X, ref) < 1e-13
    y = gpu.fft_c2r(ref, shape[2])
    assert relerr(y, np.fft.irfftn(ref, s=shape[:3], axes=(0, 1, 2)) * np.prod(shape[:3])) < 1e-13


_SIZE_CASES = [("elastic", (256, 256, 256)), ("thermal", (256, 256, 256)), ("elastic", (128, 128, 128)),
               ("thermal", (128, 128, 128)), ("elastic", (96, 96, 96)), ("thermal", (96, 96, 96)),
               ("elastic", (80, 80, 80)), ("thermal", (80, 80, 80)), ("elastic", (48, 48, 48)),
               ("thermal", (64, 64, 64)), ("elastic", (64, 64, 64)), ("thermal", (48, 48, 48))]


def _size_problem(kind, dims):
    n = dims[0]
    s = n / 256.0
    count = max(2, int(round(150 * s ** 3)))
    mk = cases.elastic if kind == "elastic" else cases.thermal
    return mk(dims=dims, seed=7, count=count, rmin=max(2.0, 8 * s), rmax=max(3.0, 20 * s))


@pytest.mark.parametrize("kind,dims", _SIZE_CASES)
def test_operators_at_size(kind, dims):
    """FP64: Green operator, residual, linear matvec and convolution; FP32 (the
    mixed form's inner solve): linear matvec and convolution -- against the
    oracle on the same grid."""
    p = _size_problem(kind, dims)
    say(f"operators at size {kind} dims={dims}")
    o = oracle.Oracle(p)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(p.dims + (p.howmany,))
    g0 = rng.standard_normal(p.n_str) * 0.1
    o.set_gradient(g0)
    r_o, kd_o, c_o = o.residual(u), o.residual(u, linear_kd=True), o.convolve(u)
    g = gpu.GpuSolver(p)
    g.set_gradient(g0)
    assert relerr(g.reference_stiffness(), o.reference_stiffness()) < 1e-15
    assert relerr(g.green(), o.green()) < 1e-12
    e = (relerr(g.residual(u), r_o), relerr(g.residual(u, linear_kd=True), kd_o), relerr(g.convolve(u), c_o))
    g.close()
    say(f"  fp64 residual {e[0]:.2e} matvec {e[1]:.2e} convolve {e[2]:.2e}")
    assert max(e) < 1e-12
    g = gpu.GpuSolver(p, precision="mixed")
    g.set_operator_fp32(True)
    e32 = (relerr(g.residual(u, linear_kd=True), kd_o), relerr(g.convolve(u), c_o))
    g.close()
    say(f"  fp32 matvec {e32[0]:.2e} convolve {e32[1]:.2e}")
    assert max(e32) < 1e-5


def test_headline_cg_iterations_256():
    """The headline solve (256^3 elastic, FP64) for a fixed number of CG
    iterations: error history, solution and homogenized stress against the
    oracle's run of the same iterations."""
    p = workloads.headline_problem()
    p.n_it, p.tol = 3, 0.0
    g0 = workloads.mandel_unit(6, 0)
    say("headline 256^3 fp64 CG, 3 iterations")
    g = gpu.GpuSolver(p)
    g.set_gradient(g0)
    it_g = g.solve()
    hg, ug, sg = g.error_history(), g.get_u(), g.homogenized_stress()
    g.close()
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    it_o = o.solve()
    ho = o.error_history()
    assert it_g == it_o == 3
    assert relerr(hg, ho) < 1e-12
    assert relerr(ug, o.get_u()) < 1e-11
    assert relerr(sg, o.homogenized_stress()) < 1e-12


def test_config1_exact():
    """Configuration 1: thermal 32^3, 12 spheres r in [3, 6], k = 1 / 10, load
    cases e_x, e_y, e_z, CG L-inf 1e-10, 500 iterations at most; iterations,
    homogenized flux, temperature and gradient fields against the oracle."""
    p = workloads.config1_problem()
    say(f"config 1 thermal 32^3, inclusion fraction {workloads.inclusion_fraction(p.ms):.4f}")
    g = gpu.GpuSolver(p)
    for lc in p.load_cases:
        g0 = lc.g0_path[0]
        o = oracle.Oracle(p)
        o.set_gradient(g0)
        it_o = o.solve()
        g.reset_state()
        g.set_gradient(g0)
        it_g = g.solve()
        assert abs(it_g - it_o) <= 2, (g0, it_g, it_o)
        assert g.error_history()[-1] <= p.tol
        assert relerr(g.homogenized_stress(), o.homogenized_stress()) < 1e-10
        assert relerr(g.get_u(), o.get_u()) < 1e-8
        fo, fg = o.element_fields(), g.element_fields()
        for k in ("strain", "stress", "strain_gp", "stress_gp"):
            assert relerr(fg[k], fo[k]) < 1e-8, k
    g.close()


def test_config2_full_128():
    """Configuration 2 at its full 128^3: FP64 CG against the oracle on e_11
    (iterations +-2, homogenized stress 1e-10); all six Mandel load cases on the
    GPU give a symmetric positive definite tangent; the mixed form reaches each
    FP64 column to 1e-8."""
    p = workloads.config2_problem()
    say(f"config 2 elastic 128^3, inclusion fraction {workloads.inclusion_fraction(p.ms):.4f}")
    g = gpu.GpuSolver(p)
    C, its = np.zeros((6, 6)), []
    for k, lc in enumerate(p.load_cases):
        g.reset_state()
        g.set_gradient(lc.g0_path[0])
        its.append(g.solve())
        assert g.error_history()[-1] <= p.tol
        C[:, k] = g.homogenized_stress()
    g.close()
    say(f"  fp64 iterations {its}")
    assert np.max(np.abs(C - C.T)) < 1e-6 * np.max(np.abs(C))
    assert np.min(np.linalg.eigvalsh(0.5 * (C + C.T))) > 0
    gm = gpu.GpuSolver(p, precision="mixed")
    for k, lc in enumerate(p.load_cases):
        gm.reset_state()
        gm.set_gradient(lc.g0_path[0])
        gm.solve()
        assert relerr(gm.homogenized_stress(), C[:, k]) < 1e-8, k
    gm.close()
    o = oracle.Oracle(p)
    o.set_gradient(p.load_cases[0].g0_path[0])
    it_o = o.solve()
    assert abs(its[0] - it_o) <= 2
    assert relerr(C[:, 0], o.homogenized_stress()) < 1e-10


def test_config3_j2_residual_256():
    """The J2 residual kernel at 256^3 (slot layout with the 18th stencil tile
    per direction) with yielding Gauss points, against the oracle."""
    p = workloads.config3_problem()
    say("config 3 J2 residual 256^3")
    rng = np.random.default_rng(9)
    u = rng.standard_normal(p.dims + (3,)) * 0.002
    g0 = np.array([0.01, -0.004, -0.004, 0.001, 0.0, -0.001])
    o = oracle.Oracle(p)
    o.set_gradient(g0)
    r_o = o.residual(u)
    del o
    g = gpu.GpuSolver(p)
    g.set_gradient(g0)
    e = relerr(g.residual(u), r_o)
    g.close()
    say(f"  residual {e:.2e}")
    assert e < 1e-12


def test_config3_load_steps_256():
    """Configuration 3 at 256^3, the first five load steps of the ramp (past the
    first yield): every step converges, the stress-controlled components stay
    at zero, the hardening variable is non-negative and grows, and the mixed
    form reaches the FP64 stress and macro strain to 1e-8."""
    p = workloads.config3_problem()
    mbc = workloads.config3_ramp()
    say("config 3 J2 256^3, 5 load steps fp64 + mixed")
    res = {}
    for prec in ("fp64", "mixed"):
        g = gpu.GpuSolver(p, precision=prec)
        out = []
        qmax_prev = 0.0
        for t in range(5):
            g.enable_mixed_bc(mbc, t)
            g.solve()
            assert g.error_history()[-1] <= p.tol, (prec, t)
            s = g.homogenized_stress()
            assert max(abs(s[1]), abs(s[2])) < 1e-6 * abs(s[0]), (prec, t, s)
            out.append((s.copy(), g.macro_gradient().copy()))
            if prec == "fp64":
                q = g.get_field("hardening")
                assert q.min() >= 0.0
                assert q.max() >= qmax_prev
                qmax_prev = float(q.max())
                del q
            g.extrapolate_displacement()
        g.close()
        res[prec] = out
        if prec == "fp64":
            say(f"  fp64 max hardening after 5 steps {qmax_prev:.3e}")
            assert qmax_prev > 0  # the matrix yielded
    for t, ((s64, m64), (smx, mmx)) in enumerate(zip(res["fp64"], res["mixed"])):
        assert relerr(smx, s64) < 1e-8, t
        assert relerr(mmx, m64) < 1e-8, t


def test_config5_every_group():
    """One RVE of every (edge, physics) group of the 384-RVE batch, at its own
    size (edges 32..96 including the radix-3/5 lengths 48, 80, 96).  FP64 GPU
    against the oracle with the tolerance tightened to 1e-10 (iterations +-2,
    homogenized response 1e-10); the mixed form at the batch's own 1e-8 to 1e-7
    of the oracle's converged response."""
    from dataclasses import replace
    probs = config5_rves()
    first = {}
    for i, q in enumerate(probs):
        first.setdefault((q.dims[0], q.howmany), i)
    for (n, h), i in sorted(first.items()):
        q = probs[i]
        q64 = replace(q, tol=1e-10)
        g0 = q.load_cases[0].g0_path[0]
        say(f"config 5 RVE {i}: {n}^3 h={h}")
        o = oracle.Oracle(q64)
        o.set_gradient(g0)
        it_o = o.solve()
        so = o.homogenized_stress()
        for prec, qq in (("fp64", q64), ("mixed", q)):
            g = gpu.GpuSolver(qq, precision=prec)
            g.set_gradient(g0)
            it_g = g.solve()
            sg = g.homogenized_stress()
            assert g.error_history()[-1] <= qq.tol
            g.close()
            if prec == "fp64":
                assert abs(it_g - it_o) <= 2, (i, it_g, it_o)
                assert relerr(sg, so) < 1e-10, i
            else:
                assert relerr(sg, so) < 1e-7, i
