"""The drop-in CLI `fans_amd/FANS in.json out.h5` (C++ host over the C ABI).

CPU: usage / version / input-error exit codes and messages of src/main.cpp and
src/reader.cpp:63-183, and the missing-device error.  GPU: the reference's own
test inputs (test/input_files/test_LinearThermal.json, test_LinearElastic.json
on sphere32) run end to end; the HDF5 result layout (reader.h:330-351) and
values are checked against the oracle.
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from fans_amd import hdf5io
from fans_amd.problem import Problem
from tests.conftest import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fans_amd", "FANS")
DSET = "/sphere/32x32x32/ms"


def _need_cli():
    if not os.path.exists(CLI):
        pytest.skip("fans_amd/FANS is not built (python -m fans_amd.build)")


def _run(*args, env=None):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, env=env)


@pytest.fixture()
def case_dir(tmp_path, sphere32_ms):
    ms_file = tmp_path / "sphere32.h5"
    hdf5io.write_microstructure(str(ms_file), DSET, sphere32_ms)

    def make(name, **edits):
        j = json.load(open(os.path.join(GOLDEN, f"{name}.json")))
        j["microstructure"]["filepath"] = str(ms_file)
        j["microstructure"]["datasetname"] = DSET
        j.update(edits)
        p = tmp_path / f"{name}.json"
        json.dump(j, open(p, "w"))
        return str(p), j
    return make


def test_usage_and_version():
    _need_cli()
    r = _run()
    assert r.returncode == 10 and "USAGE" in r.stderr
    r = _run("--version")
    assert r.returncode == 0 and "C ABI 1" in r.stdout


@pytest.mark.parametrize("edit,msg", [
    ({"FE_type": "HEX20"}, "FE_type must be one of"),
    ({"strain_type": "medium"}, "strain_type must be either"),
    ({"linesearch_parameters": {"max_iter": 0}}, "linesearch_parameters"),
    ({"macroscale_loading": [[[1.0, 0.0]]]}, "Invalid length of loading vector"),
    ({"materials": []}, "'materials' must be non-empty array"),
    ({"reference_material": [[1.0 if i == j else 0.0 for j in range(6)] for i in range(5)]},
     "reference_material must be 6x6"),
    # indefinite (a negative pivot) and singular: both fail the Cholesky test
    ({"reference_material": [[(-1.0 if i == 3 else 2.0) if i == j else 0.0 for j in range(6)] for i in range(6)]},
     "reference_material must be symmetric positive definite"),
    ({"reference_material": [[1.0] * 6 for _ in range(6)]},
     "reference_material must be symmetric positive definite"),
])
def test_input_errors_exit_10(case_dir, edit, msg):
    _need_cli()
    path, _ = case_dir("test_LinearElastic", **edit)
    r = _run(path, path + ".h5")
    assert r.returncode == 10, r.stderr
    assert msg in r.stderr


def test_missing_property_exit_10(case_dir):
    _need_cli()
    path, j = case_dir("test_LinearElastic")
    del j["materials"][0]["material_properties"]["shear_modulus"]
    json.dump(j, open(path, "w"))
    r = _run(path, path + ".h5")
    assert r.returncode == 10 and "Missing material property 'shear_modulus'" in r.stderr


def test_no_device_message(case_dir):
    _need_cli()
    from tests.conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    path, _ = case_dir("test_LinearElastic")
    r = _run(path, path + ".h5", env={**os.environ, "HIP_VISIBLE_DEVICES": ""})
    assert r.returncode == 1 and "no HIP device" in r.stderr


def _res(out, j, field, load=0, step=0):
    prefix = j.get("results_prefix", "")
    return hdf5io.read_dataset(out, f"{DSET}_results/{prefix}/load{load}/time_step{step}/{field}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["test_LinearThermal", "test_LinearElastic"])
def test_cli_end_to_end(case_dir, sphere32_ms, name):
    _need_cli()
    path, j = case_dir(name)
    out = path + ".h5"
    print(f"[gpu-case] cli {name}", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = Problem.from_json(path, ms=sphere32_ms)
    o = oracle.Oracle(p)
    lc = p.load_cases[0]
    o.set_gradient(lc.g0_path[0])
    it = o.solve()
    n_str, h = p.n_str, p.howmany
    s_avg, perm = _res(out, j, "stress_average")
    assert perm is None and s_avg.shape == (n_str,)
    so = o.homogenized_stress()
    assert np.max(np.abs(s_avg - so)) <= 1e-10 * np.max(np.abs(so))
    e_avg, _ = _res(out, j, "strain_average")
    assert np.allclose(e_avg, lc.g0_path[0], atol=1e-12)
    err, _ = _res(out, j, "absolute_error")
    assert abs(len(err) - (it + 1)) <= 2 and err[-1] <= p.tol
    msd, perm = _res(out, j, "microstructure", )
    assert perm == "zyx" and msd.shape == (32, 32, 32, 1)
    assert np.array_equal(msd[..., 0].transpose(2, 1, 0), sphere32_ms)
    uf, perm = _res(out, j, "displacement_fluctuation")
    assert perm == "zyx" and uf.shape == (32, 32, 32, h)
    ut, _ = _res(out, j, "displacement")
    # displacement = fluctuation + g0 . X with X relative to the cell centre (solver.h:586-648)
    x = (np.arange(32) / 32.0 - 0.5)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    g = e_avg
    if h == 1:
        lin = g[0] * X + g[1] * Y + g[2] * Z
        assert np.max(np.abs(ut[..., 0] - uf[..., 0] - lin)) < 1e-12
    else:
        r2 = np.sqrt(0.5)
        G = np.array([[g[0], r2 * g[3], r2 * g[4]], [r2 * g[3], g[1], r2 * g[5]], [r2 * g[4], r2 * g[5], g[2]]])
        for c in range(3):
            lin = G[c, 0] * X + G[c, 1] * Y + G[c, 2] * Z
            assert np.max(np.abs(ut[..., c] - uf[..., c] - lin)) < 1e-12
    fo = o.element_fields()
    for fld in ("stress", "strain"):
        v, perm = _res(out, j, fld)
        assert perm == "zyx" and v.shape == (32, 32, 32, n_str)
        ref = fo[fld].reshape(32, 32, 32, n_str).transpose(2, 1, 0, 3)
        assert np.max(np.abs(v - ref)) <= 1e-8 * np.max(np.abs(ref)), fld
    C, _ = _res(out, j, "homogenized_tangent")
    Co = o.homogenized_tangent()
    assert C.shape == (n_str, n_str)
    assert np.max(np.abs(C - Co)) <= 1e-6 * np.max(np.abs(Co))


@pytest.mark.gpu
def test_cli_j2_steps(case_dir, sphere32_ms):
    """test_J2Plasticity.json with its model swapped for J2PlasticityNew (the
    reference input's J2ViscoPlastic_NonLinearIsotropicHardening is not built;
    same moduli and yield stress, no hardening) and three of its strain steps
    (0, 0.002, 0.004 in e11, the last two past yield): per-step P_bar and the
    model's plastic_strain / isotropic_hardening_variable slabs
    (J2PlasticityNew.h:155-208) against the oracle."""
    _need_cli()
    j0 = json.load(open(os.path.join(GOLDEN, "test_J2Plasticity.json")))
    steps = [j0["macroscale_loading"][0][k] for k in (0, 20, 40)]
    mats = [{"phases": [0], "matmodel": "J2PlasticityNew_LinearIsotropicHardening",
             "material_properties": {"bulk_modulus": [62.5], "shear_modulus": [28.8462], "yield_stress": [0.1],
                                     "isotropic_hardening_parameter": [0.0]}},
            j0["materials"][1]]
    results = ["stress_average", "strain_average", "plastic_strain", "isotropic_hardening_variable"]
    path, j = case_dir("test_J2Plasticity", macroscale_loading=[steps], materials=mats, results=results)
    out = path + ".h5"
    print("[gpu-case] cli j2 3 steps", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = Problem.from_json(path, ms=sphere32_ms)
    o = oracle.Oracle(p)
    for t, g in enumerate(steps):
        o.set_gradient(g)
        o.solve()
        s_avg, _ = _res(out, j, "stress_average", step=t)
        so = o.homogenized_stress()
        assert np.max(np.abs(s_avg - so)) <= 1e-8 * np.max(np.abs(so)), t
        ep, q = o.get_state()
        pe, perm = _res(out, j, "plastic_strain", step=t)
        assert perm == "zyx" and pe.shape == (32, 32, 32, 6)
        pq, _ = _res(out, j, "isotropic_hardening_variable", step=t)
        ref_e = ep.mean(axis=1).reshape(32, 32, 32, 6).transpose(2, 1, 0, 3)
        ref_q = q.mean(axis=1).reshape(32, 32, 32, 1).transpose(2, 1, 0, 3)
        if t > 0:
            assert q.max() > 0, t  # past yield
        assert np.max(np.abs(pe - ref_e)) <= 1e-6 * max(np.max(np.abs(ref_e)), 1e-12), t
        assert np.max(np.abs(pq - ref_q)) <= 1e-6 * max(np.max(np.abs(ref_q)), 1e-12), t
        o.extrapolate_displacement()


@pytest.mark.gpu
def test_cli_j2_visco_reference_input(case_dir, sphere32_ms):
    """The reference's test_J2Plasticity.json unchanged (J2ViscoPlastic with the
    saturation law, 151 strain steps, its own result list) through the CLI.  The
    oracle follows the CLI's per-step sequence (gradient, solve, postprocess,
    extrapolation) for the first 45 steps (past yield, at e11 = 0.0044); the
    stress average, the plastic strain and both hardening variables of those
    steps match it, and every later step is written and finite."""
    _need_cli()
    path, j = case_dir("test_J2Plasticity")
    out = path + ".h5"
    print("[gpu-case] cli test_J2Plasticity.json unchanged (151 steps)", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = Problem.from_json(path, ms=sphere32_ms)
    o = oracle.Oracle(p)
    steps = j["macroscale_loading"][0]
    for t in range(45):
        o.set_gradient(steps[t])
        o.solve()
        fo = o.element_fields()
        s_avg, _ = _res(out, j, "stress_average", step=t)
        assert np.max(np.abs(s_avg - fo["stress_average"])) <= 1e-8 * np.max(np.abs(fo["stress_average"])), t
        if t % 11 == 0 or t == 44:
            ep, q = o.get_state()
            kb = o.get_kinematic_state()
            for name, ref in (("plastic_strain", ep.mean(axis=1)), ("isotropic_hardening_variable", q.mean(axis=1)),
                              ("kinematic_hardening_variable", kb.mean(axis=1))):
                got, _ = _res(out, j, name, step=t)
                ref = ref.reshape(32, 32, 32, -1).transpose(2, 1, 0, 3)
                assert np.max(np.abs(got - ref)) <= 1e-6 * max(np.max(np.abs(ref)), 1e-12), (t, name)
        o.extrapolate_displacement()
    assert q.max() > 0
    last, _ = _res(out, j, "stress_average", step=len(steps) - 1)
    assert np.all(np.isfinite(last))


@pytest.mark.gpu
def test_cli_pseudo_plastic_input(case_dir, sphere32_ms):
    """test_PseudoPlastic.json (PseudoPlasticNonLinearHardening matrix, BBAR)
    cut to three of its 51 steps: per-step P_bar and the plastic_flag slab
    (float, PseudoPlastic.h:56-64) against the oracle."""
    _need_cli()
    j0 = json.load(open(os.path.join(GOLDEN, "test_PseudoPlastic.json")))
    steps = [j0["macroscale_loading"][0][k] for k in (0, 5, 20)]
    path, j = case_dir("test_PseudoPlastic", macroscale_loading=[steps],
                       results=["stress_average", "strain_average", "plastic_flag"])
    out = path + ".h5"
    print("[gpu-case] cli pseudo-plastic 3 steps", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = Problem.from_json(path, ms=sphere32_ms)
    assert p.fe_type == "BBAR"
    o = oracle.Oracle(p)
    for t, g in enumerate(steps):
        o.set_gradient(g)
        o.solve()
        s_avg, _ = _res(out, j, "stress_average", step=t)
        so = o.homogenized_stress()
        if t == 0:
            assert np.max(np.abs(s_avg)) == 0.0
        else:
            assert np.max(np.abs(s_avg - so)) <= 1e-8 * np.max(np.abs(so)), t
        fl, perm = _res(out, j, "plastic_flag", step=t)
        assert perm == "zyx" and fl.shape == (32, 32, 32, 1)
        ref = o.plastic_flag().reshape(32, 32, 32, 1).transpose(2, 1, 0, 3).astype(np.float32).astype(np.float64)
        assert np.array_equal(fl, ref), t
        if t == 2:
            assert ref.max() == 1.0  # plastic points in the matrix (one phase: flag 0 / 1)
        o.extrapolate_displacement()


@pytest.mark.gpu
def test_cli_mixed_bc_golden(case_dir, sphere32_ms):
    """test_MixedBCs.json (PseudoPlastic matrix, HEX8) load cases 2 and 3 through the
    CLI: the stress-controlled hydrostatic steps reach the macro strains the
    reference stored as load case 3 (1e-8 relative) at the prescribed stress,
    and load case 3 driven by those strains gives that stress back."""
    _need_cli()
    j0 = json.load(open(os.path.join(GOLDEN, "test_MixedBCs.json")))
    ml = j0["macroscale_loading"]
    path, j = case_dir("test_MixedBCs", macroscale_loading=[ml[1], ml[2]],
                       results=["stress_average", "strain_average", "absolute_error"])
    out = path + ".h5"
    print("[gpu-case] cli mixed bc golden", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    golden = np.array(ml[2])
    for t in range(len(golden)):
        target = np.array(ml[1]["stress"][t])
        e_avg, _ = _res(out, j, "strain_average", load=0, step=t)
        s_avg, _ = _res(out, j, "stress_average", load=0, step=t)
        assert np.max(np.abs(e_avg[:3] - golden[t][:3])) < 1e-8 * np.abs(golden[t][:3]).max(), (t, e_avg)
        assert np.max(np.abs(e_avg[3:])) < 1e-12
        assert np.max(np.abs(s_avg - target)) < 1e-9, (t, s_avg)
        e1, _ = _res(out, j, "strain_average", load=1, step=t)
        s1, _ = _res(out, j, "stress_average", load=1, step=t)
        assert np.allclose(e1, golden[t], rtol=0, atol=1e-15)
        assert np.max(np.abs(s1 - target)) < 1e-7, (t, s1)


@pytest.mark.gpu
def test_cli_large_strain(case_dir, sphere32_ms):
    """test_CompressibleNeoHookean.json (finite strain, n_str = 9) cut to its
    first two shear steps: F_bar and P_bar per step against the oracle, the
    total displacement u + (F - I) X of solver.h:623-638 and the 9-component
    slabs."""
    _need_cli()
    j0 = json.load(open(os.path.join(GOLDEN, "test_CompressibleNeoHookean.json")))
    steps = j0["macroscale_loading"][0][:2]
    path, j = case_dir("test_CompressibleNeoHookean", macroscale_loading=[steps])
    out = path + ".h5"
    print("[gpu-case] cli large strain", flush=True)
    r = _run(path, out)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    p = Problem.from_json(path, ms=sphere32_ms)
    o = oracle.Oracle(p)
    for t, F in enumerate(steps):
        o.set_gradient(F)
        o.solve()
        P_avg, _ = _res(out, j, "stress_average", step=t)
        F_avg, _ = _res(out, j, "strain_average", step=t)
        assert P_avg.shape == (9,) and np.allclose(F_avg, F, atol=1e-12)
        Po = o.homogenized_stress()
        assert np.max(np.abs(P_avg - Po)) <= 1e-9 * np.max(np.abs(Po)), t
        if t == 0:
            uf, perm = _res(out, j, "displacement_fluctuation", step=t)
            ut, _ = _res(out, j, "displacement", step=t)
            x = (np.arange(32) / 32.0 - 0.5)
            Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
            Fm = np.asarray(F).reshape(3, 3) - np.eye(3)
            for c in range(3):
                lin = Fm[c, 0] * X + Fm[c, 1] * Y + Fm[c, 2] * Z
                assert np.max(np.abs(ut[..., c] - uf[..., c] - lin)) < 1e-12
            s, perm = _res(out, j, "stress", step=t)
            assert perm == "zyx" and s.shape == (32, 32, 32, 9)
        o.extrapolate_displacement()
