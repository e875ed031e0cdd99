"""CPU checks of the C ABI library (loads, exports every declared symbol, fails
loudly without a device) and of the reference JSON input parsing."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from fans_amd import gpu
from fans_amd.problem import Problem
from tests.conftest import GOLDEN, ROOT, gpu_available


def declared_symbols():
    with open(os.path.join(ROOT, "include", "fans_gpu.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"\b(fans_gpu_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(gpu.EXPORTS)
    assert lib.fans_gpu_abi_version() == 1


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_create_fails_loudly_without_device(sphere32_ms):
    p = Problem.from_json(os.path.join(GOLDEN, "test_LinearThermal.json"), ms=sphere32_ms)
    with pytest.raises(gpu.FansGpuError) as e:
        gpu.GpuSolver(p)
    assert e.value.code == -2  # FANS_GPU_ENODEV, no CPU fallback


def test_parse_reference_inputs(sphere32_ms):
    p = Problem.from_json(os.path.join(GOLDEN, "test_LinearThermal.json"), ms=sphere32_ms)
    assert p.howmany == 1 and p.fe_type == "HEX8R" and p.method == "cg"
    assert p.measure == "Linfinity" and p.err_type == "absolute" and p.tol == 1e-10 and p.n_it == 100
    k, n, ids, prm = p.flat_materials()
    assert list(k) == [1] and list(n) == [2] and list(ids) == [0, 1] and list(prm) == [1.0, 10.0]
    assert np.allclose(p.load_cases[0].g0_path, [[0.01, 0.02, -0.01]])
    assert "homogenized_tangent" in p.results
    e = Problem.from_json(os.path.join(GOLDEN, "test_LinearElastic.json"), ms=sphere32_ms)
    k, n, ids, prm = e.flat_materials()
    assert list(k) == [3] and list(prm) == [62.5, 28.8462, 222.222, 166.6667]
    m = Problem.from_json(os.path.join(GOLDEN, "test_MixedBCs.json"), ms=sphere32_ms)
    assert [lc.mixed for lc in m.load_cases] == [True, True, False]
    assert m.load_cases[0].mbc["strain_indices"] == [2, 3, 4, 5] and m.load_cases[0].n_steps == 12
    assert m.load_cases[1].mbc["strain_indices"] == [] and m.load_cases[1].n_steps == 2
    k, n, ids, prm = m.flat_materials()
    assert list(k) == [6, 3] and list(ids) == [0, 1]


@pytest.mark.parametrize("mutate,msg", [
    (lambda j: j.update(FE_type="QUAD4"), "FE_type"),
    (lambda j: j["materials"][0].update(phases=[0, 0]), "duplicate"),
    (lambda j: j["materials"][0].update(phases=[0, 2]), "not assigned"),
    (lambda j: j["materials"][0]["material_properties"].pop("conductivity"), "Missing"),
    (lambda j: j.update(macroscale_loading=[[[0.1, 0.2]]]), "loading vector"),
    (lambda j: j.update(linesearch_parameters={"max_iter": 0}), "linesearch"),
    (lambda j: j["error_parameters"].update(measure="Lfoo"), "measure"),
])
def test_input_errors_like_reference(sphere32_ms, mutate, msg):
    with open(os.path.join(GOLDEN, "test_LinearThermal.json")) as f:
        j = json.load(f)
    mutate(j)
    with pytest.raises(ValueError, match=msg):
        Problem.from_json(j, ms=sphere32_ms)


def test_mixed_bc_validation():
    from fans_amd.problem import mixed_bc_from_json
    with pytest.raises(ValueError, match="both"):
        mixed_bc_from_json({"strain_indices": [0, 1, 2], "stress_indices": [2, 3, 4, 5],
                            "strain": [[0, 0, 0]], "stress": [[0, 0, 0, 0]]}, 6)
    with pytest.raises(ValueError, match="either"):
        mixed_bc_from_json({"strain_indices": [0], "stress_indices": [1], "strain": [[0]], "stress": [[0]]}, 6)


def test_odd_nz_is_refused(sphere32_ms):
    """n_z must be even (the r2c half-length transform); the reference only
    warns, this build refuses before any context exists."""
    ms = np.ascontiguousarray(sphere32_ms[:, :, :31])
    with pytest.raises(ValueError, match="n_z must be even"):
        Problem.from_json(os.path.join(GOLDEN, "test_LinearThermal.json"), ms=ms)


@pytest.mark.gpu
def test_grid_factors_are_refused(sphere32_ms):
    """Grid edges with a prime factor other than 2, 3 and 5 are refused by
    fans_gpu_create (FANS_GPU_EUNSUP) with the limit in the message."""
    ms = np.ascontiguousarray(sphere32_ms[:28, :, :])  # 28 = 4 * 7
    p = Problem.from_json(os.path.join(GOLDEN, "test_LinearThermal.json"), ms=ms)
    with pytest.raises(gpu.FansGpuError, match="products of 2, 3 and 5"):
        gpu.GpuSolver(p)


def test_batch_slot_count_is_checked_before_the_device():
    """fans_gpu_create_batch refuses slot counts outside 1..4096 and methods
    other than CG with FANS_GPU_EINVAL, before it looks for a device (so also
    on a host without one)."""
    import ctypes as C
    L = gpu.lib()
    d = gpu.Desc()
    d.n_x = d.n_y = d.n_z = 16
    d.howmany = 1
    d.method = gpu.METHOD["cg"]
    d.n_it = 10
    d.L[:] = [1.0, 1.0, 1.0]
    d.tol = 1e-8
    ptr = C.c_void_p()
    for slots in (0, -3, 4097):
        assert L.fans_gpu_create_batch(C.byref(ptr), C.byref(d), slots) == -1
        assert not ptr.value
    d.method = gpu.METHOD["fp"]
    assert L.fans_gpu_create_batch(C.byref(ptr), C.byref(d), 4) == -1
