"""Batched uneven RVEs (context reuse per size class, concurrent groups,
autotuned stencil depth): every RVE reaches the oracle's homogenized response
for its own problem."""
import numpy as np
import pytest

import oracle
from fans_amd.batch import config5_rves


def test_batch_generator_is_deterministic():
    a = config5_rves(count=6, seed=3)
    b = config5_rves(count=6, seed=3)
    for p, q in zip(a, b):
        assert p.dims == q.dims and p.howmany == q.howmany
        assert np.array_equal(p.ms, q.ms)
        assert p.groups[0].properties == q.groups[0].properties
    assert {p.howmany for p in a} == {1, 3}


@pytest.mark.gpu
def test_batch_matches_oracle():
    from fans_amd.batch import solve_batch
    probs = [p for p in config5_rves(count=12, seed=5) if p.dims[0] <= 48][:6]
    print(f"[gpu-case] batch of {len(probs)} RVEs", flush=True)
    ref = []
    for p in probs:
        o = oracle.Oracle(p)
        o.set_gradient(p.load_cases[0].g0_path[0])
        o.solve()
        ref.append(o.homogenized_stress())
    for fp in ("mixed", "fp64"):
        rep = solve_batch(probs, precision=fp, workers=3)
        assert len(rep.results) == len(probs)
        for r in rep.results:
            so = ref[r.index]
            assert r.converged
            # both forms stop on the FP64 residual at 1e-8; the homogenized
            # responses then agree far below the tolerance's own scale
            assert np.max(np.abs(r.stress - so)) <= 1e-7 * np.max(np.abs(so)), (fp, r.index)


def test_batch_groups_split_on_context_settings():
    """RVEs that differ in any setting fixed at context creation (here the RVE
    lengths L, the element type and the tolerance) never share a context."""
    from dataclasses import replace

    from fans_amd.gpu import context_key
    p = config5_rves(count=1, seed=3)[0]
    keys = {context_key(p), context_key(replace(p, L=(2.0, 1.0, 1.0))), context_key(replace(p, fe_type="HEX8R")),
            context_key(replace(p, tol=1e-6)), context_key(replace(p, ms=p.ms.copy()))}
    assert len(keys) == 4


@pytest.mark.gpu
def test_load_rejects_other_context_settings():
    from dataclasses import replace

    from fans_amd.gpu import GpuSolver
    p = [q for q in config5_rves(count=8, seed=5) if q.dims[0] <= 48][0]
    print(f"[gpu-case] load() with other settings, {p.dims}", flush=True)
    g = GpuSolver(p)
    try:
        g.load(replace(p, ms=p.ms.copy()))  # same settings: accepted
        with pytest.raises(ValueError, match="L"):
            g.load(replace(p, L=(2.0, 1.0, 1.0)))
    finally:
        g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [32, 48, 80])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_context_matches_single_solves(precision, n):
    """One batched context solving 5 RVEs at once gives every RVE the iteration
    count, homogenized response and displacement field of solving it alone,
    bit for bit (same kernels and summation order).  The RVEs finish at
    different iterations, so the launch grid shrinks on the way; one slot has
    a zero load and is finished before the first iteration.  Mixed form: the
    slots that continue after a residual replacement resume at a common
    iteration index, their last direction moved in the ring.  n = 32 runs the
    compile-time z and register y / x kernels, n = 48 and 80 (radix 3 and 5)
    the generic ones."""
    from fans_amd.gpu import FansGpuError, GpuBatch, GpuSolver
    # at 1e-10 the mixed form takes two or more residual replacements, after
    # which the slots resume together from different iteration counts
    probs = config5_rves(count=10, seed=11, edges=(n,), tol=1e-10)
    for phys in (0, 1):
        ps = probs[phys::2]
        g0s = [ps[0].load_cases[0].g0_path[0]] * len(ps)
        g0s[2] = np.zeros_like(g0s[2])
        print(f"[gpu-case] batched context {precision} n={n} howmany={ps[0].howmany} slots={len(ps)}", flush=True)
        single = []
        for p, g0 in zip(ps, g0s):
            g = GpuSolver(p, precision=precision)
            try:
                g.set_gradient(g0)
                g.solve()
                single.append((int(g.stats.inner_iterations), g.homogenized_stress(), g.get_u()))
            finally:
                g.close()
        bt = GpuBatch(ps, precision=precision)
        try:
            for rep in range(2):  # the second solve replays the captured graphs
                for b, p in enumerate(ps):
                    bt.load_slot(b, p, g0s[b])
                bt.solve()
                its, conv, _ = bt.results()
                st = bt.stresses()
                assert its[2] == 0 and not np.any(st[2])
                for b in range(len(ps)):
                    assert conv[b]
                    assert its[b] == single[b][0], (rep, b, its, [s[0] for s in single])
                    assert np.array_equal(st[b], single[b][1]), (rep, b)
                    bt.select(b)
                    assert np.array_equal(bt.homogenized_stress(), single[b][1])
                    assert np.array_equal(bt.get_u(), single[b][2])
            assert len(set(int(v) for v in its)) > 1  # the slots finished apart
            if precision == "mixed":
                assert bt.stats.outer_iterations >= 2
            with pytest.raises(FansGpuError, match="batched"):
                bt.set_voxel_moduli(np.ones(int(np.prod(ps[0].dims))))
        finally:
            bt.close()


@pytest.mark.gpu
def test_fft_tile_budget_keeps_results():
    """The FFT tile width of the generic passes (LDS budget per workgroup, the
    autotuner's knob) moves only the order of the partial sums: iterations and
    homogenized response stay put across the budgets."""
    from fans_amd.gpu import FansGpuError, GpuSolver
    p = config5_rves(count=2, seed=4, edges=(48,))[1]
    g0 = p.load_cases[0].g0_path[0]
    print(f"[gpu-case] FFT LDS budgets on {p.dims} howmany={p.howmany}", flush=True)
    g = GpuSolver(p, precision="fp64")
    try:
        res = {}
        for kib in (16, 24, 40, 60):
            g.set_fft_lds_kib(kib)
            g.reset_state()
            g.set_gradient(g0)
            g.solve()
            res[kib] = (g.stats.iterations, g.homogenized_stress())
        it0, s0 = res[40]
        for kib, (it, st) in res.items():
            assert abs(it - it0) <= 1, (kib, it, it0)
            assert np.max(np.abs(st - s0)) <= 1e-10 * np.max(np.abs(s0)), kib
        with pytest.raises(FansGpuError, match="LDS"):
            g.set_fft_lds_kib(200)
    finally:
        g.close()
