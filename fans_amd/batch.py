"""Batched solves of many independent RVEs of uneven sizes.

Each RVE is one reference run (main.cpp:10-41 for one load case): its own
microstructure, phases and Green operator.  RVEs are grouped by their context
settings (grid, physics, element, solver settings).  Each group, in chunks of
at most `max_slots` RVEs, is one batched GPU context (fans_gpu_create_batch):
every RVE has its own buffers in it, and one solve runs all of them at once,
every kernel launch covering each RVE still iterating (the captured CG
iterations replay for the whole group; finished RVEs leave the launch grid).
Groups run concurrently, one host thread and one HIP stream each (the C ABI
calls release the interpreter lock).  Before the timed batch, a small
autotuner times the FFT tile width of the generic passes (their LDS budget per
workgroup) and the stencil tile depth (x planes per assembly workgroup) on
each group's context and keeps the fastest.  With batched=False each group
is one single-RVE context that loads and solves its RVEs in turn.

The synthetic batch of the differentiation-free throughput benchmark (384 RVEs,
edges {32, 48, 64, 80, 96}, half thermal, half elastic, 2-3 phases with random
moduli, random-sphere microstructures with volume fraction in [0.1, 0.4]) is
`config5_rves()`.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .gpu import GpuBatch, GpuSolver, context_key
from .microstructure import random_spheres
from .problem import LoadCase, MaterialGroup, Problem, isotropic_KG

EDGES = (32, 48, 64, 80, 96)


@dataclass
class RVEResult:
    index: int
    dims: tuple
    howmany: int
    iterations: int
    stress: np.ndarray
    solve_ms: float
    converged: bool


@dataclass
class BatchReport:
    results: list = field(default_factory=list)
    wall_s: float = 0.0
    tuned: dict = field(default_factory=dict)
    # per context: host wall ms of loading its RVEs, of the solve (and the
    # library's device time of it) and of reading the results
    units: list = field(default_factory=list)

    @property
    def rves_per_s(self):
        return len(self.results) / self.wall_s if self.wall_s > 0 else 0.0

    @property
    def voxel_iters(self):
        return sum(int(np.prod(r.dims)) * r.iterations for r in self.results)

    @property
    def voxel_iters_per_s(self):
        return self.voxel_iters / self.wall_s if self.wall_s > 0 else 0.0


def config5_rves(count=384, seed=0, tol=1e-8, edges=EDGES):
    """Deterministic synthetic batch: edge uniform from `edges`, physics
    alternating (even thermal, odd elastic), 2-3 phases, moduli E in [1, 100],
    nu in [0.2, 0.4], conductivity in [1, 50], sphere radii U[n/16, n/8] with
    the count set for a target volume fraction U[0.1, 0.4] (non-overlapping
    placement caps what is reached)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(edges))
        thermal = i % 2 == 0
        nph = int(rng.integers(2, 4))
        vf = float(rng.uniform(0.1, 0.4))
        rmin, rmax = n / 16.0, n / 8.0
        mean_vol = 4.0 / 3.0 * np.pi * (rmax ** 4 - rmin ** 4) / (4.0 * (rmax - rmin))
        k = max(1, int(round(vf * n ** 3 / mean_vol)))
        ms = random_spheres(n, k, rmin, rmax, seed=int(rng.integers(1 << 30)), n_inclusion_phases=nph - 1,
                            max_tries=20000)
        if thermal:
            cond = list(rng.uniform(1.0, 50.0, nph))
            groups = [MaterialGroup("LinearThermalIsotropic", list(range(nph)), {"conductivity": cond})]
            g0 = np.array([1.0, 0.0, 0.0])
        else:
            KG = [isotropic_KG(float(rng.uniform(1.0, 100.0)), float(rng.uniform(0.2, 0.4))) for _ in range(nph)]
            groups = [MaterialGroup("LinearElasticIsotropic", list(range(nph)),
                                    {"bulk_modulus": [a for a, _ in KG], "shear_modulus": [b for _, b in KG]})]
            g0 = np.array([1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
        p = Problem(dims=(n, n, n), howmany=1 if thermal else 3, ms=ms, groups=groups, fe_type="HEX8",
                    method="cg", measure="Linfinity", err_type="absolute", tol=tol, n_it=2000,
                    load_cases=[LoadCase(mixed=False, g0_path=g0[None, :], n_steps=1)])
        out.append(p)
    return out


# generic FFT passes: lengths other than the powers of two 16..256
def _generic_fft(dims):
    return any(n not in (16, 32, 64, 128, 256) for n in (dims[0], dims[1], dims[2] // 2))


def _autotune(solver: GpuSolver, fresh, planes=(4, 8, 16), fft_kib=(16, 24, 40, 60), repeats=1):
    """Tile settings with the smallest solve time (the library's device time
    of the solve, after `fresh()` has set up a new one): the FFT LDS budget of
    the generic passes (when the grid has such lengths), then the assembly tile
    depth.  Returns (planes, fft_kib)."""
    def timed():
        ms = []
        for _ in range(repeats):
            fresh()
            solver.solve()
            ms.append(solver.stats.solve_ms)
        return min(ms)

    def pick(values, setter):
        best, best_ms = None, float("inf")
        for v in values:
            setter(v)
            t = timed()
            if t < best_ms:
                best, best_ms = v, t
        setter(best)
        return best

    kib = pick(fft_kib, solver.set_fft_lds_kib) if _generic_fft(solver.dims) else 40
    cx = pick([c for c in planes if c <= solver.dims[0]], solver.set_asm_planes)
    return cx, kib


def solve_batch(problems, precision="mixed", workers=None, autotune=True, batched=True, max_slots=64):
    """Solve every problem's first load-case step; returns a BatchReport.
    workers: concurrent groups (default: all groups, at most 16).  batched:
    one batched context per group chunk of at most max_slots RVEs (else one
    single-RVE context per group, RVEs in turn)."""
    groups = {}
    for i, p in enumerate(problems):
        groups.setdefault(context_key(p), []).append(i)
    units = []
    for key, idx in groups.items():
        step = max_slots if batched else len(idx)
        units += [(key, tuple(idx[k:k + step])) for k in range(0, len(idx), step)]
    report = BatchReport()

    # setup, outside the timed region like the reference's plan creation:
    # the contexts and their autotuned tiles
    solvers = {}

    def setup(unit):
        key, idx = unit
        first = problems[idx[0]]
        g0 = [problems[i].load_cases[0].g0_path[0] for i in idx]
        if batched:
            s = GpuBatch([problems[i] for i in idx], precision=precision)

            def fresh():
                for b, i in enumerate(idx):
                    s.load_slot(b, problems[i], g0[b])
        else:
            s = GpuSolver(first, precision=precision)

            def fresh():
                s.reset_state()
                s.set_gradient(g0[0])
        if autotune:
            # tuned on the context itself (a batched one: its whole chunk)
            if key not in report.tuned:
                report.tuned[key] = _autotune(s, fresh)
            cx, kib = report.tuned[key]
            s.set_asm_planes(cx)
            s.set_fft_lds_kib(kib)
        solvers[unit] = s

    loaded = {}

    def load_unit(unit):
        s, (_, idx) = solvers[unit], unit
        t0 = time.perf_counter()
        for b, i in enumerate(idx):
            s.load_slot(b, problems[i], problems[i].load_cases[0].g0_path[0])
        loaded[unit] = (t0, time.perf_counter())

    def run_unit(unit):
        s, (_, idx) = solvers[unit], unit
        res = []
        if batched:
            t1 = time.perf_counter()
            s.solve()
            t2 = time.perf_counter()
            its, conv, _ = s.results()
            stress = s.stresses()
            t3 = time.perf_counter()
            p0 = problems[idx[0]]
            l0, l1 = loaded[unit]
            report.units.append({"group": f"{p0.dims[0]}^3 h={p0.howmany}", "rves": len(idx),
                                 "load_ms": round(1e3 * (l1 - l0), 2), "solve_ms": round(1e3 * (t2 - t1), 2),
                                 "device_solve_ms": round(float(s.stats.solve_ms), 2),
                                 "results_ms": round(1e3 * (t3 - t2), 2), "max_iterations": int(max(its))})
            for b, i in enumerate(idx):
                p = problems[i]
                res.append(RVEResult(i, tuple(p.dims), p.howmany, int(its[b]), stress[b].copy(),
                                     float(s.stats.solve_ms), bool(conv[b])))
            return res
        for i in idx:
            p = problems[i]
            s.load(p)
            s.set_gradient(p.load_cases[0].g0_path[0])
            s.solve()
            res.append(RVEResult(i, tuple(p.dims), p.howmany, int(s.stats.inner_iterations),
                                 s.homogenized_stress().copy(), float(s.stats.solve_ms), bool(s.stats.converged)))
        return res

    # largest units first so the tail is short
    order = sorted(units, key=lambda u: -np.prod(u[0][0]) * len(u[1]))
    try:
        for unit in order:
            setup(unit)
        t0 = time.perf_counter()
        nw = workers if workers else min(len(order), 16)
        # batched: every context loads its RVEs (host-bound table and Green
        # operator setup), then all solve; loads that ran beside other
        # contexts' solves waited on the busy GPU at every synchronisation.
        # One worker runs in the calling thread.
        if nw <= 1:
            if batched:
                list(map(load_unit, order))
            for res in map(run_unit, order):
                report.results.extend(res)
        else:
            with ThreadPoolExecutor(max_workers=nw) as ex:
                if batched:
                    list(ex.map(load_unit, order))
                for res in ex.map(run_unit, order):
                    report.results.extend(res)
        report.wall_s = time.perf_counter() - t0
    finally:
        for s in solvers.values():
            s.close()
    report.results.sort(key=lambda r: r.index)
    return report
