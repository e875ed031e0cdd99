#!/usr/bin/env python3
"""Headline benchmark: voxel-iterations/s of the FANS CG solve, 256^3 linear
elastic, on one MI355X; FP64 (drop-in) form as `value`, the mixed FP32/FP64
form reported beside it.

A step is one full solve (Solver::solve, include/solver.h:282-300) of the
256^3 problem from a fresh solver state.  Inputs (microstructure, material
tables, Green operator) are resident in HBM before the timed region starts.

    python bench.py [--steps K] [--warmup W] [--size 256] [--no-cpu-baseline] [--no-mixed]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel-iters/s at 256^3 elastic, mixed vs FP64 GPU @1e-8 err; % of HBM peak"
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s

# Algorithmic (compulsory) bytes per voxel of each kernel family, FP64, h = 3:
# the kernel's own inputs and outputs read/written once.  The spectrum holds
# (nz/2+1)/nz complex values per real value.
def family_bytes_per_voxel(nz, h=3, wb=8, green_quarter=True):
    """Per cubic grid of edge nz.  green_quarter: the Green operator is stored
    for kx <= n/2, ky <= n/2 only (mirror-symmetric reference element, as for
    the isotropic headline materials)."""
    V = wb * h
    S = 2 * wb * h * (nz // 2 + 1) / nz  # spectrum bytes per voxel
    Gp = wb * h * (h + 1) // 2 * (nz // 2 + 1) / nz  # packed Green operator, full storage
    if green_quarter:
        Gp *= (nz // 2 + 1) ** 2 / nz ** 2
    return {
        "fft_z_r2c": V + S,
        "fft_y_fwd": 2 * S,
        "fft_x_green": 2 * S + Gp,        # 54.6 B at 256^3 with quarter storage
        "fft_y_inv": 2 * S,
        # CG inverse z pass: read the spectrum, d_prev and K d_prev (for
        # <d, K d_prev>); write the new direction d = s + beta d_prev
        "fft_z_c2r": S + 3 * V,
        # CG solution update once per eight iterations: read eight directions
        # and u, write u
        "cg_update": 10 * V,
        # matvec: read d and the phase; write rnew = K d
        "assemble_lin": 2 * V + 1,
        # residual update fused into the forward z pass: read r and rnew;
        # write r and the spectrum
        "cg_update_fft_z_r2c": 3 * V + S,
    }


B_ALG_FP64 = 409.0  # compulsory bytes per voxel-iteration, FP64 h=3 (CG, single-pass FFT ideal)
B_ALG_MIXED = 205.0


def headline_problem(n=256, seed=0, tol=1e-10):
    from fans_amd.microstructure import random_spheres
    from fans_amd.problem import MaterialGroup, Problem, isotropic_KG
    s = n / 256.0
    ms = random_spheres(n, max(1, int(round(150 * s ** 3))), 8 * s, 20 * s, seed=seed)
    Km, Gm = isotropic_KG(3.0, 0.35)
    Ki, Gi = isotropic_KG(70.0, 0.22)
    groups = [MaterialGroup("LinearElasticIsotropic", [0, 1], {"bulk_modulus": [Km, Ki], "shear_modulus": [Gm, Gi]})]
    return Problem(dims=(n, n, n), howmany=3, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=2000)


def run_gpu(p, precision, steps, warmup, g0, timing_solves=2, all_load_cases=False):
    """`steps` timed solves (wall clock, synchronous library calls), then
    `timing_solves` identical solves with every launch bracketed by HIP events
    on the library's stream for the per-kernel durations (kept out of the timed
    region because the events add ~8% host overhead)."""
    from fans_amd.gpu import GpuSolver
    g = GpuSolver(p, precision=precision)
    g.set_gradient(g0)
    for _ in range(warmup):
        g.reset_state()
        g.set_gradient(g0)
        g.solve()
    iters, solve_ms, step_s = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        g.reset_state()
        g.set_gradient(g0)
        g.solve()  # synchronous: returns after the device finished
        step_s.append(time.perf_counter() - ts)
        iters.append(int(g.stats.inner_iterations))
        solve_ms.append(float(g.stats.solve_ms))
    t1 = time.perf_counter()
    g.set_timing(True)
    fam = {}
    fam_iters = 0
    for _ in range(timing_solves):
        g.reset_state()
        g.set_gradient(g0)
        g.solve()
        fam_iters += int(g.stats.inner_iterations)
        for k, (ms, n) in g.kernel_times().items():
            a, b = fam.get(k, (0.0, 0))
            fam[k] = (a + ms, b + n)
    g.set_timing(False)
    others = []
    for k in range(len(g0)) if all_load_cases else []:
        # the six Mandel unit strains, one solve each (outside the timed region)
        e = np.zeros(len(g0))
        e[k] = 1.0
        g.reset_state()
        g.set_gradient(e)
        g.solve()
        others.append((int(g.stats.inner_iterations), round(float(g.stats.solve_ms), 2)))
    g.reset_state()
    g.set_gradient(g0)
    g.solve()
    res = dict(iters=iters, solve_ms=solve_ms, step_s=step_s, wall_s=t1 - t0, fam=fam, fam_iters=fam_iters,
               load_cases=others,
               stress=g.homogenized_stress(),
               final_error=float(g.stats.final_error), outer=int(g.stats.outer_iterations),
               converged=bool(g.stats.converged))
    g.close()
    return res


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def probe_reference_fans():
    """Whether reference FANS (or the libraries it needs: FFTW3-MPI, parallel
    HDF5) exists on this host; the preferred CPU baseline runs it with mpirun."""
    import ctypes.util
    import shutil
    found = {"FANS": shutil.which("FANS"), "mpirun": shutil.which("mpirun") or shutil.which("mpiexec"),
             "libfftw3_mpi": ctypes.util.find_library("fftw3_mpi"), "libfftw3": ctypes.util.find_library("fftw3"),
             "libhdf5": ctypes.util.find_library("hdf5")}
    return {k: (v or None) for k, v in found.items()}


def cpu_baseline(p, g0, iters=2, reps=5):
    """CPU baseline beside the GPU run.  Reference FANS (mpirun over FFTW-MPI)
    is preferred; where this host lacks it (probe_reference_fans) the CPU
    restatement (oracle, OpenMP) runs the same workload: `reps` timed solves of
    `iters` CG iterations each on P = min(host CPUs available to this process,
    OMP_NUM_THREADS, n_x / 4) threads -- the reference's own rank limit of four
    x planes per rank -- and the median is reported."""
    import oracle
    probe = probe_reference_fans()
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS") or avail)
    cores = max(1, min(avail, omp, p.dims[0] // 4))
    oracle.set_threads(cores)
    from dataclasses import replace
    q = replace(p, n_it=iters, tol=0.0)
    o = oracle.Oracle(q)
    n = int(np.prod(p.dims))
    zero = np.zeros(tuple(p.dims) + (p.howmany,))
    rates, secs = [], []
    for _ in range(reps):
        o.set_u(zero)
        o.set_gradient(g0)
        t0 = time.perf_counter()
        it = o.solve()
        dt = time.perf_counter() - t0
        rates.append(n * it / dt)
        secs.append(dt)
    return {"value": float(np.median(rates)), "unit": "voxel-iters/s", "cores": cores, "kind": "port",
            "sample": f"CPU restatement (oracle, not reference FANS) of the {p.dims[0]}^3 linear elastic CG, same "
                      f"microstructure and load, {iters} iterations per solve, median of {reps} solves "
                      f"({min(secs):.1f}-{max(secs):.1f} s each)",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "cpus_available": avail, "threads": cores,
            "binding": os.environ.get("OMP_PROC_BIND", "none (OS scheduling)") + (
                f", places {os.environ['OMP_PLACES']}" if os.environ.get("OMP_PLACES") else ""),
            "reference_fans_probe": probe}


def traffic_from_profiles(family):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(family)


def run_batch_workload(args):
    """384 independent RVEs (edges 32..96, half thermal / half elastic, 2-3
    phases), mixed form to 1e-8; one JSON line with RVEs/s and voxel-iters/s."""
    from fans_amd.batch import config5_rves, solve_batch
    print("[bench] generating the 384-RVE batch", file=sys.stderr, flush=True)
    problems = config5_rves()
    workers = None
    if args.batch_group:
        # one (edge, howmany) group alone, in the calling thread
        e, h = (int(v) for v in args.batch_group.split(","))
        problems = [p for p in problems if p.dims[0] == e and p.howmany == h]
        workers = 1
    print("[bench] batch: setup + autotune, then the timed solves", file=sys.stderr, flush=True)
    rep = solve_batch(problems, precision="mixed", workers=workers)
    its = [r.iterations for r in rep.results]
    scaling = None
    if args.batch_workers:
        # the same batch with fewer concurrent groups: wall time against the
        # sum of the library's per-solve times (its share of the wall time
        # shows what the concurrency leaves idle or contends for)
        scaling = []
        for w in [int(v) for v in args.batch_workers.split(",")]:
            print(f"[bench] batch with {w} concurrent groups", file=sys.stderr, flush=True)
            rw = solve_batch(problems, precision="mixed", workers=w)
            by = {}
            for r in rw.results:
                by.setdefault(f"{r.dims[0]}^3 h={r.howmany}", []).append(r.solve_ms)
            sm = sum(r.solve_ms for r in rw.results)
            scaling.append({"workers": w, "wall_ms": round(1000 * rw.wall_s, 1), "rves_per_s": round(rw.rves_per_s, 1),
                            "sum_solve_ms": round(sm, 1),
                            "solve_ms_per_rve": {k: round(float(np.mean(v)), 3) for k, v in sorted(by.items())}})
    out = {"metric": "RVEs/s (384 uneven RVEs, mixed form, tol 1e-8)" if not args.batch_group else
           f"RVEs/s ({len(problems)} RVEs of the 384-RVE batch: edge,howmany {args.batch_group})",
           "value": round(rep.rves_per_s, 2),
           "unit": "RVEs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round(1000 * rep.wall_s, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "384 RVEs, edges {32,48,64,80,96}^3, half LinearThermalIsotropic / half "
                                  "LinearElasticIsotropic, 2-3 phases, random spheres, CG L-inf 1e-8, mixed form",
                      "rves": len(rep.results), "voxels": int(sum(np.prod(p.dims) for p in problems))},
           "voxel_iters_per_s": round(rep.voxel_iters_per_s, 1),
           "iterations": {"min": int(min(its)), "mean": round(float(np.mean(its)), 1), "max": int(max(its))},
           "converged": int(sum(r.converged for r in rep.results)),
           "autotuned_tiles": {f"{k[0][0]}^3 h={k[1]}": {"asm_planes": v[0], "fft_lds_kib": v[1]}
                               for k, v in rep.tuned.items()}}
    out["units"] = sorted(rep.units, key=lambda u: u["group"])
    if args.batch_single:
        # the same batch with one single-RVE context per group (RVEs in turn)
        print("[bench] batch with single-RVE contexts", file=sys.stderr, flush=True)
        rs = solve_batch(problems, precision="mixed", batched=False)
        out["single_rve_contexts"] = {"rves_per_s": round(rs.rves_per_s, 2), "wall_ms": round(1000 * rs.wall_s, 1)}
    if scaling:
        out["worker_scaling"] = scaling
    print(json.dumps(out), flush=True)


def j2_problem(n=256, seed=0, tol=1e-8):
    """J2 plasticity with linear isotropic hardening (matrix E=70, nu=0.33,
    yield 0.25, hardening 1.0) around elastic spheres (E=400, nu=0.2), the
    headline sphere law (150 spheres, r ~ U[8,20] at 256^3, seed 0), HEX8,
    CG L-inf 1e-8 per step."""
    from fans_amd.microstructure import random_spheres
    from fans_amd.problem import MaterialGroup, Problem, isotropic_KG
    s = n / 256.0
    ms = random_spheres(n, max(1, int(round(150 * s ** 3))), 8 * s, 20 * s, seed=seed)
    Km, Gm = isotropic_KG(70.0, 0.33)
    Ki, Gi = isotropic_KG(400.0, 0.2)
    groups = [MaterialGroup("J2PlasticityNew_LinearIsotropicHardening", [0],
                            {"bulk_modulus": [Km], "shear_modulus": [Gm], "yield_stress": [0.25],
                             "isotropic_hardening_parameter": [1.0]}),
              MaterialGroup("LinearElasticIsotropic", [1], {"bulk_modulus": [Ki], "shear_modulus": [Gi]})]
    return Problem(dims=(n, n, n), howmany=3, ms=ms, groups=groups, fe_type="HEX8", method="cg",
                   measure="Linfinity", err_type="absolute", tol=tol, n_it=2000)


def run_j2_workload(args):
    """Uniaxial strain ramp to 2 % in `--j2-steps` of 20 load steps with zero
    lateral macroscopic stress (mixed BCs), J2 matrix; FP64 and mixed form
    (FP32 Green preconditioner, FP64 residuals and return map).  One JSON line:
    voxel-iterations/s of the nonlinear CG (every line-search residual counts
    as part of its iteration) and the FP64-vs-mixed stress difference."""
    from fans_amd.gpu import GpuSolver
    n = args.size
    p = j2_problem(n)
    N = n ** 3
    nsteps = 20
    emax = 0.02
    strain = np.linspace(emax / nsteps, emax, nsteps)
    mbc = dict(strain_indices=[0, 3, 4, 5], stress_indices=[1, 2],
               strain_path=np.stack([strain, 0 * strain, 0 * strain, 0 * strain], 1),
               stress_path=np.zeros((nsteps, 2)), n_steps=nsteps)
    res = {}
    # the third run repeats the FP64 ramp with every launch bracketed by HIP
    # events (per-kernel durations; kept out of the timed runs)
    for prec, timing in (("fp64", False), ("mixed", False), ("fp64", True)):
        print(f"[bench] J2 {n}^3 {prec}{' (kernel timing)' if timing else ''}: {args.j2_steps} load steps",
              file=sys.stderr, flush=True)
        g = GpuSolver(p, precision=prec)
        g.set_timing(timing)
        its, ms, fam = [], [], {}
        for t in range(args.j2_steps):
            g.enable_mixed_bc(mbc, t)
            g.solve()
            its.append(int(g.stats.iterations))
            ms.append(float(g.stats.solve_ms))
            print(f"[bench]   step {t}: {its[-1]} iterations, {ms[-1]:.1f} ms", file=sys.stderr, flush=True)
            if timing:
                for k, (kms, cnt) in g.kernel_times().items():
                    a, b = fam.get(k, (0.0, 0))
                    fam[k] = (a + kms, b + cnt)
            g.extrapolate_displacement()
        _, q = g.get_state()
        res["timing" if timing else prec] = dict(its=its, ms=ms, stress=g.homogenized_stress().copy(),
                                                 qmax=float(q.max()), fam=fam)
        g.close()
    r64, rmx, rt = res["fp64"], res["mixed"], res["timing"]
    # J2 residual assembly: reads u (24 B) and the phase byte, writes r (24 B);
    # J2 voxels also read the committed state, 7 doubles at each of 8 Gauss
    # points (448 B); elastic voxels carry no state
    f_j2 = float(np.mean(p.ms == 0))
    res_bpv = 49 + 448 * f_j2
    ms_res, n_res = rt["fam"].get("assemble_res", (0.0, 0))
    avg_res = ms_res / max(n_res, 1)
    ach = res_bpv * N / (avg_res * 1e-3) / 1e9 if n_res else None
    rel = float(np.max(np.abs(rmx["stress"] - r64["stress"])) / np.max(np.abs(r64["stress"])))
    t64 = sum(r64["ms"]) * 1e-3
    out = {"metric": f"voxel-iters/s, J2 {n}^3 nonlinear CG (uniaxial ramp, mixed BCs)",
           "value": round(N * sum(r64["its"]) / t64, 1), "unit": "voxel-iters/s", "n_gpus": 1,
           "steps": args.j2_steps, "warmup": 0, "ms_per_step": round(1000 * t64 / args.j2_steps, 2),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{n}^3 J2PlasticityNew linear hardening matrix E=70 nu=0.33 sy=0.25 H=1.0, "
                                  "150 elastic spheres E=400 nu=0.2 r~U[8,20] seed 0, eps11 ramp to 0.02 in 20 "
                                  "steps with zero lateral stress, CG L-inf 1e-8", "load_steps": args.j2_steps},
           "fp64": {"iterations": r64["its"], "step_ms": [round(v, 1) for v in r64["ms"]]},
           "mixed": {"iterations": rmx["its"], "step_ms": [round(v, 1) for v in rmx["ms"]],
                     "time_to_solution_speedup": round(sum(r64["ms"]) / sum(rmx["ms"]), 3),
                     "homog_stress_rel_err": rel},
           "max_hardening_q": r64["qmax"],
           "roofline": {"bound": "hbm", "kernel": "assemble_res", "achieved": round(ach, 1) if ach else None,
                        "peak": 8000.0, "unit": "GB/s", "frac": round(ach / 8000.0, 4) if ach else None,
                        "traffic": traffic_from_profiles("assemble_res_j2"),
                        "alg_bytes_per_launch": res_bpv * N, "avg_launch_ms": round(avg_res, 4),
                        "j2_voxel_fraction": round(f_j2, 4)},
           "kernel_ms_per_iteration": {k: round(v[0] / max(sum(rt["its"]), 1), 4)
                                       for k, v in rt["fam"].items() if v[1]},
           "launches_per_iteration": {k: round(v[1] / max(sum(rt["its"]), 1), 2)
                                      for k, v in rt["fam"].items() if v[1]}}
    print(json.dumps(out), flush=True)


def _rel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _fields(g):
    f = g.element_fields(gp=False)
    return {"u": g.get_u(), "strain": f["strain"], "stress": f["stress"]}


def _field_errs(fa, fb):
    return {k: _rel(fa[k], fb[k]) for k in fb}


def run_accuracy_workload(args):
    """Per-configuration accuracy of the GPU forms: homogenized stress / flux
    and the fields (u, element strain, element stress) of the mixed form
    against the FP64 form, and of the FP64 form against the oracle wherever
    the oracle finishes in seconds.  One JSON line."""
    import oracle
    from dataclasses import replace

    from fans_amd import workloads
    from fans_amd.batch import config5_rves
    from fans_amd.gpu import GpuSolver
    oracle.set_threads(max(1, min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS") or 16))))
    out = {"metric": "per-configuration accuracy, mixed vs FP64 GPU and FP64 GPU vs oracle", "n_gpus": 1,
           "data": "synthetic", "configs": {}}

    def linear_config(name, p, oracle_cases):
        print(f"[bench] accuracy: {name}", file=sys.stderr, flush=True)
        g64, gmx = GpuSolver(p), GpuSolver(p, precision="mixed")
        rows = []
        for k, lc in enumerate(p.load_cases):
            g0 = lc.g0_path[0]
            row = {"load": [float(v) for v in g0]}
            res = {}
            for tag, g in (("fp64", g64), ("mixed", gmx)):
                g.reset_state()
                g.set_gradient(g0)
                g.solve()
                res[tag] = (g.homogenized_stress().copy(), _fields(g), int(g.stats.inner_iterations))
            row["iterations"] = {"fp64": res["fp64"][2], "mixed_inner": res["mixed"][2]}
            row["mixed_vs_fp64"] = {"homogenized": _rel(res["mixed"][0], res["fp64"][0]),
                                    "fields": _field_errs(res["mixed"][1], res["fp64"][1])}
            if k in oracle_cases:
                o = oracle.Oracle(p)
                o.set_gradient(g0)
                it_o = o.solve()
                fo = o.element_fields()
                fo = {"u": o.get_u(), "strain": fo["strain"], "stress": fo["stress"]}
                row["iterations"]["oracle"] = int(it_o)
                row["fp64_vs_oracle"] = {"homogenized": _rel(res["fp64"][0], o.homogenized_stress()),
                                         "fields": _field_errs(res["fp64"][1], fo)}
                del o
            rows.append(row)
        g64.close()
        gmx.close()
        return {"load_cases": rows,
                "max_mixed_vs_fp64_homogenized": max(r["mixed_vs_fp64"]["homogenized"] for r in rows),
                "max_fp64_vs_oracle_homogenized": max((r["fp64_vs_oracle"]["homogenized"] for r in rows
                                                       if "fp64_vs_oracle" in r), default=None)}

    p1 = workloads.config1_problem()
    out["configs"]["1_thermal_32"] = linear_config("config 1 (thermal 32^3, 3 load cases)", p1, range(3))
    p2 = workloads.config2_problem()
    out["configs"]["2_elastic_128"] = linear_config("config 2 (elastic 128^3, 6 load cases)", p2, [0])
    out["configs"]["2_elastic_128"]["inclusion_volume_fraction"] = workloads.inclusion_fraction(p2.ms)

    # config 3: J2 ramp, mixed vs FP64 per load step (the oracle runs the 256^3
    # residual kernel in the tests; a full 256^3 J2 step is beyond its seconds)
    n3 = args.size
    print(f"[bench] accuracy: config 3 (J2 {n3}^3, {args.j2_steps} load steps)", file=sys.stderr, flush=True)
    p3 = workloads.config3_problem(n=n3)
    mbc = workloads.config3_ramp()
    steps = {}
    for prec in ("fp64", "mixed"):
        g = GpuSolver(p3, precision=prec)
        hist = []
        for t in range(args.j2_steps):
            g.enable_mixed_bc(mbc, t)
            g.solve()
            hist.append((g.homogenized_stress().copy(), g.macro_gradient().copy(), int(g.stats.iterations)))
            g.extrapolate_displacement()
        f = _fields(g)
        f["plastic_strain"], f["hardening"] = g.get_state()
        steps[prec] = (hist, f)
        g.close()
    h64, hmx = steps["fp64"][0], steps["mixed"][0]
    out["configs"]["3_j2_256"] = {
        "grid": n3, "load_steps": args.j2_steps,
        "iterations": {"fp64": [h[2] for h in h64], "mixed": [h[2] for h in hmx]},
        "mixed_vs_fp64_homogenized_per_step": [_rel(b[0], a[0]) for a, b in zip(h64, hmx)],
        "mixed_vs_fp64_macro_strain_per_step": [_rel(b[1], a[1]) for a, b in zip(h64, hmx)],
        "mixed_vs_fp64_fields_last_step": _field_errs(steps["mixed"][1], steps["fp64"][1])}
    del steps

    # config 4: SIMP objective and its adjoint gradient, mixed vs FP64
    print("[bench] accuracy: config 4 (96^3 SIMP objective and gradient)", file=sys.stderr, flush=True)
    import torch

    from fans_amd.pyfans import voxel_problem
    rho0 = workloads.smoothed_density(96, corr=8.0, seed=0)
    params = [torch.tensor(v, dtype=torch.float64) for v in (1.0, 0.3, 1e-3, 0.3)]
    res4 = {}
    for prec in ("fp64", "mixed"):
        sv = GpuSolver(voxel_problem((96, 96, 96), tol=1e-10), precision=prec)
        rho = torch.tensor(rho0, requires_grad=True)
        J = workloads.config4_objective(sv, rho, params)
        J.backward()
        res4[prec] = (float(J.item()), rho.grad.numpy().copy(), _fields(sv))
        sv.close()
    out["configs"]["4_simp_96"] = {
        "objective_fp64": res4["fp64"][0],
        "mixed_vs_fp64_objective": abs(res4["mixed"][0] - res4["fp64"][0]) / abs(res4["fp64"][0]),
        "mixed_vs_fp64_gradient": _rel(res4["mixed"][1], res4["fp64"][1]),
        "mixed_vs_fp64_fields_adjoint_state": _field_errs(res4["mixed"][2], res4["fp64"][2])}
    del res4

    # config 5: every RVE of the batch, its mixed form (bit for bit the
    # batched contexts' results) against FP64 at 1e-10 and against FP64 at the
    # batch's own tolerance; the first RVE of each (edge, physics) group
    # against the oracle
    print("[bench] accuracy: config 5 (384 RVEs)", file=sys.stderr, flush=True)
    probs = config5_rves()
    worst, worst_same, worst64, worst_o, seen = 0.0, 0.0, 0.0, 0.0, set()
    for i, q in enumerate(probs):
        g0 = q.load_cases[0].g0_path[0]
        sig = {}
        for tag, prec, tol in (("fp64", "fp64", 1e-10), ("fp64_tol", "fp64", q.tol), ("mixed", "mixed", q.tol)):
            g = GpuSolver(replace(q, tol=tol), precision=prec)
            g.set_gradient(g0)
            g.solve()
            sig[tag] = g.homogenized_stress().copy()
            g.close()
        worst = max(worst, _rel(sig["mixed"], sig["fp64"]))
        worst_same = max(worst_same, _rel(sig["mixed"], sig["fp64_tol"]))
        worst64 = max(worst64, _rel(sig["fp64_tol"], sig["fp64"]))
        key = (q.dims[0], q.howmany)
        if key not in seen:
            seen.add(key)
            o = oracle.Oracle(replace(q, tol=1e-10))
            o.set_gradient(g0)
            o.solve()
            worst_o = max(worst_o, _rel(sig["fp64"], o.homogenized_stress()))
            del o
    out["configs"]["5_batch_384"] = {
        "rves": len(probs), "mixed_tol": probs[0].tol,
        "max_mixed_vs_fp64_homogenized": worst,
        "max_mixed_vs_fp64_same_tol_homogenized": worst_same, "max_fp64_tol_vs_fp64_1e-10_homogenized": worst64,
        "max_fp64_vs_oracle_homogenized_first_rve_per_group": worst_o, "groups_checked": len(seen)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-load-cases", action="store_true",
                    help="skip the six Mandel unit load cases reported beside the timed solves (profiling runs)")
    ap.add_argument("--workload", choices=["headline", "batch", "j2", "accuracy"], default="headline",
                    help="batch: the 384-RVE uneven-size batch (RVEs/s); j2: the J2 load-step ramp; accuracy: "
                         "per-configuration errors, mixed vs FP64 and FP64 vs oracle; none is the headline metric")
    ap.add_argument("--j2-steps", type=int, default=20, help="load steps of the j2 workload (of 20)")
    ap.add_argument("--batch-group", default="",
                    help="batch workload: only the RVEs of edge E and howmany H ('E,H'), solved in the calling thread")
    ap.add_argument("--batch-single", action="store_true",
                    help="batch workload: also run it with one single-RVE context per group (RVEs in turn)")
    ap.add_argument("--batch-workers", default="",
                    help="batch workload: also run it with these numbers of concurrent groups (e.g. 1,2,4) and "
                         "report wall time against the summed per-solve times")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="time the library's own streaming copy (16 B per lane, 2 GiB moved per copy) and print "
                         "its TB/s, the attainable bandwidth of a pure streaming pass")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (reads 1 GiB, writes 1 GiB) to scale HBM counters")
    args = ap.parse_args()
    if args.stream_ceiling:
        from fans_amd.gpu import stream_bandwidth
        for nb in (1 << 28, 1 << 30, 1 << 32):
            ms, moved = stream_bandwidth(nb, reps=20)
            print(json.dumps({"bytes_per_buffer": nb, "ms_per_copy": round(ms, 4),
                              "TB_per_s": round(moved / (ms * 1e-3) / 1e12, 3)}), flush=True)
        return
    if args.calibrate_copy:
        import torch
        a = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": a.numel() * 8}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding one RVE over GPUs "
              "(x-slabs with one-plane halos, all-to-all FFT transposes) is not built", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        run_batch_workload(args)
        return
    if args.workload == "j2":
        run_j2_workload(args)
        return
    if args.workload == "accuracy":
        run_accuracy_workload(args)
        return
    n = args.size
    p = headline_problem(n)
    N = n ** 3
    g0 = np.array([1.0, 0, 0, 0, 0, 0])
    print(f"[bench] {n}^3 elastic: FP64 form, {args.warmup} warmup + {args.steps} timed solves", file=sys.stderr,
          flush=True)
    r64 = run_gpu(p, "fp64", args.steps, args.warmup, g0, all_load_cases=not args.no_load_cases)
    tot_it = sum(r64["iters"])
    # the median of the timed solves is the reported step; the mean is
    # reported beside it
    med = int(np.argsort(r64["step_s"])[len(r64["step_s"]) // 2])
    value = N * r64["iters"][med] / r64["step_s"][med]
    ms_per_step = 1000.0 * r64["step_s"][med]
    # dominant kernel family by device time
    fam = r64["fam"]
    bpv = family_bytes_per_voxel(n)
    dom = max((k for k in fam if k in bpv), key=lambda k: fam[k][0])
    avg_ms = fam[dom][0] / max(fam[dom][1], 1)
    alg_bytes = bpv[dom] * N
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": traffic_from_profiles(dom),
                "alg_bytes_per_launch": alg_bytes, "avg_launch_ms": round(avg_ms, 4)}
    t_solve = sum(r64["solve_ms"]) * 1e-3
    iter_frac = B_ALG_FP64 * N * tot_it / t_solve / HBM_PEAK
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "voxel-iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{n}^3 linear elastic CG (HEX8, Linf abs 1e-10), 150 spheres r~U[8,20] seed 0, "
                               "matrix E=3 nu=0.35 / spheres E=70 nu=0.22, Mandel e11 load",
                   "grid": n, "cg_iterations": r64["iters"][0], "form": "fp64",
                   "inclusion_volume_fraction": round(float(np.mean(p.ms != 0)), 4)},
        "mean_ms_per_step": round(1000.0 * r64["wall_s"] / args.steps, 3),
        "mandel_unit_load_cases": {"iterations": [v[0] for v in r64["load_cases"]],
                                   "solve_ms": [v[1] for v in r64["load_cases"]]},
        "roofline": roofline,
        "iteration_hbm_frac": round(iter_frac, 4),
        "kernel_ms_per_iteration": {k: round(v[0] / max(r64["fam_iters"], 1), 4) for k, v in fam.items() if v[1]},
    }
    if not args.no_mixed:
        print(f"[bench] {n}^3 elastic: mixed form", file=sys.stderr, flush=True)
        rmx = run_gpu(p, "mixed", args.steps, args.warmup, g0)
        rel = float(np.max(np.abs(rmx["stress"] - r64["stress"])) / np.max(np.abs(r64["stress"])))
        t_mx = float(np.median(rmx["solve_ms"]))
        t_64 = float(np.median(r64["solve_ms"]))
        out["mixed"] = {"voxel_iters_per_s": round(N * sum(rmx["iters"]) / rmx["wall_s"], 1),
                        "inner_iterations": rmx["iters"][0], "outer_iterations": rmx["outer"],
                        "time_to_solution_speedup": round(t_64 / t_mx, 3), "homog_stress_rel_err": rel,
                        "converged": rmx["converged"], "solve_ms": round(t_mx, 2), "fp64_solve_ms": round(t_64, 2),
                        "iteration_hbm_frac": round(B_ALG_MIXED * N * sum(rmx["iters"]) /
                                                    (sum(rmx["solve_ms"]) * 1e-3) / HBM_PEAK, 4),
                        "kernel_ms_per_inner_iteration": {k: round(v[0] / max(rmx["fam_iters"], 1), 4)
                                                          for k, v in rmx["fam"].items() if v[1]}}
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(p, g0)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
