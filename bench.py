"""Benchmark: frequency-solves/s of the Burton–Miller BEM solve on one MI355X.

Workload (the configuration the metric is quoted on): seeded head-like mesh, ellipsoid
0.075/0.09/0.11 m radially perturbed by a real spherical-harmonic sum (degree <= 8, RMS
3 mm, seed 0), geodesic frequency 32 -> 20480 triangles; two reciprocal point sources
1 mm off the surface at the +-y ear positions; 1730-point Fibonacci evaluation grid at
1.2 m; frequencies 200..16000 Hz step 200 Hz.  A step is one frequency: dense operator
assembly, the Krylov solve of both sources (they share the operator) and the field at the
evaluation grid.  Steps walk the 80 frequencies with an even stride.

value      = frequency-solves/s of the mixed-precision form (FP32-complex bulk quadrature,
             operator and inner matvecs; FP64 singular/near integrals, Krylov dots and the
             refinement residual), counted only if every surface and evaluation-grid
             pressure is within rel-L2 1e-3 of the FP64 plain method on the same frequency.
plain      = the same steps with the FP64 plain method (the GPU drop-in).
roofline   = the dominant kernel (the operator matvec, HBM bound): algorithmic bytes per
             launch (operator once + B vectors in and out) / mean launch time (HIP events).
cpu_baseline = reference NumCalc (oracle/_ref, built from the reference sources) on one
             frequency of one source of the same mesh, ML-FMM as the configuration asks,
             on one host core.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mesh2hrtf_amd import ncg, project as P  # noqa: E402

METRIC = "frequency-solves/sec at rel-L2≤1e-3 vs FP64 oracle; p50/p99 solve latency"
C, RHO = 343.0, 1.1839
HBM_PEAK = 8.0e12


def setup(freq_nu):
    mesh = P.config2_mesh(freq_nu)
    grid = P.fibonacci_grid(1730, 1.2)
    src = P.ear_sources(mesh)
    n = len(mesh.elems)
    xyz = mesh.nodes[mesh.elems]
    prob = ncg.Problem(np.full(n, 3), xyz, np.zeros(n, np.int32), grid.nodes, device=0)
    return mesh, grid, src, prob


def step(prob, src, f, precision):
    omega = 2 * np.pi * f
    k = omega / C
    rp = RHO * omega
    amp = 1.0 / (1j * rp)
    ps = [[s[0], s[1], s[2], amp.real, amp.imag, r] for r, s in enumerate(src)]
    prob.assemble(k, nrhs=len(src), point_sources=ps, precision=precision)
    if precision == ncg.FP64:
        st = prob.solve(solver=ncg.CGS, precond=ncg.NONE, tolerance=1e-9)
    else:
        st = prob.solve(solver=ncg.GMRES, restart=50, refinement_steps=8, refinement_tolerance=1e-7,
                        max_iterations=500)
    x = prob.solution()
    phi, _ = prob.field(gradient=False)
    return x, phi, st


def cpu_baseline(mesh, grid, src, f, timeout):
    exe = os.path.join(ROOT, "oracle", "_ref", "NumCalc")
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="ncbase_")
    try:
        prj = P.Project(mesh=mesh, freqs=[f], eval_grid=grid, point_sources=[(*src[0], 1.0, 0.0)],
                        method_bem=4, c=C, rho=RHO, title="cpu baseline")
        P.write_project(d, prj)
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        return dict(value=1.0 / wall, unit="frequency-solves/s", cores=1, kind="reference",
                    sample="reference NumCalc (ML-FMM, method 4) on the same 20480-element mesh, one source, "
                           "one frequency (%.0f Hz) incl. the 1730-point grid: %.1f s wall on one core" % (f, wall))
    except subprocess.TimeoutExpired:
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nu", type=int, default=32, help="geodesic frequency of the mesh (32 -> 20480 triangles)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-timeout", type=float, default=120.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: the solve shards by independent (frequency, source) jobs with no "
              "communication, so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)
    freqs = P.config2_freqs()
    t_setup = time.perf_counter()
    mesh, grid, src, prob = setup(a.nu)
    t_setup = time.perf_counter() - t_setup
    n = len(mesh.elems)
    total = a.warmup + a.steps
    stride = max(1, len(freqs) // total)
    idx = [(i * stride + stride // 2) % len(freqs) for i in range(total)]
    warm, timed = idx[:a.warmup], idx[a.warmup:]
    print("# n=%d eval=%d sources=%d setup %.2fs; timed frequencies %s" % (
        n, len(grid.nodes), len(src), t_setup, [freqs[i] for i in timed]), file=sys.stderr, flush=True)

    out = {}
    for name, prec in (("mixed", ncg.MIXED), ("fp64", ncg.FP64)):
        for i in warm:
            step(prob, src, freqs[i], prec)
        prob.timings(reset=True)
        lat, res, mv_ms, far_ms, asm_ms, sol_ms, fld_ms = [], [], [], [], [], [], []
        t0 = time.perf_counter()
        for i in timed:
            ts = time.perf_counter()
            x, phi, st = step(prob, src, freqs[i], prec)
            lat.append(time.perf_counter() - ts)
            tm = prob.timings()
            asm_ms.append(tm["ms_assemble"])
            sol_ms.append(tm["ms_solve"])
            fld_ms.append(tm["ms_field"])
            res.append((x, phi, st["iterations"]))
            print("# %s f=%5.0f  %.1f ms  (assemble %.1f, solve %.1f [%s it], field %.1f)" % (
                name, freqs[i], lat[-1] * 1e3, tm["ms_assemble"], tm["ms_solve"], st["iterations"], tm["ms_field"]),
                file=sys.stderr, flush=True)
        elapsed = time.perf_counter() - t0
        tm = prob.timings()
        out[name] = dict(elapsed=elapsed, lat=lat, res=res, matvec_ms=tm["matvec_kernel_ms"],
                         matvec_calls=tm["matvec_calls"], far_ms=tm["far_kernel_ms"],
                         far_evals=tm["far_evaluations"], asm=np.mean(asm_ms), sol=np.mean(sol_ms),
                         fld=np.mean(fld_ms))

    # accuracy of the mixed form against the FP64 plain method, per frequency
    err_s, err_e = [], []
    for (x1, p1, _), (x0, p0, _) in zip(out["mixed"]["res"], out["fp64"]["res"]):
        err_s.append(max(np.linalg.norm(x1[r] - x0[r]) / np.linalg.norm(x0[r]) for r in range(len(src))))
        err_e.append(max(np.linalg.norm(p1[r] - p0[r]) / np.linalg.norm(p0[r]) for r in range(len(src))))
    ok = max(err_s + err_e) <= 1e-3
    chosen = "mixed" if ok else "fp64"
    o = out[chosen]
    solves = a.steps * len(src)
    value = solves / o["elapsed"]
    nrhs = len(src)
    if chosen == "mixed":
        bytes_launch = n * n * 8 + 2 * nrhs * n * 16
    else:
        bytes_launch = n * n * 16 + 2 * nrhs * n * 16
    achieved = bytes_launch / (o["matvec_ms"] * 1e-3) if o["matvec_ms"] > 0 else 0.0
    cpu = None if a.no_cpu else cpu_baseline(mesh, grid, src, 8000.0, a.cpu_timeout)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "frequency-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": o["elapsed"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if chosen == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "head-like mesh %d triangles (seed 0), 2 ear point sources, 1730-point grid at 1.2 m, "
                               "frequencies 200-16000 Hz; dense Burton-Miller BEM" % n,
                   "form": chosen, "frequencies": [freqs[i] for i in idx[a.warmup:]]},
        "p50_ms": float(np.percentile(o["lat"], 50) * 1e3),
        "p99_ms": float(np.percentile(o["lat"], 99) * 1e3),
        "plain_fp64": {"value": solves / out["fp64"]["elapsed"],
                       "ms_per_step": out["fp64"]["elapsed"] / a.steps * 1e3,
                       "p50_ms": float(np.percentile(out["fp64"]["lat"], 50) * 1e3),
                       "p99_ms": float(np.percentile(out["fp64"]["lat"], 99) * 1e3)},
        "mixed_vs_fp64_rel_l2": {"surface_max": float(max(err_s)), "eval_grid_max": float(max(err_e))},
        "phase_ms": {k: {"assemble": float(out[k]["asm"]), "solve": float(out[k]["sol"]), "field": float(out[k]["fld"])}
                     for k in out},
        "roofline": {"bound": "hbm", "kernel": "k_matvec (%s operator, %d vectors)" % (
                         "FP32" if chosen == "mixed" else "FP64", nrhs),
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "bytes_per_launch": bytes_launch, "mean_launch_ms": o["matvec_ms"],
                     "launches": o["matvec_calls"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
