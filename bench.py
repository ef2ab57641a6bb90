"""Benchmark: frequency-solves/s of the Burton-Miller BEM solve on one MI355X.

Workload (the configuration the metric is quoted on): seeded head-like mesh, ellipsoid
0.075/0.09/0.11 m radially perturbed by a real spherical-harmonic sum (degree <= 8, RMS
3 mm, seed 0), geodesic frequency 32 -> 20480 triangles; two reciprocal point sources
1 mm off the surface at the +-y ear positions; 1730-point Fibonacci evaluation grid at
1.2 m; frequencies 200..16000 Hz step 200 Hz, method as NumCalc chooses it for methodBEM 4
(SL-FMM below lambda/D = 1, two-level ML-FMM above).  A frequency step is operator assembly,
the solve of both sources and the field at the evaluation grid.  A bench *step* is one pass
over the workload: the 80-frequency sweep with both sources (160 solves), one job of the
serving runtime.

value        = frequency-solves/s (one solve = one frequency x one source) of the served
               mixed-precision form over the K timed sweeps, submitted at once to the serving
               runtime (ncg.Server: lanes of grouped steps on their own streams, problems
               resident across frequencies): FP32-complex bulk operator with FP64
               near/singular integrals, FP64 Krylov dots and refinement residual, GMRES with
               the block-Jacobi preconditioner.  Only solves checked against the FP64 plain
               method at the same frequency and within rel-L2 1e-3 of it (surface and grid
               pressure) count; checked_solves says how many were compared.
plain_fp64   = the FP64 plain method over the sweep's frequencies, one step after the other
               as the drop-in NumCalc runs them: NumCalc's method choice, CGS with its ILU
               preconditioner (IPRECOND 0, what mesh2input writes), tolerance 1e-9.
p50/p99      = per-request latency of single-frequency requests (both sources) arriving as a
               Poisson stream at 70% of the measured throughput.
roofline     = the apply kernel with the largest summed launch time in the timed region, from
               HIP events around each of its launches on the stream it runs on: algorithmic
               bytes per launch / mean launch time.  `traffic` is the measured HBM bytes per
               launch from the committed counter profile (rocprofv3 FETCH_SIZE and
               WRITE_SIZE, calibrated on a copy kernel).
cpu_baseline = reference NumCalc (oracle/_ref, built from the reference sources) run the way
               manage_numcalc runs it, through the project runner: one process per (source,
               frequency step) of a sample of the sweep, at most one per physical core
               (bounded by --cpu-cores), admitted by -estimate_ram; pool throughput, per-job
               p50/p99, and the GPU FP64 results' rel-L2 against its be.out.

--workload graded runs the same measurement on the graded large case (60320 triangles, edges
1 mm near the ears to 8 mm on the far sides, 40 frequencies 12-23.7 kHz).  --workload batched
measures the batched uneven-problems configuration instead: 256 seeded small meshes
(500-6000 triangles) x 8 frequencies x 16 plane-wave load cases, dense BEM, all steps
submitted at once; its roofline is the matrix-core multi-load-case apply.  --workload stream
serves Poisson requests over 32 head-like meshes of 5k-40k triangles.
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
TRAFFIC_PROFILE = "kg_near_spmv_head20480.txt"
TRAFFIC_PROFILE_BATCHED = "k_matvec_mfma_batched.txt"
SERVED = dict(precision=ncg.MIXED, solver=ncg.GMRES, precond=ncg.BLOCK, restart=50, max_iterations=500,
              refinement_steps=8, refinement_tolerance=1e-6)


def freq_spec(src, f):
    return P.point_source_spec(src, f)


def cpu_baseline(mesh, grid, src, f, timeout, method_bem=4, what=None):
    exe = os.path.join(ROOT, "oracle", "_ref", "NumCalc")
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="ncbase_")
    try:
        prj = P.Project(mesh=mesh, freqs=[f], eval_grid=grid, point_sources=[(*src[0], 1.0, 0.0)],
                        method_bem=method_bem, c=C, rho=RHO, title="cpu baseline")
        P.write_project(d, prj)
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=timeout)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        what = what or ("reference NumCalc (ML-FMM, methodBEM 4, ILU-CGS) on the same %d-element mesh, one source, "
                        "one frequency (%%.0f Hz) incl. the 1730-point grid" % len(mesh.elems))
        return dict(value=1.0 / wall, unit="frequency-solves/s", cores=1, kind="reference",
                    sample=(what % f) + ": %.1f s wall on one core" % wall)
    except subprocess.TimeoutExpired:
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)


def isolated_kernels(geo, src, f, launches=20):
    """per-kernel launch times of the operator apply on one problem with nothing else running"""
    prob = ncg.Problem(geo["nv"], geo["xyz"], geo["kind"], None, node_index=geo["node_index"], method_bem=4)
    try:
        sp = freq_spec(src, f)
        prob.assemble(sp["k"], nrhs=sp["nrhs"], point_sources=sp["point_sources"], precision=ncg.MIXED)
        x = np.random.default_rng(0).standard_normal((sp["nrhs"], prob.n)) + 0j
        prob.apply(x)
        prob.kernel_timing(True)
        prob.kernel_stats(reset=True)
        for _ in range(launches):
            prob.apply(x)
        return prob.kernel_stats()
    finally:
        prob.close()


# VALU peaks of MI355X: FP32 157.3 TFLOP/s counts packed (two-wide) FMAs, which this build
# does not emit, so its FP32 kernels are also quoted against the one-wide 78.6; FP64 78.6
VALU_PEAK = {"fp32": 157.3e12, "fp64": 78.6e12}
FP32_QUADRATURE = ("k_fmm_near<f32>", "k_assemble_far<f32>", "k_field<f32>")


def quadrature_roofline(kst):
    """FLOP/s of the assembly and field kernels (algorithmic FLOPs: 60 per Green's-function
    evaluation counted on the device, 16 per T plane-wave term, 20 per S entry, 7 per D
    series term) against the VALU peak of their arithmetic."""
    out = {}
    for k, v in sorted(kst.items()):
        if not v.get("flops") or not v["total_ms"]:
            continue
        prec = "fp32" if k in FP32_QUADRATURE else "fp64"
        fs = v["flops"] / (v["total_ms"] * 1e-3)
        out[k] = {"calls": v["calls"], "mean_launch_ms": round(v["total_ms"] / v["calls"], 4),
                  "gflop_per_launch": round(v["flops"] / v["calls"] / 1e9, 3), "tflops": round(fs / 1e12, 3),
                  "precision": prec, "frac_of_peak": round(fs / VALU_PEAK[prec], 4)}
        if prec == "fp32":
            out[k]["frac_of_unpacked_fp32"] = round(fs / 78.6e12, 4)
    return out


def isolated_assembly(geo, src, f):
    """kernel statistics of one mixed-form assembly and field pass at f with nothing else on
    the GPU (the evaluation counts come from a first, counting pass)"""
    prob = ncg.Problem(geo["nv"], geo["xyz"], geo["kind"], geo["eval_xyz"], node_index=geo["node_index"],
                       method_bem=geo.get("method_bem", 4))
    try:
        sp = freq_spec(src, f)
        prob.kernel_timing(True)
        for _ in range(2):
            prob.kernel_stats(reset=True)
            prob.assemble(sp["k"], nrhs=sp["nrhs"], point_sources=sp["point_sources"], precision=ncg.MIXED)
            prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-5)
            prob.field(gradient=True)
        return {k: v for k, v in prob.kernel_stats().items() if v.get("flops")}
    finally:
        prob.close()


def profiled_traffic(kernel, profile=TRAFFIC_PROFILE, bytes_per_launch=None, near_modes=(0,)):
    """measured HBM bytes per launch of `kernel`, from the committed rocprofv3 counter
    summary of this benchmark, or (None, None).  The summary holds either a line
    'traffic <kernel> <bytes>' (bytes per launch of a fixed launch) or lines
    'traffic_ratio <kernel> near_mode=<m> <ratio>' (counter bytes over algorithmic bytes of
    the same dispatches); a ratio is applied to this run's algorithmic bytes per launch,
    averaged over the near-field modes the run's launch shapes use."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None, None
    ratios = {}
    with open(path) as fh:
        for line in fh:
            w = line.split()
            if len(w) == 3 and w[0] == "traffic" and w[1] == kernel:
                return float(w[2]), "profiles/" + profile
            if len(w) == 4 and w[0] == "traffic_ratio" and w[1] == kernel and w[2].startswith("near_mode="):
                ratios[int(w[2].split("=")[1])] = float(w[3])
    use = [ratios[m] for m in sorted(set(near_modes)) if m in ratios]
    if not use or bytes_per_launch is None or len(use) != len(set(near_modes)):
        return None, None
    r = sum(use) / len(use)
    return bytes_per_launch * r, "profiles/%s (counter/algorithmic ratio %.3f)" % (profile, r)


def pick(freqs, count, offset=0):
    """count frequencies spread evenly over the sweep (all of them, cyclically, if count >= len)."""
    if count <= 0:
        return []
    if count >= len(freqs):
        return [freqs[(offset + i) % len(freqs)] for i in range(count)]
    stride = len(freqs) / count
    return [freqs[int(offset + i * stride) % len(freqs)] for i in range(count)]


def batched_main(a):
    """Batched uneven problems: every (mesh, frequency) step of the batch submitted at once to
    the serving runtime; dense BEM (methodBEM 0), mixed form, the 16 load cases applied on
    the matrix cores."""
    meshes, dirs = P.batched_problems(a.meshes)
    ns = [len(m.elems) for m, _ in meshes]
    print("# batched: %d meshes (n %d..%d, mean %d), 8 frequencies x %d plane waves each; lanes=%d" % (
        len(meshes), min(ns), max(ns), int(np.mean(ns)), len(dirs), a.lanes), file=sys.stderr, flush=True)
    srv = ncg.Server(device=0, lanes=a.lanes, group=a.group, graphs=not a.eager)
    ids = []
    for m, _ in meshes:
        n = len(m.elems)
        ids.append(srv.add_mesh(np.full(n, 3), m.nodes[m.elems], np.zeros(n, np.int32), None, node_index=m.elems,
                                method_bem=0))
    for j in [srv.submit(ids[i], [P.plane_wave_spec(dirs, meshes[i][1][0])], **SERVED) for i in range(min(4, len(ids)))]:
        j.wait()
    t0 = time.perf_counter()
    jobs = [srv.submit(mid, [P.plane_wave_spec(dirs, f) for f in fq], **SERVED) for mid, (_, fq) in zip(ids, meshes)]
    sts = [j.wait() for j in jobs]
    elapsed = time.perf_counter() - t0
    steps = sum(len(fq) for _, fq in meshes)
    print("# batched: %d steps in %.2f s" % (steps, elapsed), file=sys.stderr, flush=True)
    # the kernels' shares from a second pass over the first 32 meshes with per-launch events
    srv.set_kernel_timing(True)
    srv.kernel_stats(reset=True)
    for j in [srv.submit(mid, [P.plane_wave_spec(dirs, f) for f in fq], **SERVED)
              for mid, (_, fq) in list(zip(ids, meshes))[:32]]:
        j.wait()
    kst = srv.kernel_stats()
    info = srv.stats()
    srv.close()
    # every a.check-th mesh: its last frequency against the FP64 plain method (dense CGS)
    errs, bad = [], 0
    for i in range(0, len(meshes), max(1, a.check)):
        m, fq = meshes[i]
        n = len(m.elems)
        prob = ncg.Problem(np.full(n, 3), m.nodes[m.elems], np.zeros(n, np.int32), None, node_index=m.elems,
                           method_bem=0)
        sp = P.plane_wave_spec(dirs, fq[-1])
        prob.assemble(sp["k"], nrhs=sp["nrhs"], plane_waves=sp["plane_waves"])
        prob.solve(solver=ncg.CGS, tolerance=1e-9)
        x0 = prob.solution()
        prob.close()
        e = max(np.linalg.norm(jobs[i].solution[-1][r] - x0[r]) / np.linalg.norm(x0[r]) for r in range(len(dirs)))
        errs.append(e)
        bad += e > 1e-3
    failed = sum(s["state"] != "done" for s in sts)
    valid = 0 if (bad or failed) else steps * len(dirs)
    # the matrix-core apply on one problem of about 3000 unknowns with nothing else on the GPU
    iso = None
    k_iso = min(range(len(meshes)), key=lambda i: abs(ns[i] - 3000))
    m, fq = meshes[k_iso]
    n = len(m.elems)
    prob = ncg.Problem(np.full(n, 3), m.nodes[m.elems], np.zeros(n, np.int32), None, node_index=m.elems, method_bem=0)
    try:
        sp = P.plane_wave_spec(dirs, fq[0])
        prob.assemble(sp["k"], nrhs=sp["nrhs"], plane_waves=sp["plane_waves"], precision=ncg.MIXED)
        prob.kernel_timing(True)
        prob.kernel_stats(reset=True)
        prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-5)
        st = prob.kernel_stats().get("k_matvec_mfma<f32>")
        if st and st["calls"]:
            t = st["total_ms"] / st["calls"] * 1e-3
            flops = 2.0 * n * (2 * n) * 32          # real GEMM n x 2n x 32 per pass of 16 load cases
            ach = st["bytes"] / st["calls"] / t
            iso = dict(n=n, mean_launch_ms=t * 1e3, launches=st["calls"], bytes_per_launch=st["bytes"] / st["calls"],
                       achieved_gbs=ach / 1e9, tflops=flops / t / 1e12)
    finally:
        prob.close()
    top = max((k for k in kst if kst[k]["bytes"] > 0), key=lambda k: kst[k]["total_ms"], default=None)
    roof = None
    if iso:
        traffic, tsrc = profiled_traffic("k_matvec_mfma<f32>", TRAFFIC_PROFILE_BATCHED)
        roof = {"bound": "hbm", "kernel": "k_matvec_mfma<f32>", "achieved": iso["achieved_gbs"],
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": iso["achieved_gbs"] * 1e9 / HBM_PEAK,
                "traffic": traffic, "traffic_source": tsrc,
                "mfma": {"achieved": iso["tflops"], "peak": 157.3, "unit": "TFLOP/s", "frac": iso["tflops"] / 157.3,
                         "dtype": "fp32 (v_mfma_f32_32x32x2f32)"},
                "mean_launch_ms": iso["mean_launch_ms"], "launches": iso["launches"],
                "bytes_per_launch": iso["bytes_per_launch"],
                "workload": "one problem of %d unknowns, 16 load cases, mixed GMRES; HIP events around each "
                            "launch of the matrix-core product kernel" % iso["n"],
                "served_top_kernel": top,
                "served_share_source": "a second pass over the first %d meshes with per-launch HIP events" % min(
                    32, len(meshes)),
                "share_of_timed_kernel_ms": {k: round(kst[k]["total_ms"], 2) for k in kst}}
    cpu = None
    if not a.no_cpu:
        m, fq = meshes[k_iso]
        cpu = cpu_baseline(m, None, np.zeros((1, 3)), fq[0], a.cpu_timeout, method_bem=0,
                           what="reference NumCalc (dense TBEM, ILU-CGS) on one batched mesh of %d elements, one "
                                "point-source load case, one frequency (%%.0f Hz)" % len(m.elems))
    line = {
        "metric": METRIC, "value": valid / elapsed, "unit": "frequency-solves/s", "n_gpus": 1, "steps": steps,
        "warmup": 4, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "batched uneven dense problems: %d meshes (%d-%d triangles), 8 frequencies in 0.2-8 kHz "
                               "each, 16 plane-wave load cases" % (len(meshes), min(ns), max(ns)),
                   "form": "served mixed precision, %d lanes, dense BEM, GMRES + block-Jacobi, matrix-core "
                           "multi-load-case apply" % a.lanes},
        "p50_ms": float(np.percentile([s["latency_ms"] for s in sts], 50)),
        "p99_ms": float(np.percentile([s["latency_ms"] for s in sts], 99)),
        "valid_solves": valid, "solves": steps * len(dirs),
        "mixed_vs_fp64_rel_l2": {"max": float(max(errs)) if errs else None, "meshes_checked": len(errs)},
        "server": {k: info[k] for k in ("peak_resident_bytes", "steps_done", "solves_done")},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


def stream_main(a):
    """Serving stream: 32 seeded head-like meshes (5k-40k triangles, log-uniform), each with
    the 1730-point grid; requests = (mesh, contiguous band of 1-50 frequencies in 100 Hz
    steps within 100 Hz-20 kHz, both ear sources).  The capacity comes from a batch of
    requests submitted at once; the latency from Poisson arrivals at 70% of it."""
    rng = np.random.default_rng(0)
    meshes = []
    for i in range(a.stream_meshes):
        fr = int(round(np.sqrt(np.exp(rng.uniform(np.log(5000.0), np.log(40000.0))) / 20.0)))
        meshes.append(P.head_like_mesh(freq=fr, seed=i))
    grid = P.fibonacci_grid(1730, 1.2)
    srcs = [P.ear_sources(m) for m in meshes]
    ns = [len(m.elems) for m in meshes]
    reqs = []
    for _ in range(a.requests):
        i = int(rng.integers(0, len(meshes)))
        nf = int(rng.integers(1, 51))
        f0 = 100.0 * int(rng.integers(1, 201 - nf))
        reqs.append((i, [f0 + 100.0 * q for q in range(nf)]))
    print("# stream: %d meshes (n %d..%d), %d requests of 1-50 frequencies, lanes=%d" % (
        len(meshes), min(ns), max(ns), len(reqs), a.lanes), file=sys.stderr, flush=True)
    srv = ncg.Server(device=0, lanes=a.lanes)
    ids = [srv.add_mesh(np.full(len(m.elems), 3), m.nodes[m.elems], np.zeros(len(m.elems), np.int32), grid.nodes,
                        node_index=m.elems, method_bem=4) for m in meshes]
    srv.submit(ids[reqs[0][0]], [freq_spec(srcs[reqs[0][0]], reqs[0][1][0])], **SERVED).wait()
    # capacity: the first quarter of the requests at once
    ncap = max(1, len(reqs) // 4)
    t0 = time.perf_counter()
    jobs = [srv.submit(ids[i], [freq_spec(srcs[i], f) for f in fq], **SERVED) for i, fq in reqs[:ncap]]
    sts = [j.wait() for j in jobs]
    cap_el = time.perf_counter() - t0
    cap_solves = sum(len(fq) for _, fq in reqs[:ncap]) * 2
    capacity = cap_solves / cap_el
    failed = sum(s["state"] != "done" for s in sts)
    # Poisson arrivals of the other requests at 70% of the capacity
    rest = reqs[ncap:]
    mean_solves = np.mean([len(fq) * 2 for _, fq in rest]) if rest else 1.0
    rate = 0.7 * capacity / mean_solves
    arrivals = np.cumsum(np.random.default_rng(1).exponential(1.0 / rate, len(rest)))
    pj = []
    t0 = time.perf_counter()
    for t_arr, (i, fq) in zip(arrivals, rest):
        dt = t_arr - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        pj.append((i, fq, srv.submit(ids[i], [freq_spec(srcs[i], f) for f in fq], **SERVED)))
    pst = [j.wait() for _, _, j in pj]
    p_el = time.perf_counter() - t0
    failed += sum(s["state"] != "done" for s in pst)
    for (i, fq), st in list(zip(reqs[:ncap], sts)) + [((i, fq), st) for (i, fq, _), st in zip(pj, pst)]:
        if st["state"] != "done":
            print("# failed request: mesh %d (n %d), %.0f-%.0f Hz, after %d frequencies: %s" % (
                i, ns[i], fq[0], fq[-1], st["freq_done"], st["error"]), file=sys.stderr, flush=True)
    lat = [s["latency_ms"] for s in pst] or [float("nan")]
    info = srv.stats()
    srv.close()
    print("# stream: capacity %.1f solves/s (%d solves in %.2f s); Poisson %.2f req/s: p50 %.0f ms p99 %.0f ms" % (
        capacity, cap_solves, cap_el, rate, np.percentile(lat, 50), np.percentile(lat, 99)), file=sys.stderr, flush=True)
    # a sample of the Poisson requests against the FP64 plain method at their first frequency
    errs = []
    for i, fq, j in pj[:: max(1, len(pj) // a.stream_checks)][:a.stream_checks]:
        m = meshes[i]
        prob = ncg.Problem(np.full(len(m.elems), 3), m.nodes[m.elems], np.zeros(len(m.elems), np.int32), grid.nodes,
                           node_index=m.elems, method_bem=4)
        sp = freq_spec(srcs[i], fq[0])
        prob.assemble(sp["k"], nrhs=2, point_sources=sp["point_sources"])
        prob.solve(solver=ncg.CGS, precond=ncg.ILU, tolerance=1e-9)
        x0 = prob.solution()
        p0, _ = prob.field(gradient=False)
        prob.close()
        errs.append(max(max(np.linalg.norm(j.solution[0][r] - x0[r]) / np.linalg.norm(x0[r]),
                            np.linalg.norm(j.phi[0][r] - p0[r]) / np.linalg.norm(p0[r])) for r in range(2)))
    valid = not failed and all(e <= 1e-3 for e in errs)
    line = {
        "metric": METRIC, "value": capacity if valid else 0.0, "unit": "frequency-solves/s", "n_gpus": 1,
        "steps": sum(len(fq) for _, fq in reqs), "warmup": 1, "ms_per_step": cap_el / (cap_solves / 2) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "serving stream: %d head-like meshes (%d-%d triangles) with the 1730-point grid, "
                               "requests of 1-50 frequencies in 100 Hz-20 kHz x 2 ear sources" % (
                                   len(meshes), min(ns), max(ns)),
                   "form": "served mixed precision, %d lanes, GMRES + block-Jacobi" % a.lanes},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "latency": {"poisson_rate_per_s": rate, "requests": len(rest), "offered_solves_per_s": rate * mean_solves,
                    "served_solves_per_s": sum(len(fq) * 2 for _, fq in rest) / p_el},
        "capacity": {"requests": ncap, "solves": cap_solves, "seconds": cap_el},
        "mixed_vs_fp64_rel_l2": {"max": float(max(errs)) if errs else None, "requests_checked": len(errs)},
        "failed_requests": int(failed),
        "server": {k: info[k] for k in ("peak_resident_bytes", "steps_done", "solves_done")},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


def cpu_pool(mesh, grid, src, freqs, cores, timeout):
    """Reference NumCalc (oracle/_ref, built from the reference sources) run as
    manage_numcalc runs it: a project with one source folder per ear source, one process per
    (source, frequency step) through the project runner, at most `cores` at once, admitted
    by the -estimate_ram figures.  Returns the pool's figures and the reference be.out of
    every step, or None without the reference build."""
    exe = os.path.join(ROOT, "oracle", "_ref", "NumCalc")
    if not os.path.exists(exe):
        return None
    from mesh2hrtf_amd import runner
    d = tempfile.mkdtemp(prefix="ncpool_")
    try:
        prj = P.Project(mesh=mesh, freqs=list(freqs), eval_grid=grid, method_bem=4, c=C, rho=RHO,
                        title="CPU baseline pool")
        folders = P.write_sources(d, prj, src)
        t_est = time.perf_counter()
        every, todo = runner.jobs_of(d, exe)
        est_s = time.perf_counter() - t_est
        res = runner.run(d, exe, max_instances=cores, jobs=todo)
        out = {}
        for k, folder in enumerate(folders):
            for step, f in enumerate(freqs, start=1):
                path = os.path.join(folder, "be.out", "be.%d" % step)
                if os.path.isfile(os.path.join(path, "pEvalGrid")):
                    out[(k, f)] = (P.read_be_file(os.path.join(path, "pBoundary"))[1],
                                   P.read_be_file(os.path.join(path, "pEvalGrid"))[1])
        return res, out, est_s
    finally:
        shutil.rmtree(d, ignore_errors=True)


def host_cpu():
    """lscpu figures of this host: sockets, cores per socket, threads per core, model"""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "Model name", "CPU(s)"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def head_main(a):
    if a.workload == "graded":
        freqs, mesh, iso_f = P.config3_freqs(), P.graded_head_mesh(), 18000.0
    else:
        freqs, mesh, iso_f = P.config2_freqs(), P.config2_mesh(a.nu), 8000.0
    grid = P.fibonacci_grid(1730, 1.2)
    src = P.ear_sources(mesh)
    n, m, nr = len(mesh.elems), len(grid.nodes), len(src)
    geo = dict(nv=np.full(n, 3), xyz=mesh.nodes[mesh.elems], kind=np.zeros(n, np.int32), eval_xyz=grid.nodes,
               node_index=mesh.elems, method_bem=4)
    nf = len(freqs)
    print("# n=%d eval=%d sources=%d lanes=%d group=%d; a step is one pass over %d frequencies; %d timed steps" % (
        n, m, nr, a.lanes, a.group, nf, a.steps), file=sys.stderr, flush=True)

    # ---- served mixed form: every step (sweep) submitted at once ------------------------
    srv = ncg.Server(device=0, lanes=a.lanes, kernel_timing=False, group=a.group, autotune=not a.no_autotune,
                     tuning_file=a.tuning_file, graphs=not a.eager)
    mid = srv.add_mesh(**geo)
    sweep = [freq_spec(src, f) for f in freqs]
    for j in [srv.submit(mid, sweep, **SERVED) for _ in range(a.warmup)]:
        j.wait()
    lane0 = dict(srv.stats()["lane_ms"])
    mono = {}
    mono["timed"] = [time.monotonic_ns()]
    t0 = time.perf_counter()
    jobs = [srv.submit(mid, sweep, **SERVED) for _ in range(a.steps)]
    sts = [j.wait() for j in jobs]
    elapsed = time.perf_counter() - t0
    mono["timed"].append(time.monotonic_ns())
    lane1 = srv.stats()["lane_ms"]
    failed = [s for s in sts if s["state"] != "done"]
    if failed:
        print("# served steps failed: %s" % failed[:3], file=sys.stderr, flush=True)
    print("# served: %d sweeps (%d frequency steps) in %.2f s, %.2f ms per frequency step" % (
        a.steps, a.steps * nf, elapsed, elapsed / (a.steps * nf) * 1e3), file=sys.stderr, flush=True)

    # ---- the same sweeps again with HIP events around every kernel launch (the roofline's
    # timing; events time single launches, so the Arnoldi steps run kernel by kernel here)
    nroof = max(1, min(a.steps, a.roofline_steps))
    srv.set_kernel_timing(True)
    srv.kernel_stats(reset=True)
    mono["per_launch_pass"] = [time.monotonic_ns()]
    t0r = time.perf_counter()
    rjobs = [srv.submit(mid, sweep, **SERVED) for _ in range(nroof)]
    rsts = [j.wait() for j in rjobs]
    elapsed_r = time.perf_counter() - t0r
    mono["per_launch_pass"].append(time.monotonic_ns())
    kst = srv.kernel_stats()
    srv.set_kernel_timing(False)
    print("# timed-kernel pass: %d sweeps in %.2f s, %.2f ms per frequency step" % (
        nroof, elapsed_r, elapsed_r / (nroof * nf) * 1e3), file=sys.stderr, flush=True)

    # ---- latency: single-frequency Poisson requests at 70% of the measured rate ---------
    rate = 0.7 * a.steps * nf / elapsed
    rng = np.random.default_rng(1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, a.poisson))
    pfreqs = [freqs[i] for i in rng.integers(0, nf, a.poisson)]
    pjobs = []
    t0p = time.perf_counter()
    for t_arr, f in zip(arrivals, pfreqs):
        dt = t_arr - (time.perf_counter() - t0p)
        if dt > 0:
            time.sleep(dt)
        pjobs.append(srv.submit(mid, [freq_spec(src, f)], **SERVED))
    pl = [j.wait()["latency_ms"] for j in pjobs] or [float("nan")]
    print("# poisson %.1f req/s: latency p50 %.1f ms p99 %.1f ms" % (rate, np.percentile(pl, 50), np.percentile(pl, 99)),
          file=sys.stderr, flush=True)
    info = srv.stats()
    tunings = srv.tunings()
    srv.close()

    # ---- FP64 plain method (the drop-in), one frequency after the other -----------------
    nplain = nf if a.plain_steps < 0 else min(a.plain_steps, nf)
    plain_f = pick(freqs, nplain)
    ref, plain_lat, plain_elapsed = {}, [], 0.0
    if plain_f:
        prob = ncg.Problem(geo["nv"], geo["xyz"], geo["kind"], geo["eval_xyz"], node_index=geo["node_index"],
                           method_bem=4)
        sp = freq_spec(src, plain_f[0])
        prob.assemble(sp["k"], nrhs=nr, point_sources=sp["point_sources"])
        prob.solve(solver=ncg.CGS, precond=ncg.ILU, tolerance=1e-9)
        t0 = time.perf_counter()
        for f in plain_f:
            ts = time.perf_counter()
            sp = freq_spec(src, f)
            prob.assemble(sp["k"], nrhs=nr, point_sources=sp["point_sources"])
            prob.solve(solver=ncg.CGS, precond=ncg.ILU, tolerance=1e-9)
            x0 = prob.solution()
            p0, _ = prob.field(gradient=False)
            plain_lat.append((time.perf_counter() - ts) * 1e3)
            ref[f] = (x0, p0)
        plain_elapsed = time.perf_counter() - t0
        prob.close()
    print("# plain FP64: %d frequencies in %.1f s" % (nplain, plain_elapsed), file=sys.stderr, flush=True)

    # every timed solve against the plain solution at its frequency; only checked solves count
    err_s, err_e, checked, valid = [], [], 0, 0
    for j, s in zip(jobs, sts):
        for q, f in enumerate(freqs):
            if f not in ref or s["state"] != "done":
                continue
            x0, p0 = ref[f]
            es = max(np.linalg.norm(j.solution[q][r] - x0[r]) / np.linalg.norm(x0[r]) for r in range(nr))
            ee = max(np.linalg.norm(j.phi[q][r] - p0[r]) / np.linalg.norm(p0[r]) for r in range(nr))
            err_s.append(es)
            err_e.append(ee)
            checked += nr
            valid += nr if (es <= 1e-3 and ee <= 1e-3) else 0
    value = valid / elapsed

    # ---- roofline: the apply kernel with the largest share of the timed region ----------
    top = max((k for k in kst if kst[k]["bytes"] > 0), key=lambda k: kst[k]["total_ms"], default=None)
    roof = None
    if top:
        s = kst[top]
        ach = s["bytes"] / (s["total_ms"] * 1e-3)
        modes = [t["near_mode"] for _, t in tunings] if tunings else [0]
        traffic, tsrc = profiled_traffic(top, bytes_per_launch=s["bytes"] / s["calls"], near_modes=modes)
        roof = {"bound": "hbm", "kernel": top, "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": traffic, "traffic_source": tsrc,
                "bytes_per_launch": s["bytes"] / s["calls"], "mean_launch_ms": s["total_ms"] / s["calls"],
                "launches": s["calls"],
                "timing": "HIP events around each launch on the group's stream, over %d further sweeps of the "
                          "same workload run right after the timed region with per-launch events (%.2f ms per "
                          "frequency step there); %d lanes sharing the GPU" % (
                              nroof, elapsed_r / (nroof * nf) * 1e3, a.lanes),
                "share_of_timed_kernel_ms": {k: round(kst[k]["total_ms"], 2) for k in sorted(kst)},
                "timed_kernels": {k: {"calls": kst[k]["calls"], "total_ms": round(kst[k]["total_ms"], 3),
                                      "algorithmic_bytes": kst[k]["bytes"]} for k in sorted(kst) if kst[k]["bytes"]},
                "quadrature_served": quadrature_roofline(kst)}
        if not a.no_isolated:
            mono["isolated"] = [time.monotonic_ns()]
            iso = isolated_kernels(geo, src, iso_f)
            roof["isolated_single_problem"] = {
                k: {"mean_launch_ms": v["total_ms"] / v["calls"], "achieved_gbs": v["bytes"] / v["total_ms"] / 1e6}
                for k, v in iso.items() if v["calls"] and v["bytes"]}
            roof["quadrature_isolated"] = {"%.0f Hz" % f: quadrature_roofline(isolated_assembly(geo, src, f))
                                           for f in (iso_f, freqs[-1])}
            mono["isolated"].append(time.monotonic_ns())

    # ---- CPU baseline: the reference NumCalc as a manage_numcalc-style pool -------------
    cpu = None
    if not a.no_cpu:
        from mesh2hrtf_amd import runner
        cores = min(runner.physical_cores(), a.cpu_cores)
        cf = pick(freqs, max(1, cores // nr))
        pool = cpu_pool(mesh, grid, src, cf, cores, a.cpu_timeout)
        if pool:
            res, out, est_s = pool
            sm = res.summary()
            e_ref = []
            for (k, f), (pB, pE) in out.items():
                if f in ref:
                    x0, p0 = ref[f]
                    rp = RHO * 2 * np.pi * f
                    e_ref.append(max(np.linalg.norm(1j * rp * x0[k] - pB) / np.linalg.norm(pB),
                                     np.linalg.norm(1j * rp * p0[k] - pE) / np.linalg.norm(pE)))
            cpu = {"value": sm["jobs"] / sm["wall_s"], "unit": "frequency-solves/s", "cores": cores,
                   "kind": "reference",
                   "sample": "reference NumCalc (oracle/_ref, methodBEM 4: SL/ML-FMM, ILU-CGS) on the same %d-element "
                             "mesh with the 1730-point grid, run like manage_numcalc: one process per (source, "
                             "frequency step), %d at once, %d jobs (%d frequencies %s Hz x %d sources), admitted by "
                             "-estimate_ram; wall from the first start to the last exit" % (
                                 n, cores, sm["jobs"], len(cf), "/".join("%.0f" % f for f in cf), nr),
                   "wall_s": sm["wall_s"], "job_p50_s": sm["p50_s"], "job_p99_s": sm["p99_s"],
                   "max_running": sm["max_running"], "failed_jobs": sm["failed"], "estimate_ram_s": est_s,
                   "host": host_cpu(),
                   "gpu_fp64_vs_reference_rel_l2_max": float(max(e_ref)) if e_ref else None}
            print("# CPU pool: %d jobs in %.1f s on %d cores; GPU FP64 vs reference max rel-L2 %s" % (
                sm["jobs"], sm["wall_s"], cores, cpu["gpu_fp64_vs_reference_rel_l2_max"]), file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "frequency-solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": ("head-like mesh %d triangles (seed 0), 2 ear point sources, 1730-point grid at 1.2 m, "
                                "frequencies 200-16000 Hz step 200; NumCalc methodBEM 4 (SL/ML-FMM); one step = one "
                                "pass over the %d frequencies x 2 sources" % (n, nf))
                   if a.workload == "head" else
                   ("graded head-like mesh %d triangles (seed 0; 1 mm edges near the ears to 8 mm), 2 ear point "
                    "sources, 1730-point grid at 1.2 m, frequencies 12000-23700 Hz step 300; NumCalc methodBEM 4; "
                    "one step = one pass over the %d frequencies x 2 sources" % (n, nf)),
                   "form": "served mixed precision: FP32 operator and inner GMRES(50) + block-Jacobi, FP64 refinement "
                           "to relative residual %g (residual with FP64 accumulation and FMM signatures); %d lanes, "
                           "%d steps per grouped solve; Arnoldi steps %s" % (
                               SERVED["refinement_tolerance"], a.lanes, a.group,
                               "launched kernel by kernel" if a.eager else "replayed as captured HIP graphs"),
                   "frequencies": nf},
        "p50_ms": float(np.percentile(pl, 50)),
        "p99_ms": float(np.percentile(pl, 99)),
        "latency": {"requests": a.poisson, "poisson_rate_per_s": rate,
                    "what": "single-frequency requests (both sources) arriving as a Poisson stream at 70% of the "
                            "measured rate; submit to result"},
        "ms_per_frequency_step": elapsed / (a.steps * nf) * 1e3,
        "valid_solves": valid,
        "checked_solves": checked,
        "solves": a.steps * nf * nr,
        "plain_fp64": None if not nplain else {"value": nplain * nr / plain_elapsed, "ms_per_frequency": plain_elapsed / nplain * 1e3,
                       "p50_ms": float(np.percentile(plain_lat, 50)), "p99_ms": float(np.percentile(plain_lat, 99)),
                       "frequencies": nplain,
                       "form": "the drop-in's FP64 plain method: NumCalc's method choice, ILU-CGS to 1e-9, one "
                               "frequency at a time"},
        "speedup_vs_plain": value / (nplain * nr / plain_elapsed) if nplain else None,
        "refinement": {"tolerance": SERVED["refinement_tolerance"],
                       "steps_max": max(s["refinement_steps_max"] for s in sts),
                       "relative_residual_max": max(s["relative_residual_max"] for s in sts),
                       "sweeps_with_every_step_within_tolerance": "%d of %d" % (
                           sum(1 for s in sts if s["state"] == "done" and s["relative_residual_max"] <= SERVED[
                               "refinement_tolerance"]), a.steps)},
        "mixed_vs_fp64_rel_l2": {"surface_max": float(max(err_s)) if err_s else None,
                                 "eval_grid_max": float(max(err_e)) if err_e else None},
        "server": {"peak_resident_bytes": info["peak_resident_bytes"],
                   "arnoldi_steps": {"replayed_from_graphs": info["graph_steps"], "kernel_by_kernel": info["eager_steps"],
                                     "note": "whole run: the timed sweeps replay captured graphs, the per-launch "
                                             "timed pass and the tuning runs launch kernel by kernel"},
                   "lane_ms_timed": {k: round(lane1[k] - lane0[k], 1) for k in lane1},
                   "launch_shapes": {k: t for k, t in tunings} if tunings else "defaults (autotune off)"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "regions_monotonic_ns": mono,
    }
    print(json.dumps(line, default=lambda o: o.item() if hasattr(o, "item") else str(o)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3,
                    help="timed steps; a step is one pass over the workload's frequency sweep (80 frequencies x 2 "
                         "sources for the head mesh)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed sweeps")
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--group", type=int, default=4, help="served steps a lane solves together (ncg_solve_group)")
    ap.add_argument("--nu", type=int, default=32, help="geodesic frequency of the mesh (32 -> 20480 triangles)")
    ap.add_argument("--plain-steps", type=int, default=-1,
                    help="frequencies the FP64 plain method runs (-1: every frequency of the sweep)")
    ap.add_argument("--poisson", type=int, default=160, help="requests of the latency run")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=16, help="CPU baseline: at most this many NumCalc processes")
    ap.add_argument("--no-autotune", action="store_true",
                    help="served run: keep the default launch shapes instead of measuring them per operator shape")
    ap.add_argument("--refinement-tolerance", type=float, default=None,
                    help="served mixed form: FP64 refinement residual target (default %g)" % SERVED["refinement_tolerance"])
    ap.add_argument("--tuning-file", default=None,
                    help="served run: launch shapes per operator shape to load (and to save what is tuned)")
    ap.add_argument("--no-isolated", action="store_true", help="skip the isolated kernel timing")
    ap.add_argument("--eager", action="store_true",
                    help="served run: launch the GMRES Arnoldi steps kernel by kernel instead of replaying graphs")
    ap.add_argument("--roofline-steps", type=int, default=5,
                    help="sweeps of the per-launch timed pass that gives the roofline (at most --steps)")
    ap.add_argument("--cpu-timeout", type=float, default=120.0)
    ap.add_argument("--stream-meshes", type=int, default=32, help="stream: number of meshes")
    ap.add_argument("--requests", type=int, default=96, help="stream: number of requests")
    ap.add_argument("--stream-checks", type=int, default=4, help="stream: requests verified against FP64")
    ap.add_argument("--workload", choices=("head", "graded", "batched", "stream"), default="head",
                    help="head: the served head-mesh sweep (default); graded: the 60k-triangle graded head mesh "
                         "over 12-24 kHz; batched: many small dense problems with 16 plane-wave load cases each; "
                         "stream: Poisson requests over 32 head-like meshes of 5k-40k triangles")
    ap.add_argument("--meshes", type=int, default=256, help="batched: number of meshes")
    ap.add_argument("--check", type=int, default=16, help="batched: every check-th mesh is verified against FP64")
    a = ap.parse_args()
    if a.refinement_tolerance is not None:
        SERVED["refinement_tolerance"] = a.refinement_tolerance
    if a.gpus != 1:
        print("bench.py measures one GPU: the solve shards by independent (frequency, source) jobs with no "
              "communication, so only --gpus 1 is run here", file=sys.stderr)
        sys.exit(2)
    if a.workload == "batched":
        return batched_main(a)
    if a.workload == "stream":
        return stream_main(a)
    return head_main(a)


if __name__ == "__main__":
    main()
