"""Serving runtime (ncg_server_* through the ctypes binding): continuous batching of
frequency steps over lanes, resident problems, memory admission.

* a job's results are the same alone and in a mix of concurrent jobs on other meshes
  and frequencies (the steps are deterministic: equal to 1e-12);
* served steps equal the problem API's steps on the same inputs (1e-12), and the mixed
  form stays within rel-L2 1e-4 of the FP64 plain method (the metric's bar is 1e-3);
* a job whose step cannot fit the memory budget fails with NCG_E_MEMORY and leaves the
  server usable; statistics count every finished step and load case.
"""
import numpy as np
import pytest

import ncref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ncg():
    from mesh2hrtf_amd import ncg as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


def mesh_args(case, method_bem):
    pa = R.problem_arrays(case, case["spec"]["freqs"][0])
    return dict(nv=pa["nv"], xyz=pa["xyz"], kind=pa["kind"], eval_xyz=case["eval_nodes"], node_index=case["elems"],
                method_bem=method_bem)


def freq_specs(case, freqs, nrhs=1):
    out = []
    for f in freqs:
        pa = R.problem_arrays(case, f)
        ps = [list(v) + [r] for r in range(nrhs) for v in pa["ps"]]
        pw = [list(v) + [r] for r in range(nrhs) for v in pa["pw"]]
        bc = np.repeat(pa["bc"][:, 0][None, :], nrhs, axis=0)
        out.append(dict(k=pa["k"], bc=bc, adm=pa["adm"], point_sources=ps, plane_waves=pw, nrhs=nrhs))
    return out


def problem_reference(ncg, case, method_bem, specs, precision, **solver):
    a = mesh_args(case, method_bem)
    prob = ncg.Problem(a["nv"], a["xyz"], a["kind"], a["eval_xyz"], node_index=a["node_index"], method_bem=method_bem)
    xs, phis = [], []
    for f in specs:
        prob.assemble(f["k"], bc=f["bc"], adm=f["adm"], point_sources=f["point_sources"],
                      plane_waves=f["plane_waves"], precision=precision)
        prob.solve(**solver)
        xs.append(prob.solution())
        phis.append(prob.field(gradient=False)[0])
    prob.close()
    return np.array(xs), np.array(phis)


SOLVER = dict(solver=1, precond=3, max_iterations=500, restart=50, refinement_steps=8, refinement_tolerance=1e-5)


def test_results_same_alone_and_in_the_mix(ncg):
    head = R.load_case("fmm_ml_head")
    sph = R.load_case("fmm_sl_point")
    dense = R.load_case("head8")
    srv = ncg.Server(lanes=3)
    try:
        mh = srv.add_mesh(**mesh_args(head, 4))
        ms = srv.add_mesh(**mesh_args(sph, 1))
        md = srv.add_mesh(**mesh_args(dense, 0))
        fa = freq_specs(head, [3000.0, 5000.0, 8000.0], nrhs=2)
        print("server: job A alone", flush=True)
        a0 = srv.submit(mh, fa, **SOLVER)
        st = a0.wait(timeout=120)
        assert st["state"] == "done", st
        print("server: job A in a mix", flush=True)
        jobs = [srv.submit(ms, freq_specs(sph, [1000.0, 2000.0, 3000.0]), **SOLVER),
                srv.submit(mh, fa, **SOLVER),
                srv.submit(md, freq_specs(dense, [2000.0, 6000.0]), solver=0, precond=2, precision=ncg.FP64),
                srv.submit(mh, freq_specs(head, [4000.0, 6000.0]), **SOLVER)]
        for j in jobs:
            s = j.wait(timeout=180)
            assert s["state"] == "done", s
        a1 = jobs[1]
        assert np.max(np.abs(a1.solution - a0.solution)) <= 1e-12 * np.max(np.abs(a0.solution))
        assert np.max(np.abs(a1.phi - a0.phi)) <= 1e-12 * np.max(np.abs(a0.phi))
        # served steps equal the problem API's
        x_ref, phi_ref = problem_reference(ncg, head, 4, fa, ncg.MIXED, **SOLVER)
        assert np.max(np.abs(a0.solution - x_ref)) <= 1e-12 * np.max(np.abs(x_ref))
        assert np.max(np.abs(a0.phi - phi_ref)) <= 1e-12 * np.max(np.abs(phi_ref))
        xd, phid = problem_reference(ncg, dense, 0, freq_specs(dense, [2000.0, 6000.0]), ncg.FP64, solver=0,
                                     precond=2)
        assert np.max(np.abs(jobs[2].solution - xd)) <= 1e-12 * np.max(np.abs(xd))
        # mixed form against the FP64 plain method
        x64, phi64 = problem_reference(ncg, head, 4, fa, ncg.FP64, solver=0, precond=0)
        for i in range(len(fa)):
            assert R.rel_l2(a0.solution[i], x64[i]) < 1e-4
            assert R.rel_l2(a0.phi[i], phi64[i]) < 1e-4
        info = srv.stats()
        assert info["jobs_done"] == 5 and info["jobs_failed"] == 0
        assert info["steps_done"] == 3 + 3 + 3 + 2 + 2
        assert info["solves_done"] == 6 + 3 + 6 + 2 + 2
        assert 0 < info["p50_ms"] <= info["p99_ms"]
        assert info["peak_resident_bytes"] > 0
    finally:
        srv.close()


def test_memory_admission(ncg):
    head = R.load_case("fmm_ml_head")
    srv = ncg.Server(lanes=2, memory_budget=64.0 * 2 ** 20)
    try:
        mh = srv.add_mesh(**mesh_args(head, 4))
        print("server: job beyond the budget", flush=True)
        j = srv.submit(mh, freq_specs(head, [8000.0]), **SOLVER)
        st = j.wait(timeout=60)
        assert st["state"] == "failed" and st["error_code"] == -3 and "budget" in st["error"]
        info = srv.stats()
        assert info["jobs_failed"] == 1 and info["queued_steps"] == 0
    finally:
        srv.close()
    srv = ncg.Server(lanes=1)
    try:
        mh = srv.add_mesh(**mesh_args(head, 4))
        j = srv.submit(mh, freq_specs(head, [3000.0]), **SOLVER)
        assert j.wait(timeout=60)["state"] == "done"
    finally:
        srv.close()


def test_failed_step_ends_job_after_running_steps(ncg):
    """One frequency of a multi-frequency job fails (a point source with a load-case index
    out of range, rejected by ncg_assemble on the host) on a 2-lane server: the job reports
    failed with that error, and only after every running step of it has finished, so no
    output is written after wait returns; the server then serves the next job."""
    case = R.load_case("head8")
    srv = ncg.Server(lanes=2)
    try:
        md = srv.add_mesh(**mesh_args(case, 0))
        specs = freq_specs(case, [1000.0, 1500.0, 2000.0, 2500.0, 3000.0, 3500.0])
        bad = dict(specs[2])
        bad["point_sources"] = [list(v[:5]) + [7] for v in bad["point_sources"]]
        specs[2] = bad
        print("server: job with a failing step", flush=True)
        j = srv.submit(md, specs, solver=0, precond=2, precision=ncg.FP64)
        st = j.wait(timeout=120)
        assert st["state"] == "failed" and st["error_code"] == -1 and "rhs index" in st["error"], st
        snap_x, snap_phi = j.solution.copy(), j.phi.copy()
        import time
        time.sleep(1.0)
        assert np.array_equal(j.solution, snap_x) and np.array_equal(j.phi, snap_phi)
        info = srv.stats()
        assert info["jobs_failed"] == 1 and info["queued_steps"] == 0
        ok = srv.submit(md, freq_specs(case, [2000.0]), solver=0, precond=2, precision=ncg.FP64)
        assert ok.wait(timeout=60)["state"] == "done"
    finally:
        srv.close()
