"""Launch-shape tuning (ncg_autotune, ncg_set_tuning, the server's per-shape tunings).

* the near-field work-group mode, the near-field assembly mapping (thread per column or
  wave per collocation row) and the dense far-field rows per work-group change no value:
  results bit for bit equal;
* the far-field waves per work-group and the matrix-core splits change only the summation
  order: solutions within rel-L2 1e-10 (FP64) of the default shapes';
* ncg_autotune picks valid shapes from measurements on a group and sets them on every
  member; a server with autotune keeps one tuning per operator shape and writes / reads its
  tuning file.
"""
import numpy as np
import pytest

import ncref as R

pytestmark = pytest.mark.gpu


def _fmm_group(ncg, name="fmm_ml_head", freqs=(3000.0, 5000.0, 8000.0), precision=None):
    case = R.load_case(name)
    probs = []
    for f in freqs:
        pa = R.problem_arrays(case, f)
        p = ncg.Problem(pa["nv"], pa["xyz"], pa["kind"], case["eval_nodes"], node_index=case["elems"],
                        method_bem=case["spec"]["method_bem"])
        ps = [list(v) + [r] for r in range(2) for v in pa["ps"]]
        p.assemble(pa["k"], nrhs=2, point_sources=ps, precision=ncg.FP64 if precision is None else precision)
        probs.append(p)
    return probs


def test_near_mode_and_waves():
    from mesh2hrtf_amd import ncg
    probs = _fmm_group(ncg)
    opts = dict(solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-10)
    res = {}
    for key, tune in (("default", {}), ("cluster", dict(near_mode=1)), ("queue", dict(near_mode=2)),
                      ("queue4", dict(near_mode=3)), ("waves4", dict(t_waves=4, d_waves=2, s_waves=4)),
                      ("rows", dict(near_assembly=1))):
        for p in probs:
            p.set_tuning(**tune)
        if key == "rows":   # the near field assembled again with a wave per collocation row
            for p, f in zip(probs, (3000.0, 5000.0, 8000.0)):
                pa = R.problem_arrays(R.load_case("fmm_ml_head"), f)
                p.assemble(pa["k"], nrhs=2, point_sources=[list(v) + [r] for r in range(2) for v in pa["ps"]],
                           precision=ncg.FP64)
        print("tuning", key, flush=True)
        st = ncg.solve_group(probs, **opts)
        assert all(s["converged"] for s in st)
        res[key] = [p.solution() for p in probs]
    for other in ("cluster", "queue", "queue4"):
        for a, b in zip(res["default"], res[other]):
            assert np.array_equal(a, b), other
    for a, b in zip(res["default"], res["rows"]):
        assert np.array_equal(a, b)
    for a, b in zip(res["default"], res["waves4"]):
        for r in range(2):
            assert R.rel_l2(b[r], a[r]) < 1e-10
    t = ncg.autotune(probs, reps=2)
    print("chosen", t, flush=True)
    assert t["tuned"] == 1 and t["t_waves"] in (1, 2, 4, 8) and t["d_waves"] in (1, 2, 4, 8) and t["s_waves"] in (1, 2, 4, 8)
    assert t["near_mode"] in (0, 1, 2, 3) and t["near_assembly"] in (0, 1)
    for p in probs:
        assert p.tuning() == t
    st = ncg.solve_group(probs, **opts)
    assert all(s["converged"] for s in st)
    for p, a in zip(probs, res["default"]):
        x = p.solution()
        for r in range(2):
            assert R.rel_l2(x[r], a[r]) < 1e-10
    for p in probs:
        p.close()


def test_near_modes_mixed_precision():
    """the FP32-valued near field in every work-group mode: the same mixed-form solutions bit
    for bit (the persistent launches included; the work-queue mode runs twice in a row so
    the second solve starts from the queues the first left)"""
    from mesh2hrtf_amd import ncg
    probs = _fmm_group(ncg, precision=ncg.MIXED)
    opts = dict(solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-6, refinement_tolerance=1e-6)
    res = {}
    for key, mode in (("chunk", 0), ("cluster", 1), ("queue", 2), ("queue4", 3), ("queue_again", 2)):
        for p in probs:
            p.set_tuning(near_mode=mode)
        print("near_mode", key, flush=True)
        st = ncg.solve_group(probs, **opts)
        assert all(s["converged"] for s in st)
        res[key] = [p.solution() for p in probs]
    for other in ("cluster", "queue", "queue4", "queue_again"):
        for a, b in zip(res["chunk"], res[other]):
            assert np.array_equal(a, b), other
    for p in probs:
        p.close()


def test_dense_rows_and_splits():
    from mesh2hrtf_amd import ncg
    case = R.load_case("head8")
    f = case["spec"]["freqs"][-1]
    pa = R.problem_arrays(case, f)
    rng = np.random.default_rng(2)
    dirs = rng.standard_normal((8, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pw = [list(dirs[r]) + [1e-3, 1e-4 * r, r] for r in range(8)]
    prob = ncg.Problem(pa["nv"], pa["xyz"], pa["kind"], None, node_index=case["elems"], method_bem=0)
    systems = []
    for rows in (16, 64):
        prob.set_tuning(assemble_rows=rows)
        prob.assemble(pa["k"], nrhs=8, plane_waves=pw)
        systems.append(prob.system()[0])
    assert np.array_equal(systems[0], systems[1])
    x = rng.standard_normal((8, pa["n"])) + 1j * rng.standard_normal((8, pa["n"]))
    ys = []
    for splits in (0, 1, 3, 16):
        prob.set_tuning(mfma_splits=splits)
        ys.append(prob.apply(x))
    for y in ys[1:]:
        for r in range(8):
            assert R.rel_l2(y[r], ys[0][r]) < 1e-12
    t = ncg.autotune([prob], reps=2)
    assert t["tuned"] == 1 and 0 <= t["mfma_splits"] <= 16 and t["assemble_rows"] in (16, 32, 64, 128)
    prob.close()


def test_server_autotune_and_tuning_file(tmp_path):
    from mesh2hrtf_amd import ncg
    from test_gpu_server import freq_specs, mesh_args
    case = R.load_case("fmm_ml_head")
    path = str(tmp_path / "tuning.txt")
    srv = ncg.Server(lanes=2, group=2, autotune=True, tuning_file=path)
    mid = srv.add_mesh(**mesh_args(case, 4))
    j = srv.submit(mid, freq_specs(case, [3000.0, 5000.0, 6000.0, 8000.0], nrhs=2), refinement_tolerance=1e-5)
    assert j.wait(timeout=120)["state"] == "done"
    tun = srv.tunings()
    print("server tunings", tun, flush=True)
    assert tun and all(t["tuned"] == 1 for _, t in tun)
    srv.close()
    text = open(path).read()
    assert all(k in text for k, _ in tun)
    srv2 = ncg.Server(lanes=1, tuning_file=path)
    assert [k for k, _ in srv2.tunings()] == [k for k, _ in tun]
    srv2.close()
