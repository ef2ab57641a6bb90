"""GMRES Arnoldi steps replayed as captured HIP graphs (ncg_set_graphs, ncg_graph_stats).

Every launch of an Arnoldi step reads the step index from device memory, so one captured
graph serves every step of a solve, and a later solve whose launches are the same reuses it
from the leader's cache.  Replaying the graph issues the same launches as launching kernel
by kernel, so solutions, iteration counts and refinement steps are equal bit for bit:

* a group of three FMM problems (two-level ML-FMM, 2 load cases) in FP64 GMRES and in the
  mixed form (inner FP32 GMRES, FP64 refinement), each solved twice with graphs: every
  solve replays the step graph many times with eager launches and host waits between the
  replays (cycle starts, flag reads), and an eager apply and a lone eager solve run between
  the two group solves; the second group solve captures nothing new;
* a group of two dense problems of different size with 16 plane-wave load cases in the
  mixed form (FP32 matrix-core products inside the graph);
* an FMM group with more than 4 load cases is not captured and runs kernel by kernel.
"""
import numpy as np
import pytest

import ncref as R

pytestmark = pytest.mark.gpu
FREQS = (3000.0, 5000.0, 8000.0)


def _fmm_group(ncg, precision, nrhs=2):
    case = R.load_case("fmm_ml_head")
    probs = []
    for f in FREQS:
        pa = R.problem_arrays(case, f)
        p = ncg.Problem(pa["nv"], pa["xyz"], pa["kind"], case["eval_nodes"], node_index=case["elems"],
                        method_bem=case["spec"]["method_bem"])
        ps = [list(v) + [r] for r in range(nrhs) for v in pa["ps"]]
        p.assemble(pa["k"], nrhs=nrhs, point_sources=[[s[0], s[1], s[2], s[3] * (1 + 0.1 * s[5]), s[4], s[5]]
                                                       for s in ps], precision=precision)
        probs.append(p)
    return probs


def _solve(ncg, probs, graphs, **opts):
    for p in probs:
        p.graphs(graphs)
    st = ncg.solve_group(probs, **opts)
    assert all(s["converged"] for s in st), st
    return [p.solution() for p in probs], [(s["iterations"], s["refinement_steps"]) for s in st]


@pytest.mark.parametrize("mixed", [False, True])
def test_fmm_group_graph_replays_equal_eager(mixed):
    from mesh2hrtf_amd import ncg
    precision = ncg.MIXED if mixed else ncg.FP64
    probs = _fmm_group(ncg, precision)
    opts = dict(solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-10, refinement_tolerance=1e-6)
    print("eager group solve, mixed=%d" % mixed, flush=True)
    x_e, it_e = _solve(ncg, probs, False, **opts)
    s0 = probs[0].graph_stats()
    assert s0["captured"] == 0 and s0["replayed"] == 0 and s0["eager"] > 0
    print("graph group solve 1", flush=True)
    x_g, it_g = _solve(ncg, probs, True, **opts)
    s1 = probs[0].graph_stats()
    assert s1["captured"] >= 1 and s1["replayed"] > 0 and s1["eager"] == s0["eager"], (s0, s1)
    assert it_g == it_e
    for a, b in zip(x_e, x_g):
        assert np.array_equal(a, b)
    # eager work on the leader's stream between the graph solves
    rng = np.random.default_rng(0)
    v = rng.standard_normal((2, probs[0].n)) + 1j * rng.standard_normal((2, probs[0].n))
    y0 = probs[0].apply(v)
    print("lone eager solve between the graph solves", flush=True)
    x_lone, it_lone = _solve(ncg, probs[1:2], False, **opts)
    print("graph group solve 2", flush=True)
    x_g2, it_g2 = _solve(ncg, probs, True, **opts)
    s2 = probs[0].graph_stats()
    assert s2["captured"] == s1["captured"] and s2["replayed"] > s1["replayed"], (s1, s2)
    assert it_g2 == it_e
    for a, b in zip(x_e, x_g2):
        assert np.array_equal(a, b)
    assert np.array_equal(probs[0].apply(v), y0)
    # the lone solve took the same path as a lone graph solve
    x_lone_g, it_lone_g = _solve(ncg, probs[1:2], True, **opts)
    assert it_lone_g == it_lone and np.array_equal(x_lone_g[0], x_lone[0])
    for p in probs:
        p.close()


def test_dense_group_sixteen_load_cases_graph_equal_eager():
    from mesh2hrtf_amd import ncg
    rng = np.random.default_rng(3)
    dirs = rng.standard_normal((16, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    probs = []
    for name in ("head8", "soft_pw"):
        case = R.load_case(name)
        pa = R.problem_arrays(case, case["spec"]["freqs"][-1])
        p = ncg.Problem(pa["nv"], pa["xyz"], pa["kind"], None, node_index=case["elems"], method_bem=0)
        p.assemble(pa["k"], nrhs=16, plane_waves=[list(dirs[r]) + [1.0, 0.0, r] for r in range(16)],
                   precision=ncg.MIXED)
        probs.append(p)
    assert probs[0].n != probs[1].n
    opts = dict(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-6)
    x_e, it_e = _solve(ncg, probs, False, **opts)
    for rep in range(2):
        print("dense graph solve %d" % rep, flush=True)
        x_g, it_g = _solve(ncg, probs, True, **opts)
        assert it_g == it_e
        for a, b in zip(x_e, x_g):
            assert np.array_equal(a, b)
    s = probs[0].graph_stats()
    assert s["captured"] >= 1 and s["replayed"] > 0, s
    for p in probs:
        p.close()


def test_fmm_many_load_cases_runs_kernel_by_kernel():
    from mesh2hrtf_amd import ncg
    probs = _fmm_group(ncg, ncg.FP64, nrhs=6)[:2]
    x_g, it_g = _solve(ncg, probs, True, solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-8)
    s = probs[0].graph_stats()
    assert s["captured"] == 0 and s["replayed"] == 0 and s["eager"] > 0, s
    x_e, it_e = _solve(ncg, probs, False, solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-8)
    assert it_g == it_e
    for a, b in zip(x_e, x_g):
        assert np.array_equal(a, b)
    for p in probs:
        p.close()
