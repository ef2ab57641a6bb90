"""Batched uneven dense problems, reduced: 8 seeded meshes of 500-2420 triangles
(project.batched_problems) x 2 frequencies x 16 plane-wave load cases, every step submitted
at once to the serving runtime in the mixed form (FP32 operator with FP32 matrix-core
products for the 16 load cases, GMRES + block-Jacobi, FP64 refinement).

* every load case of every step within rel-L2 1e-4 of the single-problem FP64 plain solve
  (dense CGS, tolerance 1e-9) of the same step;
* mesh 3 (a perturbed sphere of 1620 triangles) also against reference NumCalc's be.out
  for its first plane wave (fixture case_c4_mesh3.npz): surface pressure within 1e-4.
"""
import numpy as np
import pytest

import ncref as R
from mesh2hrtf_amd import project as P

pytestmark = pytest.mark.gpu
TOL = 1e-4


def test_batched_uneven_dense_problems():
    from mesh2hrtf_amd import ncg
    meshes, dirs = P.batched_problems(8)
    srv = ncg.Server(device=0, lanes=4)
    ids = []
    for m, _ in meshes:
        n = len(m.elems)
        ids.append(srv.add_mesh(np.full(n, 3), m.nodes[m.elems], np.zeros(n, np.int32), None, node_index=m.elems,
                                method_bem=0))
    jobs = [srv.submit(mid, [P.plane_wave_spec(dirs, f) for f in fq[:2]], precision=ncg.MIXED, solver=ncg.GMRES,
                       precond=ncg.BLOCK, restart=50, max_iterations=500, refinement_steps=8,
                       refinement_tolerance=1e-6)
            for mid, (_, fq) in zip(ids, meshes)]
    sts = [j.wait() for j in jobs]
    srv.close()
    assert all(s["state"] == "done" for s in sts), sts
    worst = 0.0
    for i, ((m, fq), job) in enumerate(zip(meshes, jobs)):
        n = len(m.elems)
        prob = ncg.Problem(np.full(n, 3), m.nodes[m.elems], np.zeros(n, np.int32), None, node_index=m.elems,
                           method_bem=0)
        for s, f in enumerate(fq[:2]):
            sp = P.plane_wave_spec(dirs, f)
            prob.assemble(sp["k"], nrhs=16, plane_waves=sp["plane_waves"])
            assert prob.solve(solver=ncg.CGS, tolerance=1e-9)["converged"]
            x0 = prob.solution()
            errs = [R.rel_l2(job.solution[s][r], x0[r]) for r in range(16)]
            print("mesh %d n=%d f=%.0f max rel-L2 %.2e" % (i, n, f, max(errs)), flush=True)
            worst = max(worst, max(errs))
        prob.close()
    assert worst < TOL, worst
    # mesh 3 against the reference's be.out (first plane wave, amplitude 1)
    case = R.load_case("c4_mesh3")
    assert np.allclose(case["nodes"], meshes[3][0].nodes) and np.array_equal(case["elems"], meshes[3][0].elems)
    for s, f in enumerate(case["spec"]["freqs"]):
        assert f == meshes[3][1][s]
        rp = 1.1839 * 2 * np.pi * f
        e = R.rel_l2(1j * rp * jobs[3].solution[s][0], case["pBoundary"][s])
        print("mesh 3 f=%.0f vs reference NumCalc: %.2e" % (f, e), flush=True)
        assert e < TOL, e
