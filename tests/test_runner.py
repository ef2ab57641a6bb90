"""The project runner (mesh2hrtf_amd.runner, the manage_numcalc counterpart) on the CPU,
driving the reference NumCalc compiled from its sources (oracle/_ref; test infrastructure):

* every (source, step) of a two-source project runs as its own process, at most
  ``max_instances`` at once, and the be.out files equal those of one full NumCalc run;
* resume: only steps without be.out/be.N/pEvalGrid, or whose log lacks "End time:", run
  again (manage_numcalc.py:552-583);
* memory admission by the -estimate_ram figures: a budget below one step's estimate is an
  error, a budget that fits one step at a time runs them one after the other.
"""
import filecmp
import os
import shutil
import subprocess

import pytest

import ncref as R
from mesh2hrtf_amd import project as P, runner

REF = os.path.join(R.ROOT, "oracle", "_ref", "NumCalc")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("reference NumCalc not built (make -C oracle ref)")
    return REF


def small_project(folder):
    prj = P.Project(mesh=P.icosphere(2, 0.1), freqs=[500.0, 1200.0, 2500.0], eval_grid=P.fibonacci_grid(40, 0.6),
                    title="runner test")
    return P.write_sources(str(folder), prj, [(0.0, 0.0, 0.25), (0.2, 0.1, 0.0)])


def test_runner_runs_every_step_within_the_instance_limit(tmp_path, ref):
    folders = small_project(tmp_path / "prj")
    res = runner.run(str(tmp_path / "prj"), ref, max_instances=2)
    s = res.summary()
    assert s["jobs"] == 6 and s["failed"] == 0 and s["unfinished"] == 0
    assert 1 <= res.max_running <= 2
    # the same files as one process over all steps
    for d in folders:
        full = str(tmp_path / ("full_" + os.path.basename(d)))
        shutil.copytree(d, full, ignore=shutil.ignore_patterns("be.out", "NC*.out", "NC*_log.txt", "Memory.txt"))
        subprocess.run([ref], cwd=full, check=True, stdout=subprocess.DEVNULL)
        for step in (1, 2, 3):
            for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
                a = os.path.join(d, "be.out", "be.%d" % step, key)
                b = os.path.join(full, "be.out", "be.%d" % step, key)
                assert filecmp.cmp(a, b, shallow=False), (d, step, key)


def test_runner_resumes_unfinished_steps(tmp_path, ref):
    folders = small_project(tmp_path / "prj")
    runner.run(str(tmp_path / "prj"), ref, max_instances=3)
    os.remove(os.path.join(folders[0], "be.out", "be.2", "pEvalGrid"))
    log = os.path.join(folders[1], "NC3-3.out")
    with open(log) as fh:
        txt = fh.read().replace("End time:", "")
    with open(log, "w") as fh:
        fh.write(txt)
    _, todo = runner.jobs_of(str(tmp_path / "prj"), ref)
    assert sorted((j.source, j.step) for j in todo) == [(1, 2), (2, 3)]
    res = runner.run(str(tmp_path / "prj"), ref, max_instances=3)
    assert sorted((j.source, j.step) for j in res.jobs) == [(1, 2), (2, 3)]
    assert res.unfinished == []
    assert runner.jobs_of(str(tmp_path / "prj"), ref)[1] == []


def test_runner_memory_admission(tmp_path, ref):
    small_project(tmp_path / "prj")
    every, _ = runner.jobs_of(str(tmp_path / "prj"), ref)
    top = max(j.ram_gb for j in every)
    with pytest.raises(MemoryError):
        runner.run(str(tmp_path / "prj"), ref, max_instances=4, ram_budget_gb=0.5 * top)
    res = runner.run(str(tmp_path / "prj"), ref, max_instances=4, ram_budget_gb=1.5 * top * 1.05)
    assert res.max_running == 1 and res.summary()["jobs"] == 6


def test_start_order_alternates_high_and_low():
    jobs = [runner.Job("p", 1, s, 100.0 * s, gb) for s, gb in enumerate([0.5, 0.1, 0.9, 0.3, 0.7], start=1)]
    assert [j.ram_gb for j in runner.start_order(jobs)] == [0.9, 0.1, 0.7, 0.3, 0.5]
    assert [j.ram_gb for j in runner.start_order(jobs, "low")] == [0.1, 0.3, 0.5, 0.7, 0.9]
    assert runner.physical_cores() >= 1
