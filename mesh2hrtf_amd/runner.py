"""Project runner: one NumCalc process per (source, frequency step), as many at once as the
cores and the memory budget allow, resuming what is not finished.

Counterpart of mesh2hrtf's manage_numcalc (mesh2hrtf/NumCalc/manage_numcalc.py:10-430),
which drives any executable that keeps NumCalc's command line and file contract: the
drop-in ``mesh2hrtf_amd/bin/NumCalc`` or the reference NumCalc.

* Projects: a Mesh2HRTF project folder holds ``NumCalc/source_<k>/NC.inp`` per source
  (manage_numcalc.py:120-150, 534-538); a folder of such projects is searched one level
  down.
* Jobs: one per (source, frequency step) with the RAM estimate of ``NumCalc
  -estimate_ram`` (Memory.txt, written once per source folder; manage_numcalc.py:528-545).
* Done: a step is finished when ``be.out/be.<step>/pEvalGrid`` exists and its log
  ``NC<step>-<step>.out``, if present, holds ``End time:`` (manage_numcalc.py:552-583);
  everything else runs again (resume).
* Admission: at most ``max_instances`` processes (default: the physical cores) and the sum
  of the running jobs' estimates, times ``ram_safety``, within ``ram_budget_gb`` (default:
  the available memory).  Jobs start highest estimate first, alternating with the lowest
  (the reference's "alternate" order).  Unlike manage_numcalc, admission accounts the
  estimates of the jobs this runner started instead of sampling system RAM and CPU load
  after a fixed wait, so it needs no wait time and ignores unrelated processes.
* Result: per-job wall time (process start to exit), the whole run's wall time, p50/p99,
  and the steps that did not finish.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import subprocess
import sys
import time
from dataclasses import dataclass, field


@dataclass
class Job:
    project: str
    source: int
    step: int
    freq: float
    ram_gb: float

    @property
    def folder(self) -> str:
        return os.path.join(self.project, "NumCalc", "source_%d" % self.source)


@dataclass
class RunResult:
    jobs: list = field(default_factory=list)          # Job, in start order
    seconds: list = field(default_factory=list)       # per-job wall time
    returncodes: list = field(default_factory=list)
    wall_s: float = 0.0
    unfinished: list = field(default_factory=list)    # (project, source, step) after the run
    max_running: int = 0

    def summary(self) -> dict:
        import numpy as np
        t = np.asarray(self.seconds) if self.seconds else np.zeros(1)
        return dict(jobs=len(self.jobs), wall_s=self.wall_s,
                    jobs_per_s=len(self.jobs) / self.wall_s if self.wall_s > 0 else None,
                    p50_s=float(np.percentile(t, 50)), p99_s=float(np.percentile(t, 99)),
                    max_running=self.max_running, failed=sum(1 for r in self.returncodes if r != 0),
                    unfinished=len(self.unfinished))


def physical_cores() -> int:
    """Physical cores of this host (logical CPUs if that is unknown), within the CPUs
    this process may run on."""
    try:
        import psutil
        phys = psutil.cpu_count(logical=False) or psutil.cpu_count()
    except ImportError:
        phys = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return max(1, min(phys, avail))


def available_ram_gb() -> float:
    try:
        import psutil
        return psutil.virtual_memory().available / 2 ** 30
    except ImportError:
        return float("inf")


def find_projects(path: str) -> list:
    """Project folders at ``path``: the folder itself if it holds NumCalc/source_*, else
    its subfolders that do."""
    def is_project(p):
        return bool(glob.glob(os.path.join(p, "NumCalc", "source_*", "NC.inp")))
    if is_project(path):
        return [path]
    return sorted(os.path.join(path, d) for d in os.listdir(path)
                  if os.path.isdir(os.path.join(path, d)) and is_project(os.path.join(path, d)))


def read_memory_txt(path: str) -> list:
    """Memory.txt lines ``step frequency GB`` (src/NC_Main.cpp:722-731)."""
    out = []
    with open(path) as fh:
        for line in fh:
            w = line.split()
            if len(w) == 3:
                out.append((int(w[0]), float(w[1]), float(w[2])))
    return out


def source_folders(project: str) -> list:
    """NumCalc/source_1 .. source_N in source order."""
    n = len(glob.glob(os.path.join(project, "NumCalc", "source_*")))
    return [(k, os.path.join(project, "NumCalc", "source_%d" % k)) for k in range(1, n + 1)]


def estimates(folder: str, numcalc: str) -> list:
    """The source folder's RAM estimates, running ``numcalc -estimate_ram`` once."""
    mem = os.path.join(folder, "Memory.txt")
    if not os.path.isfile(mem):
        subprocess.run([numcalc, "-estimate_ram"], cwd=folder, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, check=True)
    return read_memory_txt(mem)


def step_finished(folder: str, step: int) -> bool:
    if not os.path.isfile(os.path.join(folder, "be.out", "be.%d" % step, "pEvalGrid")):
        return False
    log = os.path.join(folder, "NC%d-%d.out" % (step, step))
    if os.path.isfile(log):
        with open(log, errors="replace") as fh:
            return "End time:" in fh.read()
    return True


def jobs_of(project: str, numcalc: str):
    """(all jobs, unfinished jobs) of a project."""
    every, todo = [], []
    for k, folder in source_folders(project):
        for step, freq, gb in estimates(folder, numcalc):
            j = Job(project, k, step, freq, gb)
            every.append(j)
            if not step_finished(folder, step):
                todo.append(j)
    return every, todo


def start_order(jobs: list, order: str = "alternate") -> list:
    """Highest estimate first ('high'), lowest first ('low'), or alternating between the
    two ends ('alternate')."""
    s = sorted(jobs, key=lambda j: j.ram_gb)
    if order == "low":
        return s
    if order == "high":
        return s[::-1]
    out, lo, hi, take_hi = [], 0, len(s) - 1, True
    while lo <= hi:
        if take_hi:
            out.append(s[hi])
            hi -= 1
        else:
            out.append(s[lo])
            lo += 1
        take_hi = not take_hi
    return out


def run(path: str, numcalc: str, max_instances: int | None = None, ram_budget_gb: float | None = None,
        ram_safety: float = 1.05, order: str = "alternate", jobs: list | None = None, env: dict | None = None,
        poll_s: float = 0.02, log=None) -> RunResult:
    """Runs the unfinished steps of every project at ``path`` (or the given ``jobs``) with
    ``numcalc``; returns the per-job and whole-run timings."""
    if not os.path.basename(numcalc).startswith("NumCalc"):
        raise ValueError("the executable must be a NumCalc (its file name starts with 'NumCalc'): %s" % numcalc)
    if not os.access(numcalc, os.X_OK):
        raise FileNotFoundError("NumCalc executable not found: %s" % numcalc)
    numcalc = os.path.abspath(numcalc)
    if jobs is None:
        jobs = []
        for prj in find_projects(path):
            jobs += jobs_of(prj, numcalc)[1]
    max_instances = max_instances or physical_cores()
    budget = available_ram_gb() if ram_budget_gb is None else ram_budget_gb
    for j in jobs:
        if j.ram_gb * ram_safety > budget:
            raise MemoryError("source %d step %d of %s needs %.2f GB, beyond the %.2f GB budget"
                              % (j.source, j.step, j.project, j.ram_gb * ram_safety, budget))
    queue = start_order(jobs, order)
    res = RunResult()
    running = []   # (Popen, job, start time, log handle)
    t0 = time.perf_counter()
    while queue or running:
        # reap finished processes
        still = []
        for p, j, ts, fh in running:
            rc = p.poll()
            if rc is None:
                still.append((p, j, ts, fh))
                continue
            fh.close()
            res.jobs.append(j)
            res.seconds.append(time.perf_counter() - ts)
            res.returncodes.append(rc)
            if log:
                log("finished source %d step %d (%.0f Hz) in %.1f s, rc %d" % (j.source, j.step, j.freq,
                                                                            res.seconds[-1], rc))
        running = still
        # admit: the first queued job (in start order) that fits the memory left
        used = sum(j.ram_gb for _, j, _, _ in running) * ram_safety
        started = False
        if len(running) < max_instances:
            for i, j in enumerate(queue):
                if used + j.ram_gb * ram_safety <= budget:
                    fh = open(os.path.join(j.folder, "NC%d-%d_log.txt" % (j.step, j.step)), "w")
                    p = subprocess.Popen([numcalc, "-istart", str(j.step), "-iend", str(j.step)], cwd=j.folder,
                                         stdout=fh, stderr=subprocess.STDOUT, env=env)
                    running.append((p, j, time.perf_counter(), fh))
                    res.max_running = max(res.max_running, len(running))
                    queue.pop(i)
                    started = True
                    if log:
                        log("started source %d step %d (%.0f Hz, %.2f GB estimated)" % (j.source, j.step, j.freq,
                                                                                        j.ram_gb))
                    break
        if not started:
            time.sleep(poll_s)
    res.wall_s = time.perf_counter() - t0
    res.unfinished = [(j.project, j.source, j.step) for j in jobs if not step_finished(j.folder, j.step)]
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="run the unfinished NumCalc steps of Mesh2HRTF projects")
    ap.add_argument("path", nargs="?", default=os.getcwd(), help="a project folder or a folder of projects")
    ap.add_argument("--numcalc", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin", "NumCalc"))
    ap.add_argument("--instances", type=int, default=None, help="parallel processes (default: physical cores)")
    ap.add_argument("--ram", type=float, default=None, help="memory budget in GB (default: available RAM)")
    ap.add_argument("--ram-safety", type=float, default=1.05)
    ap.add_argument("--order", choices=("alternate", "high", "low"), default="alternate")
    a = ap.parse_args(argv)
    res = run(a.path, a.numcalc, a.instances, a.ram, a.ram_safety, a.order, log=lambda m: print(m, flush=True))
    s = res.summary()
    print(json.dumps(s))
    if res.unfinished:
        for prj, k, step in res.unfinished:
            print("unfinished: %s source %d step %d" % (prj, k, step), file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
