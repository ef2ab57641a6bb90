"""The drop-in NumCalc executable: command line, file contracts, end-to-end parity.

CPU tests cover what runs without a device: the command line (src/NC_Main.cpp:150-233),
the RAM estimate (Memory.txt), the C-ABI exports.  GPU tests run whole projects through
``mesh2hrtf_amd/bin/NumCalc`` and compare its be.out with the reference's.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import ncref as R
from mesh2hrtf_amd import project as P

ROOT = R.ROOT
NUMCALC = os.path.join(ROOT, "mesh2hrtf_amd", "bin", "NumCalc")
HEADER = os.path.join(ROOT, "include", "ncg.h")


def _need_binary():
    if not os.path.exists(NUMCALC):
        pytest.fail("drop-in NumCalc not built: run __graft_entry__.build()")


def project_from_case(folder, name, steps=None):
    c = R.load_case(name)
    spec = c["spec"]
    mesh = P.Mesh(c["nodes"], c["elems"], elem_ids=c["elem_ids"])
    grid = None
    if len(c["eval_nodes"]):
        from scipy.spatial import ConvexHull
        pts = c["eval_nodes"]
        grid = P.Mesh(pts, ConvexHull(pts).simplices, node_ids=P.EVAL_ID0 + np.arange(len(pts)),
                      elem_ids=P.EVAL_ID0 + np.arange(len(ConvexHull(pts).simplices)))
    prj = P.Project(mesh=mesh, freqs=spec["freqs"] if steps is None else [spec["freqs"][s] for s in steps],
                    eval_grid=grid, boundary=[P.Boundary(*b) for b in spec["boundary"]],
                    point_sources=[tuple(s) for s in spec["point_sources"]],
                    plane_waves=[tuple(s) for s in spec["plane_waves"]], method_solver=spec["method_solver"],
                    method_bem=spec["method_bem"], precond=spec["precond"], c=spec["c"], rho=spec["rho"],
                    symmetry=spec.get("symmetry"), test_point_source=spec.get("test_point_source", False))
    P.write_project(folder, prj)
    return c


def test_header_symbols_exported():
    """libncg exports every entry point include/ncg.h declares (no device needed)."""
    from mesh2hrtf_amd import ncg as N
    lib = N.load()
    declared = set(re.findall(r"^\w[\w\s\*]*?\b(ncg_\w+)\s*\(", open(HEADER).read(), re.M))
    assert declared == set(N.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_cli_unknown_flag(tmp_path):
    _need_binary()
    r = subprocess.run([NUMCALC, "-bogus"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 255
    assert "NumCalc was called with an unknown parameter or flag. Use NumCalc -h for help." in r.stderr


def test_cli_help_and_bad_ranges(tmp_path):
    _need_binary()
    r = subprocess.run([NUMCALC, "-h"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0 and "-estimate_ram" in r.stdout
    r = subprocess.run([NUMCALC, "-istart", "-1"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 255
    r = subprocess.run([NUMCALC, "-istart", "3", "-iend", "2"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 255 and "istart is bigger than iend" in r.stderr


def test_missing_end_is_a_parse_error(tmp_path):
    _need_binary()
    project_from_case(str(tmp_path), "quad")
    txt = open(tmp_path / "NC.inp").read().replace("\nEND\n", "\n")
    open(tmp_path / "NC.inp", "w").write(txt)
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 255 and "Error parsing the input file" in r.stdout


def test_estimate_ram_memory_txt(tmp_path):
    """TBEM RAM estimate and Memory.txt format (src/NC_Main.cpp:535-541, 722-731)."""
    _need_binary()
    case = project_from_case(str(tmp_path), "c1")
    os.makedirs(tmp_path / "be.out" / "be.1")
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = open(tmp_path / "Memory.txt").read().splitlines()
    n = len(case["elems"])
    assert len(lines) == len(case["spec"]["freqs"])
    for i, (line, f) in enumerate(zip(lines, case["spec"]["freqs"])):
        step, freq, gb = line.split()
        assert int(step) == i + 1 and float(freq) == f
        assert float(gb) == pytest.approx((n + n * n) * 16 / 1e9, rel=1e-5)
    assert lines[0] == "1 100 0.0262349"      # ostream default formatting, as the reference prints
    assert os.path.isdir(tmp_path / "be.out" / "be.1")   # estimate keeps be.out


def check_tables(ours, ref, skip_nonzeros=False):
    """NC.out analysis lines of every step equal the reference's."""
    assert len(ours) == len(ref)
    for s, (a, b) in enumerate(zip(ours, ref)):
        if skip_nonzeros:
            b = [l for l in b if "onzeros" not in l]
        assert len(a) == len(b), (s, a, b)
        for x, y in zip(a, b):
            assert R.same_table_line(x, y), (s, x, y)


@pytest.mark.parametrize("name", ["c1", "quad", "fmm_sl_point", "fmm_sl_adm", "fmm_ml_head", "fmm_ml_patch"])
def test_estimate_ram_matches_reference(tmp_path, name):
    """-estimate_ram: Memory.txt and the method / cluster / expansion tables of NC.out
    equal the reference's (TBEM, SL-FMM and ML-FMM estimates of src/NC_Main.cpp:524-731;
    the cluster plan comes from the host planner, no device needed)."""
    _need_binary()
    case = project_from_case(str(tmp_path), name)
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert open(tmp_path / "Memory.txt").read() == str(case["memory_txt"])
    check_tables(R.nc_tables(open(tmp_path / "NC.out").read()), json.loads(str(case["nc_tables"])),
                 skip_nonzeros=True)


def test_symmetry_input_errors(tmp_path):
    """SYMMETRY block checks of src/NC_Input.cpp:459-495, 678-699 (before any device use)."""
    _need_binary()
    case = project_from_case(str(tmp_path), "sym_half")
    txt = open(tmp_path / "NC.inp").read()
    open(tmp_path / "NC.inp", "w").write(txt.replace("SYMMETRY\n0 0 1\n", "SYMMETRY\n0 0 5\n"))
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "must be equal to 0, 1, 2, 3, or 4" in r.stdout
    open(tmp_path / "NC.inp", "w").write(txt.replace("SYMMETRY\n0 0 1\n", "SYMMETRY\n1 0 1\n"))
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "j must equal to numSymmetricPlanes_" in r.stdout
    # a plane through the mesh: x = 0 cuts the half sphere
    open(tmp_path / "NC.inp", "w").write(txt.replace("SYMMETRY\n0 0 1\n", "SYMMETRY\n1 0 0\n"))
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0 and "All BE nodes must be located on the same side of a symmetric plane!" in r.stdout
    del case


@pytest.mark.parametrize("name", ["sym_half", "sym_fmm_ml"])
def test_symmetry_estimate_ram_runs(tmp_path, name):
    """-estimate_ram on a symmetric project plans the mesh unfolded across its planes."""
    _need_binary()
    case = project_from_case(str(tmp_path), name)
    r = subprocess.run([NUMCALC, "-estimate_ram"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    lines = open(tmp_path / "Memory.txt").read().splitlines()
    assert len(lines) == len(case["spec"]["freqs"])
    log = open(tmp_path / "NC.out").read()
    assert "Number of symmetric planes .................. = %d" % sum(1 for t in case["spec"]["symmetry"][0] if t) in log
    assert "Number of equations = %d" % len(case["elems"]) in log


# rel-L2 bounds against the reference be.out: dense 1e-6 (7 printed digits); FMM 1e-5 (CGS
# iteration-path rounding).  Symmetry-plane projects are solved unfolded across the planes:
# dense ones to 1e-6; FMM ones differ from the reference's reflected-cluster FMM by the two
# FMM truncation errors (the reference's own runs are 1.4e-4 / 3.5e-4 from the exact surface
# solution and 3e-5 on the grid)
DROPIN_TOL = {"fmm_sl_point": 1e-5, "fmm_ml_head": 1e-5, "sym_fmm_sl": 2e-3, "sym_fmm_ml": 2e-3}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["soft_pw", "mixed_pres", "quad", "gauss", "velo_patch", "fmm_sl_point",
                                  "fmm_ml_head", "sym_half", "sym_quarter_anti", "sym_octant", "sym_fmm_sl",
                                  "sym_fmm_ml", "test_source_1pw", "test_source_2pw"])
def test_dropin_matches_reference_be_out(tmp_path, name):
    """be.out of every step and, without symmetry planes, the CGS iteration counts and the
    NC.out analysis tables (the unfolded symmetric system is larger than NumCalc's)."""
    _need_binary()
    case = project_from_case(str(tmp_path), name)
    tol = DROPIN_TOL.get(name, 1e-6)
    print("drop-in case", name, flush=True)
    r = subprocess.run([NUMCALC], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    log = open(tmp_path / "NC.out").read()
    assert "End time:" in log
    for s in range(len(case["spec"]["freqs"])):
        assert "Step %d, Frequency = " % (s + 1) in log
        for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
            ids, vals = P.read_be_file(str(tmp_path / "be.out" / ("be.%d" % (s + 1)) / key))
            assert np.array_equal(ids, case[key + "_ids"])
            ref = case[key][s]
            if np.linalg.norm(ref) == 0:
                assert np.abs(vals).max() < 1e-12
            else:
                assert R.rel_l2(vals, ref) < tol, (key, s)
    if name.startswith("sym_"):
        return
    if case["spec"]["method_solver"] != 4:
        its = [int(v) for v in re.findall(r"CGS solver: number of iterations = (\d+), relative error = ", log)]
        assert its == list(case["cgs_iterations"])
    check_tables(R.nc_tables(log), json.loads(str(case["nc_tables"])))


@pytest.mark.gpu
def test_dropin_step_range_and_log_names(tmp_path):
    _need_binary()
    case = project_from_case(str(tmp_path), "adm_ps")
    r = subprocess.run([NUMCALC, "-istart", "2", "-iend", "2"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0
    assert os.path.exists(tmp_path / "NC2-2.out")
    assert not os.path.exists(tmp_path / "be.out" / "be.1")
    ids, vals = P.read_be_file(str(tmp_path / "be.out" / "be.2" / "pEvalGrid"))
    assert R.rel_l2(vals, case["pEvalGrid"][1]) < 1e-6
    r = subprocess.run([NUMCALC, "-nitermax", "0", "-istart", "1", "-iend", "1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    log = open(tmp_path / "NC1-1.out").read()
    assert "CGS solver: number of iterations = 0" in log
    assert "Warning: Maximum number of iterations is reached!" in log


@pytest.mark.gpu
def test_runner_drives_the_dropin_like_the_reference(tmp_path):
    """The project runner (the manage_numcalc counterpart) drives the drop-in: one process per
    (source, step), three at once; every be.out equals the reference NumCalc's, run the same
    way, to its printed digits."""
    _need_binary()
    from mesh2hrtf_amd import runner
    ref_exe = os.path.join(ROOT, "oracle", "_ref", "NumCalc")
    if not os.path.exists(ref_exe):
        pytest.fail("reference NumCalc not built (make -C oracle ref)")
    prj = P.Project(mesh=P.icosphere(2, 0.1), freqs=[600.0, 1800.0, 3000.0], eval_grid=P.fibonacci_grid(40, 0.6),
                    title="runner drop-in")
    srcs = [(0.0, 0.0, 0.25), (0.15, -0.1, 0.05)]
    ours = P.write_sources(str(tmp_path / "ours"), prj, srcs)
    refs = P.write_sources(str(tmp_path / "ref"), prj, srcs)
    print("runner: drop-in", flush=True)
    r1 = runner.run(str(tmp_path / "ours"), NUMCALC, max_instances=3)
    r2 = runner.run(str(tmp_path / "ref"), ref_exe, max_instances=3)
    assert r1.unfinished == [] and r2.unfinished == [] and r1.summary()["failed"] == 0
    for a, b in zip(ours, refs):
        for step in (1, 2, 3):
            for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
                ia, va = P.read_be_file(os.path.join(a, "be.out", "be.%d" % step, key))
                ib, vb = P.read_be_file(os.path.join(b, "be.out", "be.%d" % step, key))
                assert np.array_equal(ia, ib)
                if np.linalg.norm(vb) == 0:
                    assert np.abs(va).max() < 1e-12
                else:
                    assert R.rel_l2(va, vb) < 1e-6, (a, step, key)
