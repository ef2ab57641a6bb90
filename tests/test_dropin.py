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
                    symmetry=spec.get("symmetry"), test_point_source=spec.get("test_point_source", False),
                    interior=bool(spec.get("interior")), curves=[(c, [tuple(p) for p in pts]) for c, pts in
                                                                 spec.get("curves", [])],
                    point_source_curves=[tuple(c) for c in spec.get("point_source_curves", [])],
                    plane_wave_curves=[tuple(c) for c in spec.get("plane_wave_curves", [])])
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
DROPIN_TOL = {"fmm_sl_point": 1e-5, "fmm_ml_head": 1e-5, "sym_fmm_sl": 2e-3, "sym_fmm_ml": 2e-3, "curves_fmm": 1e-5,
              "interior_fmm": 1e-5, **{"scan%d_%s" % (d, m): 1e-5 for d in (1, 2, 3) for m in ("sl", "ml")}}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["soft_pw", "mixed_pres", "quad", "gauss", "velo_patch", "fmm_sl_point",
                                  "fmm_ml_head", "sym_half", "sym_quarter_anti", "sym_octant", "sym_fmm_sl",
                                  "sym_fmm_ml", "test_source_1pw", "test_source_2pw", "curves", "curves_fmm",
                                  "interior", "interior_fmm"] +
                         ["scan%d_%s" % (d, m) for d in (1, 2, 3) for m in ("tbem", "sl", "ml")])
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


# ---- text layout of be.out and NC.out against the reference's SHTF project ------------------
# tests/golden/shtf_text.npz holds the reference's own source_1 input (NC.inp, evaluation grid
# files) and output text (NC1-20.out; first and last lines of be.out/be.1 and be.12); the
# object mesh is rebuilt from tests/golden/shtf.npz.  be.out data lines are printed as
# NC_WriteResultsObjectMesh / NC_WriteResultsEvaluationGrid print them
# (src/NC_PostProcessing.cpp:447-546, 688-749):
BE_FORMATS = {"pBoundary": "%5d % E % E", "vBoundary": "%5d % E % E", "pEvalGrid": "%5d % E % E",
              "vEvalGrid": "%5d %E %E %E %E %E %E"}
_NUM = re.compile(r"[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?")


def _be_line_ok(line, fmt):
    w = line.split()
    try:
        vals = (int(w[0]),) + tuple(float(v) for v in w[1:])
        return fmt % vals == line
    except (ValueError, TypeError, IndexError):
        return False


def log_skeleton(text):
    """NC.out with every number replaced by '#', runs of blanks by one, and the wall-clock
    stamps cut: the text a reader of the log relies on, independent of the values."""
    out = []
    for line in text.splitlines():
        if line.startswith(("Start time:", "End time:")):
            line = line.split(":")[0]
        out.append(re.sub(r" +", " ", _NUM.sub("#", line)).rstrip())
    return out


def report_fields(text):
    """Per step what mesh2hrtf's project report reads from NC*.out
    (mesh2hrtf/Output2HRTF/write_output_report.py:161-240): frequency, the iteration-limit
    flag, iteration count, relative error and the presence of the four time lines."""
    out = {}
    for blk in text.split(">> S T E P   N U M B E R   A N D   F R E Q U E N C Y <<")[1:]:
        step = int(re.search(r"Step (\d+),", blk).group(1))
        f = re.search(r"Frequency = ([\d.eE+-]+) Hz", blk)
        it = re.search(r"number of iterations = (\d+),", blk)
        err = re.search(r"relative error = (.+)", blk)
        times = [bool(re.search(p + r"\d+", blk)) for p in (
            r"Assembling the equation system         : ", r"Solving the equation system            : ",
            r"Post processing                        : ", r"Total                                  : ")]
        out[step] = dict(freq=float(f.group(1)) if f else None, maxit="Maximum number of iterations is reached!" in blk,
                         its=int(it.group(1)) if it else None, err=float(err.group(1)) if err else None, times=times)
    return out


def shtf_project(root):
    """The reference's SHTF project folder, source_1, under ``root``; returns its path."""
    z = np.load(os.path.join(R.GOLDEN, "shtf.npz"))
    t = np.load(os.path.join(R.GOLDEN, "shtf_text.npz"))
    om = os.path.join(root, "ObjectMeshes", "Reference")
    os.makedirs(om)
    with open(os.path.join(om, "Nodes.txt"), "w") as fh:
        fh.write("%d\n" % len(z["nodes"]))
        for i, p in zip(z["node_ids"], z["nodes"]):
            fh.write("%d %.6f %.6f %.6f\n" % (i, p[0], p[1], p[2]))
    with open(os.path.join(om, "Elements.txt"), "w") as fh:
        fh.write("%d\n" % len(z["elems"]))
        for i, e in zip(z["elem_ids"], z["elems"]):
            a, b, c = z["node_ids"][e]
            fh.write("%d %d %d %d 0 0 0\n" % (i, a, b, c))
    grid = os.path.join(root, "EvaluationGrids", "FourPointHorPlane_r100cm")
    os.makedirs(grid)
    with open(os.path.join(grid, "Nodes.txt"), "w") as fh:
        fh.write(str(t["grid_nodes"]))
    with open(os.path.join(grid, "Elements.txt"), "w") as fh:
        fh.write(str(t["grid_elements"]))
    src = os.path.join(root, "NumCalc", "source_1")
    os.makedirs(src)
    with open(os.path.join(src, "NC.inp"), "w") as fh:
        fh.write(str(t["nc_inp"]))
    return src


def test_be_formats_reproduce_the_reference_text():
    """The printf layouts above print every stored reference be.out data line back to the
    same text, so the layout check of the drop-in's files pins them to the reference's."""
    t = np.load(os.path.join(R.GOLDEN, "shtf_text.npz"))
    checked = 0
    for s in (1, 12):
        for key, fmt in BE_FORMATS.items():
            head = str(t["be%d_%s_head" % (s, key)]).split("\n")
            tail = [l for l in str(t["be%d_%s_tail" % (s, key)]).split("\n") if l]
            assert head[0] == "Mesh2HRTF 1.0.0"
            for line in [l for l in head[3:] if l] + tail:
                assert _be_line_ok(line, fmt), (s, key, line)
                checked += 1
    assert checked == 2 * (2 * (10 + 2) + 2 * 6)   # boundary files: 10 + 2 lines; grid files: all 4 + 2


def test_log_skeleton_masks_values_only():
    ref = str(np.load(os.path.join(R.GOLDEN, "shtf_text.npz"))["nc_out"])
    sk = log_skeleton(ref)
    assert "Step #, Frequency = # Hz" in sk and "CGS solver: number of iterations = #, relative error = #" in sk
    assert "End time" in sk and "Single level fast multipole BEM" in sk
    f = report_fields(ref)
    assert sorted(f) == list(range(1, 21)) and f[1]["its"] == 29 and f[12]["freq"] == 1200.0
    assert all(all(v["times"]) for v in f.values())


@pytest.mark.gpu
def test_dropin_shtf_text_layout(tmp_path):
    """The drop-in on the reference's SHTF project, steps 1-20 in one process as the
    reference ran them: NC1-20.out has the reference's text line for line with the values
    masked (input echo, frequency table, per-step method, cluster and expansion tables,
    nonzeros, solver line, time statistics, end stamp), the same report fields (frequency,
    iterations, no iteration limit, relative error below 1e-9), and be.out/be.1 and be.12
    have the reference's header lines, line counts, element / node ids and printf layout."""
    import difflib
    _need_binary()
    t = np.load(os.path.join(R.GOLDEN, "shtf_text.npz"))
    src = shtf_project(str(tmp_path))
    print("drop-in SHTF steps 1-20", flush=True)
    r = subprocess.run([NUMCALC, "-istart", "1", "-iend", "20"], cwd=src, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    ours_text = open(os.path.join(src, "NC1-20.out")).read()
    ref_text = str(t["nc_out"])
    ours, ref = log_skeleton(ours_text), log_skeleton(ref_text)
    diff = list(difflib.unified_diff(ref, ours, "reference", "drop-in", lineterm="", n=1))
    assert not diff, "\n".join(diff[:80])
    fo, fr = report_fields(ours_text), report_fields(ref_text)
    assert sorted(fo) == sorted(fr)
    for s in fr:
        assert fo[s]["freq"] == fr[s]["freq"] and fo[s]["its"] == fr[s]["its"], (s, fo[s], fr[s])
        assert not fo[s]["maxit"] and all(fo[s]["times"]) and fo[s]["err"] < 1e-9
    for s in (1, 12):
        for key, fmt in BE_FORMATS.items():
            lines = open(os.path.join(src, "be.out", "be.%d" % s, key)).read().split("\n")
            head = str(t["be%d_%s_head" % (s, key)]).split("\n")
            tail = str(t["be%d_%s_tail" % (s, key)]).split("\n")
            assert len(lines) == int(t["be%d_%s_lines" % (s, key)]), (s, key)
            assert lines[:3] == head[:3], (s, key, lines[:3])
            assert [l.split()[0] for l in lines[3:len(head)] if l] == [l.split()[0] for l in head[3:] if l]
            assert [l.split()[0] for l in lines[-3:] if l] == [l.split()[0] for l in tail if l]
            bad = [l for l in lines[3:] if l and not _be_line_ok(l, fmt)]
            assert not bad, (s, key, bad[:3])
