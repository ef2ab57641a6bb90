"""Write the golden fixtures of tests/golden/*.npz.

Each case is a deterministic synthetic NumCalc project (mesh2hrtf_amd.project) run
through the reference NumCalc compiled from its own sources by ``make -C oracle ref``
(oracle/_ref/NumCalc).  The fixture stores the inputs needed to rebuild the problem
(mesh, evaluation points, BC/source spec as JSON) and the reference be.out values.

Run from the repository root:  python tests/golden/make_golden.py [case ...]
"""
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from mesh2hrtf_amd import project as P  # noqa: E402
from ncref import c2_project, mesh_digest, nc_tables  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "NumCalc")
OUT = os.path.dirname(os.path.abspath(__file__))


def cube_sphere(n, radius):
    """Quad mesh: every cube face split n x n, projected on a sphere (outward order)."""
    pts, index, quads = [], {}, []

    def pid(p):
        key = tuple(np.round(p, 12))
        if key not in index:
            index[key] = len(pts)
            pts.append(p)
        return index[key]

    g = np.linspace(-1.0, 1.0, n + 1)
    for axis in range(3):
        for sign in (-1.0, 1.0):
            ids = {}
            for a in range(n + 1):
                for b in range(n + 1):
                    p = np.zeros(3)
                    p[axis] = sign
                    p[(axis + 1) % 3] = g[a]
                    p[(axis + 2) % 3] = g[b]
                    ids[a, b] = pid(p / np.linalg.norm(p) * radius)
            for a in range(n):
                for b in range(n):
                    q = [ids[a, b], ids[a + 1, b], ids[a + 1, b + 1], ids[a, b + 1]]
                    quads.append(q)
    nodes = np.asarray(pts)
    quads = np.asarray(quads)
    # NumCalc quad normal = dx/ds x dx/dt with s along 1->0 ... orient by centroid test
    out = []
    for q in quads:
        x = nodes[q]
        c = x.mean(axis=0)
        xs = 0.25 * ((x[0] - x[1]) + (x[3] - x[2]))
        xt = 0.25 * ((x[0] - x[3]) + (x[1] - x[2]))
        if np.dot(np.cross(xs, xt), c) < 0:
            q = q[::-1]
        out.append(q)
    return P.Mesh(nodes, np.asarray(out))


def cut_mesh(mesh, planes, shift=(0.0, 0.0, 0.0), triangles=False):
    """The part of ``mesh`` (moved by ``shift``) on the + side of every plane (axis,
    coordinate): elements whose centroid lies beyond each plane.  Quads are split into two
    triangles when ``triangles``."""
    nodes = mesh.nodes + np.asarray(shift)
    cen = nodes[mesh.elems].mean(axis=1)
    keep = np.ones(len(mesh.elems), bool)
    for axis, c in planes:
        keep &= cen[:, axis] > c
    el = mesh.elems[keep]
    if triangles:
        el = np.concatenate([el[:, [0, 1, 2]], el[:, [0, 2, 3]]])
    used = np.unique(el)
    remap = -np.ones(len(nodes), np.int64)
    remap[used] = np.arange(len(used))
    return P.Mesh(nodes[used], remap[el])


def cases():
    ico2 = P.icosphere(2, 0.1)
    ne = len(ico2.elems)
    C = {}
    C["c1"] = dict(prj=P.config1_project([100.0, 1000.0, 2000.0]))
    C["soft_pw"] = dict(prj=P.Project(
        mesh=ico2, freqs=[500.0, 2500.0], eval_grid=P.fibonacci_grid(100, 0.5),
        boundary=[P.Boundary(0, ne - 1, "PRES", 0.0, 0.0)],
        plane_waves=[(0.0, -1.0, 0.0, 1.0, 0.0)], title="soft sphere plane wave"))
    C["adm_ps"] = dict(prj=P.Project(
        mesh=ico2, freqs=[800.0, 3000.0], eval_grid=P.fibonacci_grid(100, 0.6),
        boundary=[P.Boundary(0, ne - 1, "ADMI", 0.5, 0.2)],
        point_sources=[(0.05, 0.0, 0.25, 1.0, 0.5)], title="admittance sphere point source"))
    C["velo_patch"] = dict(prj=P.Project(
        mesh=ico2, freqs=[1000.0, 4000.0], eval_grid=P.fibonacci_grid(60, 1.0),
        boundary=[P.Boundary(0, ne - 1, "VELO", 0.0, 0.0), P.Boundary(17, 19, "VELO", 0.1, 0.0)],
        title="rigid sphere with vibrating patch"))
    C["mixed_pres"] = dict(prj=P.Project(
        mesh=ico2, freqs=[1500.0], eval_grid=P.fibonacci_grid(60, 0.8),
        boundary=[P.Boundary(0, 99, "PRES", 0.3, -0.1), P.Boundary(200, 210, "VELO", 0.05, 0.02),
                  P.Boundary(250, 260, "IMPE", 2.0, 1.0)],
        point_sources=[(0.0, 0.2, 0.0, 1.0, 0.0)], title="mixed BCs"))
    C["quad"] = dict(prj=P.Project(
        mesh=cube_sphere(6, 0.1), freqs=[700.0, 2100.0], eval_grid=P.fibonacci_grid(80, 0.7),
        point_sources=[(0.0, 0.0, 0.3, 1.0, 0.0)], title="quad cube-sphere"))
    head = P.head_like_mesh(freq=8, seed=0)
    src = P.ear_sources(head)[0]
    C["head8"] = dict(prj=P.Project(
        mesh=head, freqs=[2000.0, 6000.0], eval_grid=P.fibonacci_grid(150, 1.2),
        point_sources=[(src[0], src[1], src[2], 1.0, 0.0)], title="head-like mesh nu=8"))
    for pc in (1, 2):
        C["c1_precond%d" % pc] = dict(prj=P.Project(
            mesh=P.icosphere(3, 0.1), freqs=[1000.0], eval_grid=P.fibonacci_grid(2030, 1.2),
            point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)], precond=pc))
    C["gauss"] = dict(prj=P.Project(
        mesh=ico2, freqs=[1200.0], eval_grid=P.fibonacci_grid(50, 0.9), method_solver=4,
        point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)], title="direct solver"))
    # fast multipole projects (NC.inp methodBEM 1 = SL-FMM, 4 = ML-FMM)
    ico3 = P.icosphere(3, 0.1)
    n3 = len(ico3.elems)
    C["fmm_sl_point"] = dict(prj=P.Project(
        mesh=ico3, freqs=[1000.0, 3000.0], eval_grid=P.fibonacci_grid(100, 0.5), method_bem=1,
        point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)], title="SL-FMM rigid sphere point source"))
    C["fmm_sl_adm"] = dict(prj=P.Project(
        mesh=ico3, freqs=[2000.0], eval_grid=P.fibonacci_grid(60, 0.8), method_bem=1,
        boundary=[P.Boundary(0, n3 - 1, "ADMI", 0.3, 0.1)],
        plane_waves=[(0.0, 0.6, 0.8, 1.0, 0.0)], point_sources=[(0.1, 0.1, 0.15, 0.5, 0.5)],
        title="SL-FMM admittance sphere"))
    C["fmm_sl_soft"] = dict(prj=P.Project(
        mesh=ico3, freqs=[1500.0], eval_grid=P.fibonacci_grid(60, 0.7), method_bem=1,
        boundary=[P.Boundary(0, n3 - 1, "PRES", 0.0, 0.0)], plane_waves=[(1.0, 0.0, 0.0, 1.0, 0.0)],
        title="SL-FMM soft sphere plane wave"))
    head12 = P.head_like_mesh(freq=12, seed=0)
    src12 = P.ear_sources(head12)[0]
    C["fmm_ml_head"] = dict(prj=P.Project(
        mesh=head12, freqs=[3000.0, 8000.0], eval_grid=P.fibonacci_grid(120, 1.2), method_bem=4,
        point_sources=[(src12[0], src12[1], src12[2], 1.0, 0.0)], title="ML-FMM head-like mesh nu=12"))
    n12 = len(head12.elems)
    C["fmm_ml_patch"] = dict(prj=P.Project(
        mesh=head12, freqs=[5000.0], eval_grid=P.fibonacci_grid(80, 1.0), method_bem=4,
        boundary=[P.Boundary(0, n12 - 1, "VELO", 0.0, 0.0), P.Boundary(100, 103, "VELO", 0.1, 0.0)],
        title="ML-FMM head-like mesh vibrating patch"))
    # symmetry planes (NC.inp SYMMETRY block): mesh parts on one side of the planes
    cs6, cs8 = cube_sphere(6, 0.1), cube_sphere(8, 0.1)
    half = cut_mesh(cs6, [(2, 0.03)], shift=(0.0, 0.0, 0.03))
    nh = len(half.elems)
    C["sym_half"] = dict(prj=P.Project(
        mesh=half, freqs=[800.0, 2500.0], eval_grid=P.fibonacci_grid(60, 0.6), symmetry=((0, 0, 1), (0.0, 0.0, 0.03)),
        boundary=[P.Boundary(0, nh - 1, "VELO", 0.0, 0.0), P.Boundary(5, 9, "VELO", 0.1, 0.02),
                  P.Boundary(40, 60, "ADMI", 0.4, 0.1)],
        point_sources=[(0.02, 0.01, 0.3, 1.0, 0.0)], plane_waves=[(0.6, 0.0, 0.8, 0.5, 0.2)],
        title="half cube-sphere, symmetric plane z=0.03"))
    quarter = cut_mesh(cs8, [(0, 0.0), (1, 0.0)], triangles=True)
    nq = len(quarter.elems)
    C["sym_quarter_anti"] = dict(prj=P.Project(
        mesh=quarter, freqs=[1800.0], eval_grid=P.fibonacci_grid(60, 0.7), symmetry=((2, 1, 0), (0.0, 0.0, 0.0)),
        boundary=[P.Boundary(0, nq - 1, "VELO", 0.0, 0.0), P.Boundary(10, 14, "VELO", 0.05, -0.02)],
        point_sources=[(0.05, 0.12, 0.1, 1.0, 0.3)], plane_waves=[(0.0, 0.6, -0.8, 1.0, 0.0)],
        title="quarter sphere (triangles), antisymmetric x=0, symmetric y=0"))
    octant = cut_mesh(cs8, [(0, 0.0), (1, 0.0), (2, 0.0)])
    no = len(octant.elems)
    C["sym_octant"] = dict(prj=P.Project(
        mesh=octant, freqs=[2200.0], eval_grid=P.fibonacci_grid(50, 0.8), symmetry=((1, 4, 3), (0.0, 0.0, 0.0)),
        boundary=[P.Boundary(0, no - 1, "VELO", 0.0, 0.0), P.Boundary(3, 4, "VELO", 0.1, 0.0),
                  P.Boundary(8, 12, "IMPE", 3.0, 1.0)],
        point_sources=[(0.1, 0.2, 0.05, 1.0, 0.0)], title="octant, planes x sym, y antisym (4), z sym (3)"))
    cs12, cs16 = cube_sphere(12, 0.1), cube_sphere(16, 0.1)
    C["sym_fmm_sl"] = dict(prj=P.Project(
        mesh=cut_mesh(cs12, [(2, 0.0)]), freqs=[2000.0], eval_grid=P.fibonacci_grid(60, 0.6), method_bem=1,
        symmetry=((0, 0, 1), (0.0, 0.0, 0.0)), point_sources=[(0.03, 0.0, 0.2, 1.0, 0.0)],
        title="SL-FMM half sphere, symmetric z=0"))
    ml = cut_mesh(cs16, [(1, 0.0)])
    nm = len(ml.elems)
    C["sym_fmm_ml"] = dict(prj=P.Project(
        mesh=ml, freqs=[4000.0], eval_grid=P.fibonacci_grid(60, 0.9), method_bem=4,
        symmetry=((0, 2, 0), (0.0, 0.0, 0.0)),
        boundary=[P.Boundary(0, nm - 1, "VELO", 0.0, 0.0), P.Boundary(30, 35, "VELO", 0.1, 0.0)],
        plane_waves=[(0.0, 0.6, 0.8, 1.0, 0.0)], title="ML-FMM half sphere, antisymmetric y=0"))
    # point-source count -1 (the reference's "test point source"): with one plane wave it
    # reaches the evaluation grid but not the right-hand side; with two it reaches both
    C["test_source_1pw"] = dict(prj=P.Project(
        mesh=ico2, freqs=[900.0], eval_grid=P.fibonacci_grid(40, 0.6), test_point_source=True,
        boundary=[P.Boundary(0, ne - 1, "VELO", 0.0, 0.0), P.Boundary(30, 33, "VELO", 0.1, 0.0)],
        plane_waves=[(0.0, 0.0, 1.0, 1.0, 0.0)], title="test point source, one plane wave"))
    C["test_source_2pw"] = dict(prj=P.Project(
        mesh=ico2, freqs=[900.0], eval_grid=P.fibonacci_grid(40, 0.6), test_point_source=True,
        plane_waves=[(0.0, 0.0, 1.0, 1.0, 0.0), (0.6, 0.8, 0.0, 0.5, 0.5)], title="test point source, two plane waves"))
    # frequency CURVES (src/NC_Input.cpp:1090-1145, src/NC_EquationSystem.cpp:1815-2009):
    # steps below, inside and above the curves' points; admittance and source amplitudes
    # scaled per part, a velocity line carrying curve ids (which NumCalc ignores)
    curves = [(1, [(600.0, 0.5), (1500.0, 1.2), (2500.0, 0.8)]),
              (2, [(300.0, 2.0), (2200.0, 0.25)]),
              (3, [(800.0, 1.0), (1200.0, 3.0), (1800.0, -1.0), (2600.0, 0.5)]),
              (4, [(1000.0, 0.7)])]
    C["curves"] = dict(prj=P.Project(
        mesh=ico2, freqs=[400.0, 1000.0, 2000.0, 3000.0], eval_grid=P.fibonacci_grid(60, 0.6),
        boundary=[P.Boundary(0, ne - 1, "ADMI", 0.5, 0.2, 1, 2), P.Boundary(30, 33, "VELO", 0.1, 0.0, 3, 4)],
        point_sources=[(0.05, 0.0, 0.25, 1.0, 0.5)], point_source_curves=[(3, 2)],
        plane_waves=[(0.0, 0.6, 0.8, 0.5, 0.2)], plane_wave_curves=[(4, 1)], curves=curves,
        title="frequency curves on admittance and sources"))
    C["curves_fmm"] = dict(prj=P.Project(
        mesh=ico3, freqs=[700.0, 1400.0, 2400.0], eval_grid=P.fibonacci_grid(60, 0.8), method_bem=1,
        boundary=[P.Boundary(0, n3 - 1, "ADMI", 0.3, 0.1, 2, 1)],
        point_sources=[(0.1, 0.1, 0.15, 0.5, 0.5)], point_source_curves=[(1, 3)], curves=curves,
        title="SL-FMM frequency curves"))
    # interior problems (IFINTE 1: tau = -1, no Burton-Miller term; src/NC_Input.cpp:346,
    # src/NC_EquationSystem.cpp:86): a source and the grid inside an absorbing sphere
    C["interior"] = dict(prj=P.Project(
        mesh=ico2, freqs=[500.0, 1700.0], eval_grid=P.fibonacci_grid(40, 0.05), interior=True,
        boundary=[P.Boundary(0, ne - 1, "ADMI", 0.5, 0.2)], point_sources=[(0.02, 0.01, -0.03, 1.0, 0.0)],
        title="interior problem, absorbing sphere"))
    C["interior_fmm"] = dict(prj=P.Project(
        mesh=ico3, freqs=[1200.0], eval_grid=P.fibonacci_grid(40, 0.06), interior=True, method_bem=1,
        boundary=[P.Boundary(0, n3 - 1, "ADMI", 0.4, 0.1), P.Boundary(10, 14, "VELO", 0.1, 0.0)],
        point_sources=[(0.0, 0.03, 0.02, 1.0, 0.0)], title="SL-FMM interior problem"))
    # ILU scanning degrees (IPRECOND -1..-3: src/NC_Input.cpp:306-320, thresholds
    # src/NC_CommonFunctions.cpp:1366-1404) on the dense, SL- and ML-FMM paths
    for d in (1, 2, 3):
        C["scan%d_tbem" % d] = dict(prj=P.Project(
            mesh=ico3, freqs=[1000.0], eval_grid=P.fibonacci_grid(60, 1.2), precond=-d,
            point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)], title="dense ILU scanning degree %d" % d))
        C["scan%d_sl" % d] = dict(prj=P.Project(
            mesh=ico3, freqs=[2000.0], eval_grid=P.fibonacci_grid(60, 0.8), method_bem=1, precond=-d,
            point_sources=[(0.0, 0.0, 0.2, 1.0, 0.0)], title="SL-FMM ILU scanning degree %d" % d))
        C["scan%d_ml" % d] = dict(prj=P.Project(
            mesh=head12, freqs=[6000.0], eval_grid=P.fibonacci_grid(60, 1.2), method_bem=4, precond=-d,
            point_sources=[(src12[0], src12[1], src12[2], 1.0, 0.0)], title="ML-FMM ILU scanning degree %d" % d))
    # one mesh of the batched uneven dense problems (project.batched_problems: index 3, a
    # perturbed sphere of 1620 triangles) at its first two frequencies, first plane wave
    bm, dirs = P.batched_problems(8)
    mesh3, f3 = bm[3]
    C["c4_mesh3"] = dict(prj=P.Project(
        mesh=mesh3, freqs=[float(f3[0]), float(f3[1])], eval_grid=P.fibonacci_grid(50, 0.5),
        plane_waves=[(float(dirs[0][0]), float(dirs[0][1]), float(dirs[0][2]), 1.0, 0.0)],
        title="batched dense problem 3 (perturbed sphere), plane wave 0"))
    return C


def spec_of(prj):
    return dict(freqs=list(prj.freqs),
                boundary=[(b.lo, b.hi, b.kind, b.re, b.im, b.re_curve, b.im_curve) for b in prj.boundary],
                point_sources=[list(map(float, s)) for s in prj.point_sources],
                plane_waves=[list(map(float, s)) for s in prj.plane_waves],
                method_bem=prj.method_bem, method_solver=prj.method_solver,
                precond=prj.precond, c=prj.c, rho=prj.rho,
                symmetry=None if prj.symmetry is None else [list(prj.symmetry[0]), list(prj.symmetry[1])],
                test_point_source=prj.test_point_source, interior=prj.interior,
                curves=[[cid, [list(p) for p in pts]] for cid, pts in prj.curves],
                point_source_curves=[list(c) for c in prj.point_source_curves],
                plane_wave_curves=[list(c) for c in prj.plane_wave_curves])


def run_case(name, prj):
    with tempfile.TemporaryDirectory() as d:
        P.write_project(d, prj)
        t0 = time.time()
        r = subprocess.run([REF], cwd=d, capture_output=True, text=True)
        wall = time.time() - t0
        if r.returncode != 0:
            raise RuntimeError("%s: NumCalc failed\n%s" % (name, r.stdout[-2000:]))
        out = {}
        nout = open(os.path.join(d, "NC.out")).read()
        iters = [int(l.split("=")[1].split(",")[0]) for l in nout.splitlines()
                 if l.startswith("CGS solver: number of iterations")]
        relerr = [float(l.split("relative error =")[1]) for l in nout.splitlines()
                  if l.startswith("CGS solver: number of iterations")]
        for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
            vals = []
            for s in range(len(prj.freqs)):
                ids, v = P.read_be_file(os.path.join(d, "be.out", "be.%d" % (s + 1), key))
                vals.append(v)
            out[key] = np.asarray(vals)
            out[key + "_ids"] = ids
        out["nc_tables"] = np.array(json.dumps(nc_tables(nout)))
        r = subprocess.run([REF, "-estimate_ram"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("%s: NumCalc -estimate_ram failed\n%s" % (name, r.stdout[-2000:]))
        out["memory_txt"] = np.array(open(os.path.join(d, "Memory.txt")).read())
    m = prj.mesh
    g = prj.eval_grid
    np.savez_compressed(
        os.path.join(OUT, "case_%s.npz" % name),
        nodes=m.nodes, elems=m.elems, elem_ids=m.elem_ids,
        eval_nodes=g.nodes if g is not None else np.zeros((0, 3)),
        spec=np.array(json.dumps(spec_of(prj))),
        cgs_iterations=np.asarray(iters, dtype=np.int64), cgs_relerr=np.asarray(relerr),
        ref_wall_s=np.float64(wall), **out)
    print("%-12s N=%d M=%d steps=%d wall=%.1fs iters=%s" % (
        name, len(m.elems), 0 if g is None else len(g.nodes), len(prj.freqs), wall, iters))


# ---- the head-like configuration at its full size ---------------------------------------
# ncref.c2_project(): 20480-element seeded head-like mesh, one ear source, the 1730-point
# grid, four frequencies.  The mesh is not stored: the generator is deterministic and the
# fixture keeps a digest of it.  Values are stored as complex64 (the files print 7 digits).


def run_c2_case(name="c2_head"):
    prj = c2_project()
    mesh, grid = prj.mesh, prj.eval_grid
    C2_FREQS = prj.freqs
    with tempfile.TemporaryDirectory() as d:
        P.write_project(d, prj)
        # one reference process per frequency step, side by side (as manage_numcalc runs them)
        t0 = time.time()
        procs = [subprocess.Popen([REF, "-istart", str(s + 1), "-iend", str(s + 1)], cwd=d,
                                  stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                 for s in range(len(C2_FREQS))]
        codes = [p.wait() for p in procs]
        wall = time.time() - t0
        if any(codes):
            raise RuntimeError("%s: NumCalc failed %s" % (name, codes))
        iters, relerr, tables = [], [], []
        for s in range(len(C2_FREQS)):
            nout = open(os.path.join(d, "NC%d-%d.out" % (s + 1, s + 1))).read()
            its = [l for l in nout.splitlines() if l.startswith("CGS solver: number of iterations")]
            iters.append(int(its[0].split("=")[1].split(",")[0]))
            relerr.append(float(its[0].split("relative error =")[1]))
            tables.append(nc_tables(nout))
        out = {}
        for key in ("pBoundary", "pEvalGrid", "vEvalGrid"):
            vals = []
            for s in range(len(C2_FREQS)):
                ids, v = P.read_be_file(os.path.join(d, "be.out", "be.%d" % (s + 1), key))
                vals.append(v)
            out[key] = np.asarray(vals).astype(np.complex64)
            out[key + "_ids"] = ids
    np.savez_compressed(
        os.path.join(OUT, "case_%s.npz" % name), mesh_sha256=np.array(mesh_digest(mesh)),
        eval_nodes=grid.nodes, spec=np.array(json.dumps(spec_of(prj))),
        cgs_iterations=np.asarray(iters, dtype=np.int64), cgs_relerr=np.asarray(relerr),
        nc_tables=np.array(json.dumps(tables)), ref_wall_s=np.float64(wall), **out)
    print("%-12s N=%d M=%d steps=%d wall=%.1fs iters=%s" % (name, len(mesh.elems), len(grid.nodes),
                                                             len(C2_FREQS), wall, iters))


def main(argv):
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    C = cases()
    for name in (argv or list(C) + ["c2_head"]):
        if name == "c2_head":
            run_c2_case(name)
        else:
            run_case(name, C[name]["prj"])


if __name__ == "__main__":
    main(sys.argv[1:])
