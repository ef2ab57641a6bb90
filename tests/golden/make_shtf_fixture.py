"""Fixture from the reference's SHTF test project (tests/resources/SHTF of the reference).

The project is a 2412-element sphere (r ~ 0.0875 m) with a reciprocal velocity source on
element 2410 (source_1) or 2411 (source_2) and 4 evaluation points at 1 m, 100..6000 Hz;
NumCalc solves steps 1-11 with the single-level FMM and 12-60 with the 2-level direct
ML-FMM.  This script copies the data the reference's tests hold (mesh, be.out values of
selected steps, the cluster/expansion/nonzero figures of NC*.out) into
tests/golden/shtf.npz, and the text of source_1's input and outputs (NC.inp, the evaluation
grid files, NC1-20.out, the first and last lines of be.out/be.1 and be.12) into
tests/golden/shtf_text.npz for the layout tests.  Run from the repository root, in a container that has the
reference checkout.
"""
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mesh2hrtf_amd import project as P  # noqa: E402

SRC = "/root/reference/tests/resources/SHTF"
STEPS = [1, 5, 11, 12, 30, 41, 60]


def read_table(path, cols):
    with open(path) as fh:
        lines = fh.read().split("\n")
    n = int(lines[0].split()[0])
    rows = [l.split() for l in lines[1:1 + n]]
    return rows


def log_figures(text):
    """Per step: method, clusters/radii per level, L and P per level, nonzeros."""
    out = {}
    for blk in text.split(">> S T E P   N U M B E R   A N D   F R E Q U E N C Y <<")[1:]:
        m = re.search(r"Step (\d+),", blk)
        step = int(m.group(1))
        d = {}
        if "Single level fast multipole BEM" in blk:
            d["method"] = 1
        elif "Direct multilevel fast multipole BEM" in blk:
            d["method"] = 3
        cl = re.findall(r"^\s+(\d+)\s+(\d+)\s+([\d.e+-]+)\s+([\d.e+-]+)\s*$", blk.split("Informations about the FMM")[0],
                        re.M)
        if cl:
            d["clusters"] = [[int(a), int(b), float(c), float(e)] for a, b, c, e in cl]
        ex = re.findall(r"^\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s+(\d+)\s*$", blk, re.M)
        if ex:
            d["expansion"] = [[int(v) for v in row] for row in ex]
        for key, pat in (("near", r"Nonzeros of the near field matrix \.+ = (\d+)"),
                         ("t", r"Nonzeros of the T-matrix \.+ = (\d+)"),
                         ("d", r"Nonzeros of the D-matri(?:x|ces) \.+ = (\d+)"),
                         ("s", r"Nonzeros of the S-matrix \.+ = (\d+)")):
            mm = re.search(pat, blk)
            if mm:
                d[key] = int(mm.group(1))
        for key, pat in (("t_lev", r"Nonzeros of T_lev_(\d+) \.+ = (\d+)"), ("d_lev", r"Nonzeros of D_lev_(\d+) \.+ = (\d+)")):
            mm = re.findall(pat, blk)
            if mm:
                d[key] = [int(v) for _, v in mm]
        it = re.search(r"CGS solver: number of iterations = (\d+)", blk)
        if it:
            d["cgs_iterations"] = int(it.group(1))
        out[step] = d
    return out


def main():
    nodes = read_table(os.path.join(SRC, "ObjectMeshes/Reference/Nodes.txt"), 4)
    elems = read_table(os.path.join(SRC, "ObjectMeshes/Reference/Elements.txt"), 7)
    enodes = read_table(os.path.join(SRC, "EvaluationGrids/FourPointHorPlane_r100cm/Nodes.txt"), 4)
    node_ids = np.array([int(r[0]) for r in nodes])
    xyz = np.array([[float(v) for v in r[1:4]] for r in nodes])
    idx = {i: k for k, i in enumerate(node_ids)}
    conn = np.array([[idx[int(v)] for v in r[1:4]] for r in elems])
    elem_ids = np.array([int(r[0]) for r in elems])
    exyz = np.array([[float(v) for v in r[1:4]] for r in enodes])
    eids = np.array([int(r[0]) for r in enodes])
    freqs = [100.0 * s for s in range(1, 61)]
    data = {}
    figures = {}
    for src in (1, 2):
        d = os.path.join(SRC, "NumCalc", "source_%d" % src)
        for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
            vals = []
            for s in STEPS:
                ids, v = P.read_be_file(os.path.join(d, "be.out", "be.%d" % s, key))
                vals.append(v)
            data["s%d_%s" % (src, key)] = np.asarray(vals)
        fig = {}
        for f in sorted(os.listdir(d)):
            if f.startswith("NC") and f.endswith(".out"):
                fig.update(log_figures(open(os.path.join(d, f)).read()))
        figures[src] = {s: fig[s] for s in fig}
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "shtf.npz"),
                        nodes=xyz, node_ids=node_ids, elems=conn, elem_ids=elem_ids, eval_nodes=exyz, eval_ids=eids,
                        steps=np.asarray(STEPS), freqs=np.asarray(freqs),
                        figures=np.array(json.dumps(figures)), **data)
    print("wrote tests/golden/shtf.npz", {k: v.shape for k, v in data.items()})


TEXT_STEPS = [1, 12]
HEAD_LINES = 13


def text_fixture():
    d = os.path.join(SRC, "NumCalc", "source_1")
    grid = os.path.join(SRC, "EvaluationGrids", "FourPointHorPlane_r100cm")
    out = dict(nc_inp=open(os.path.join(d, "NC.inp")).read(),
               grid_nodes=open(os.path.join(grid, "Nodes.txt")).read(),
               grid_elements=open(os.path.join(grid, "Elements.txt")).read(),
               nc_out=open(os.path.join(d, "NC1-20.out")).read())
    for s in TEXT_STEPS:
        for key in ("pBoundary", "vBoundary", "pEvalGrid", "vEvalGrid"):
            lines = open(os.path.join(d, "be.out", "be.%d" % s, key)).read().split("\n")
            out["be%d_%s_head" % (s, key)] = "\n".join(lines[:HEAD_LINES])
            out["be%d_%s_tail" % (s, key)] = "\n".join(lines[-3:])
            out["be%d_%s_lines" % (s, key)] = len(lines)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "shtf_text.npz"),
                        **{k: np.array(v) for k, v in out.items()})
    print("wrote tests/golden/shtf_text.npz")


if __name__ == "__main__":
    main()
    text_fixture()
