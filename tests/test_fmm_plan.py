"""The FMM plan (host, no device) against the figures reference NumCalc logs.

tests/golden/shtf.npz holds the reference's own SHTF project (2412-element sphere) and the
NC*.out figures of every step: method (SL-FMM 1-11, direct ML-FMM 12-60), cluster counts
and radii per level, expansion lengths and sphere points, and the nonzeros of the near
field, T, D and S matrices (src/NC_Main.cpp:1051-1153, src/NC_Addresses.cpp:1476-1654).
"""
import json

import numpy as np
import pytest

import ncref as R
from mesh2hrtf_amd import ncg


@pytest.fixture(scope="module")
def shtf():
    z = np.load(R.GOLDEN + "/shtf.npz")
    return {k: z[k] for k in z.files}


def test_plan_matches_numcalc_logs_every_step(shtf):
    figs = json.loads(str(shtf["figures"]))["1"]
    xyz = shtf["nodes"][shtf["elems"]]
    nv = np.full(len(xyz), 3)
    last = None
    for step in range(1, 61):
        f = 100.0 * step
        k = 2 * np.pi * f / 343.0
        info = ncg.fmm_plan(nv, xyz, k, shtf["eval_nodes"], node_index=shtf["elems"], method_bem=4)
        ref = figs[str(step)]
        if "method" in ref:
            last = ref
        assert info["method"] == last["method"], step
        ncl = [c[1] for c in last["clusters"]]
        assert info["clusters"] == ncl
        np.testing.assert_allclose(info["radius_max"], [c[2] for c in last["clusters"]], rtol=1e-5)
        np.testing.assert_allclose(info["radius_min"], [c[3] for c in last["clusters"]], rtol=1e-5)
        assert info["expansion_terms"] == [e[1] for e in ref["expansion"]]
        assert info["points_sphere"] == [e[4] for e in ref["expansion"]]
        assert info["nnz_near"] == ref["near"]
        assert sum(info["nnz_t"]) == ref["t"]
        assert sum(info["nnz_d"]) == ref["d"]
        assert sum(info["nnz_s"]) == ref["s"]
        if "d_lev" in ref:
            assert info["nnz_d"] == ref["d_lev"]
        assert info["eval_clusters"] == 4


def test_plan_method_thresholds(shtf):
    """Dense below lambda/D > 80, SL for methodBEM 1, dense for methodBEM 0."""
    xyz = shtf["nodes"][shtf["elems"]]
    nv = np.full(len(xyz), 3)
    low = 2 * np.pi * 5.0 / 343.0          # 5 Hz: lambda/D ~ 390
    assert ncg.fmm_plan(nv, xyz, low, method_bem=4)["method"] == 0
    k = 2 * np.pi * 4000.0 / 343.0
    assert ncg.fmm_plan(nv, xyz, k, method_bem=1)["method"] == 1
    assert ncg.fmm_plan(nv, xyz, k, method_bem=0)["method"] == 0
