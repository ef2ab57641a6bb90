"""GPU parity: libncg (HIP, gfx950) against the oracle and reference NumCalc's be.out.

Tolerances (FP64 path): operator and right-hand side entrywise within 1e-9 of the oracle
(same quadrature decisions, different rounding order); solved unknowns within 1e-8 of the
oracle (CGS stops at a 1e-9 residual); outputs within 1e-6 of reference NumCalc (its
be.out prints 7 significant digits).  Mixed precision: within 1e-4 of the FP64 path
(the metric's bar is 1e-3).
"""
import numpy as np
import pytest

import ncref as R

pytestmark = pytest.mark.gpu

SMALL = ["soft_pw", "adm_ps", "velo_patch", "mixed_pres", "quad", "gauss"]


@pytest.fixture(scope="module")
def ncg():
    from mesh2hrtf_amd import ncg as N
    if N.device_count() < 1:
        pytest.fail("no HIP device visible")
    return N


@pytest.fixture(scope="module")
def oracle():
    return R.Oracle()


def gpu_problem(N, case, pa):
    return N.Problem(pa["nv"], pa["xyz"], pa["kind"], case["eval_nodes"])


def sources(pa, rhs=0):
    ps = [list(s) + [rhs] for s in pa["ps"]]
    pw = [list(s) + [rhs] for s in pa["pw"]]
    return ps, pw


def assemble(prob, pa, precision=0):
    ps, pw = sources(pa)
    prob.assemble(pa["k"], bc=pa["bc"][:, 0][None, :], adm=pa["adm"], point_sources=ps, plane_waves=pw,
                  precision=precision)


@pytest.mark.parametrize("name", SMALL + ["head8"])
def test_assembly_matches_oracle(ncg, oracle, name):
    case = R.load_case(name)
    f = case["spec"]["freqs"][-1]
    pa = R.problem_arrays(case, f)
    print("case", name, "f", f, "n", pa["n"], flush=True)
    A0, b0 = oracle.assemble(pa)
    prob = gpu_problem(ncg, case, pa)
    assemble(prob, pa)
    A, b = prob.system()
    scale = np.abs(A0).max()
    assert np.abs(A - A0).max() / scale < 1e-9
    assert np.abs(b[0] - b0).max() / np.abs(b0).max() < 1e-9


@pytest.mark.parametrize("name", SMALL + ["head8", "c1"])
def test_solution_and_field_match_reference(ncg, oracle, name):
    case = R.load_case(name)
    prob = None
    for s, f in enumerate(case["spec"]["freqs"]):
        pa = R.problem_arrays(case, f)
        print("case", name, "step", s, "f", f, flush=True)
        if prob is None:
            prob = gpu_problem(ncg, case, pa)
        assemble(prob, pa)
        solver = ncg.LU if case["spec"]["method_solver"] == 4 else ncg.CGS
        st = prob.solve(solver=solver)
        assert st["converged"]
        x = prob.solution()[0]
        pB, vB = R.surface_outputs(pa, x)
        for ours, ref in ((pB, case["pBoundary"][s]), (vB, case["vBoundary"][s])):
            if np.linalg.norm(ref) == 0:
                assert np.abs(ours).max() < 1e-12
            else:
                assert R.rel_l2(ours, ref) < 1e-6
        phi, grad = prob.field()
        assert R.rel_l2(1j * pa["rpfact"] * phi[0], case["pEvalGrid"][s]) < 1e-6
        assert R.rel_l2(grad[0], case["vEvalGrid"][s]) < 1e-6
        # the field kernel against the oracle on the same solved unknowns
        if pa["n"] <= 400:
            phi0, grad0 = oracle.field(pa, x, case["eval_nodes"])
            assert R.rel_l2(phi[0], phi0) < 1e-11
            assert R.rel_l2(grad[0], grad0) < 1e-11


def test_cgs_follows_numcalc_iteration(ncg, oracle):
    """Same algorithm as the oracle's restatement: iteration count and solution agree."""
    case = R.load_case("adm_ps")
    pa = R.problem_arrays(case, 800.0)
    A0, b0 = oracle.assemble(pa)
    x0, it0, err0 = oracle.cgs(A0, b0, precond=2)
    prob = gpu_problem(ncg, case, pa)
    assemble(prob, pa)
    st = prob.solve(solver=ncg.CGS, precond=ncg.NONE)
    assert abs(st["iterations"][0] - it0) <= 1
    assert R.rel_l2(prob.solution()[0], x0) < 1e-8
    x1, it1, _ = oracle.cgs(A0, b0, precond=1)
    assemble(prob, pa)
    st1 = prob.solve(solver=ncg.CGS, precond=ncg.SCALING)
    assert abs(st1["iterations"][0] - it1) <= 1
    assert R.rel_l2(prob.solution()[0], x1) < 1e-8


@pytest.mark.parametrize("name", ["soft_pw", "adm_ps", "velo_patch", "mixed_pres", "quad", "head8", "c1"])
def test_cgs_ilu_matches_reference(ncg, oracle, name):
    """Incomplete-LU CGS (IPRECOND 0): the reference's iteration count and relative
    error at every step, and the oracle's ILU-CGS solution."""
    case = R.load_case(name)
    prob = None
    for s, f in enumerate(case["spec"]["freqs"]):
        pa = R.problem_arrays(case, f)
        print("ILU case", name, "f", f, flush=True)
        if prob is None:
            prob = gpu_problem(ncg, case, pa)
        assemble(prob, pa)
        st = prob.solve(solver=ncg.CGS, precond=ncg.ILU)
        assert st["converged"]
        assert st["iterations"][0] == case["cgs_iterations"][s], (s, st)
        assert abs(st["relative_residual"][0] - case["cgs_relerr"][s]) < 0.2 * case["cgs_relerr"][s]
        assert st["ilu_nonzeros"] >= pa["n"]
        A0, b0 = oracle.assemble(pa)
        x0, it0, _ = oracle.cgs(A0, b0, precond=0)
        assert R.rel_l2(prob.solution()[0], x0) < 1e-8


def test_solvers_agree(ncg):
    case = R.load_case("head8")
    pa = R.problem_arrays(case, 6000.0)
    prob = gpu_problem(ncg, case, pa)
    sols = {}
    for solver in (ncg.CGS, ncg.GMRES, ncg.LU):
        assemble(prob, pa)
        prob.solve(solver=solver, tolerance=1e-10)
        sols[solver] = prob.solution()[0]
    assert R.rel_l2(sols[ncg.GMRES], sols[ncg.LU]) < 1e-8
    assert R.rel_l2(sols[ncg.CGS], sols[ncg.LU]) < 1e-8


def test_multiple_rhs_equal_single_solves(ncg):
    """Load cases sharing one operator give the same answers as separate solves."""
    case = R.load_case("velo_patch")
    pa = R.problem_arrays(case, 4000.0)
    n = pa["n"]
    rng = np.random.default_rng(3)
    bc = np.zeros((3, n), complex)
    bc[0, 17:20] = 0.1
    bc[1, 100:104] = 0.05j
    ps = [[0.0, 0.0, 0.3, 1e-3, 2e-4, 2], [0.1, 0.2, 0.0, 0.0, 1e-3, 1]]
    pw = [[0.0, 0.6, 0.8, 1e-3, 0.0, 2]]
    prob = gpu_problem(ncg, case, pa)
    prob.assemble(pa["k"], bc=bc, point_sources=ps, plane_waves=pw)
    prob.solve(solver=ncg.CGS)
    X = prob.solution()
    phi, grad = prob.field()
    for r in range(3):
        prob.assemble(pa["k"], bc=bc[r:r + 1], point_sources=[s[:5] + [0] for s in ps if s[5] == r],
                      plane_waves=[s[:5] + [0] for s in pw if s[5] == r])
        prob.solve(solver=ncg.CGS)
        assert R.rel_l2(X[r], prob.solution()[0]) < 1e-8
        p1, g1 = prob.field()
        assert R.rel_l2(phi[r], p1[0]) < 1e-8
    del rng


def test_mixed_precision_within_tolerance(ncg):
    case = R.load_case("head8")
    for f in case["spec"]["freqs"]:
        pa = R.problem_arrays(case, f)
        prob = gpu_problem(ncg, case, pa)
        assemble(prob, pa, precision=ncg.FP64)
        prob.solve(solver=ncg.CGS)
        x64 = prob.solution()[0]
        p64, _ = prob.field()
        assemble(prob, pa, precision=ncg.MIXED)
        with pytest.raises(ncg.NcgError):
            prob.solve(solver=ncg.CGS)          # the mixed form is GMRES refinement only
        st = prob.solve(solver=ncg.GMRES, refinement_tolerance=1e-7)
        assert st["converged"]
        x32 = prob.solution()[0]
        p32, _ = prob.field()
        # stated tolerance of the mixed form: rel-L2 <= 1e-3 vs FP64; observed far below
        assert R.rel_l2(x32, x64) < 1e-4
        assert R.rel_l2(p32[0], p64[0]) < 1e-4


def test_dense_block_jacobi(ncg):
    """The served form's block-Jacobi preconditioner on a dense (TBEM) operator: blocks of
    consecutive unknowns; GMRES with it reaches the FP64 CGS solution, in both precisions."""
    case = R.load_case("head8")
    f = case["spec"]["freqs"][-1]
    pa = R.problem_arrays(case, f)
    prob = gpu_problem(ncg, case, pa)
    assemble(prob, pa, precision=ncg.FP64)
    prob.solve(solver=ncg.CGS, tolerance=1e-10)
    x64 = prob.solution()[0]
    st = prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, tolerance=1e-10)
    assert st["converged"]
    assert R.rel_l2(prob.solution()[0], x64) < 1e-8
    assemble(prob, pa, precision=ncg.MIXED)
    st = prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-7)
    assert st["converged"]
    assert R.rel_l2(prob.solution()[0], x64) < 1e-4


def test_dense_many_load_cases_mixed(ncg):
    """16 load cases on a dense operator in the mixed form (the FP32 operator is applied to
    all 16 in one pass on the matrix cores): each solution within 1e-4 of the FP64 solve."""
    case = R.load_case("head8")
    f = case["spec"]["freqs"][-1]
    pa = R.problem_arrays(case, f)
    n = pa["n"]
    rng = np.random.default_rng(5)
    dirs = rng.standard_normal((16, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pw = [list(dirs[r]) + [1e-3 * (1 + 0.1 * r), 1e-4 * r, r] for r in range(16)]
    bc = np.repeat(pa["bc"][:, 0][None, :], 16, axis=0)
    prob = gpu_problem(ncg, case, pa)
    prob.assemble(pa["k"], bc=bc, adm=pa["adm"], plane_waves=pw, precision=ncg.FP64)
    prob.solve(solver=ncg.CGS, tolerance=1e-10)
    x64 = prob.solution()
    for nr in (16, 7):
        prob.assemble(pa["k"], bc=bc[:nr], adm=pa["adm"], plane_waves=pw[:nr], precision=ncg.MIXED)
        st = prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-7)
        print("load cases", nr, "n", n, "iterations", st["iterations"], flush=True)
        assert st["converged"]
        x32 = prob.solution()
        for r in range(nr):
            assert R.rel_l2(x32[r], x64[r]) < 1e-4, r


def test_dense_twenty_load_cases(ncg):
    """20 load cases on a small dense operator (n = 320, less than one 128-row tile of the
    matrix cores in places): two passes (16 + 4) of the matrix-core applies, FP64 and mixed,
    against one-load-case solves."""
    case = R.load_case("soft_pw")
    pa = R.problem_arrays(case, case["spec"]["freqs"][-1])
    rng = np.random.default_rng(11)
    dirs = rng.standard_normal((20, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    pw = [list(dirs[r]) + [1e-3, 2e-4 * r, r] for r in range(20)]
    bc = np.repeat(pa["bc"][:, 0][None, :], 20, axis=0)
    prob = gpu_problem(ncg, case, pa)
    prob.assemble(pa["k"], bc=bc, adm=pa["adm"], plane_waves=pw, precision=ncg.FP64)
    prob.solve(solver=ncg.CGS, tolerance=1e-11)
    X = prob.solution()
    prob.assemble(pa["k"], bc=bc, adm=pa["adm"], plane_waves=pw, precision=ncg.MIXED)
    st = prob.solve(solver=ncg.GMRES, precond=ncg.BLOCK, refinement_tolerance=1e-8)
    assert st["converged"]
    Xm = prob.solution()
    for r in (0, 7, 15, 16, 19):
        prob.assemble(pa["k"], bc=bc[r:r + 1], adm=pa["adm"], plane_waves=[pw[r][:5] + [0]], precision=ncg.FP64)
        prob.solve(solver=ncg.CGS, tolerance=1e-11)
        x1 = prob.solution()[0]
        assert R.rel_l2(X[r], x1) < 1e-8, r
        assert R.rel_l2(Xm[r], x1) < 1e-4, r


def test_apply_matches_system(ncg):
    case = R.load_case("quad")
    pa = R.problem_arrays(case, 2100.0)
    prob = gpu_problem(ncg, case, pa)
    assemble(prob, pa)
    A, _ = prob.system()
    x = np.random.default_rng(0).standard_normal((1, pa["n"])) * (1 + 1j)
    assert R.rel_l2(prob.apply(x)[0], A @ x[0]) < 1e-13


def test_errors_are_reported(ncg):
    case = R.load_case("quad")
    pa = R.problem_arrays(case, 2100.0)
    prob = gpu_problem(ncg, case, pa)
    with pytest.raises(ncg.NcgError):
        prob.solve()                       # before assemble
    with pytest.raises(ncg.NcgError):
        prob.assemble(-1.0)                # invalid wave number
    assemble(prob, pa)
    st = prob.solve(max_iterations=1, allow_noconv=True)
    assert not st["converged"]
