"""Pins of the CPU oracle (oracle/numcalc_port.c) before it is trusted as a checker.

* end to end against reference NumCalc's own be.out (tests/golden, written by the
  reference compiled from its sources);
* quadrature rules against closed forms (polynomial exactness, numpy's Gauss-Legendre);
* singular integrals against the closed-form Laplace potentials of a flat triangle;
* subdivision against its invariants (the subelements tile the element);
* the dense solvers against numpy.linalg.solve;
* the rigid-sphere point-source series (a sanity bound: discretisation error).
"""
import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn, eval_legendre

import ncref as R

# rel-L2 bound against reference be.out: the files print 7 significant digits (%E)
GOLDEN_TOL = 1e-6


@pytest.fixture(scope="module")
def oracle():
    return R.Oracle()


def _solve(oracle, case, pa):
    A, b = oracle.assemble(pa)
    if case["spec"]["method_solver"] == 4:
        return oracle.lu_solve(A, b)
    # the project's preconditioner (IPRECOND), so the CGS stopping point is the reference's
    x, it, err = oracle.cgs(A, b, precond=case["spec"]["precond"])
    assert err < 1e-9
    return x


def _check(ours, ref, tol):
    ours, ref = np.asarray(ours), np.asarray(ref)
    if np.linalg.norm(ref) == 0:
        assert np.abs(ours).max() < 1e-12
    else:
        assert R.rel_l2(ours, ref) < tol


@pytest.mark.parametrize("name", [n for n in R.case_names()
                                  if not n.startswith(("c1", "c2_")) and R.load_case(n)["spec"]["method_bem"] == 0])
def test_oracle_matches_reference_numcalc(oracle, name):
    """Dense oracle solve against the reference's be.out.  Symmetry-plane projects (sym_*)
    are solved unfolded across their planes (one, two and three planes; symmetric and
    antisymmetric; plane types 3 and 4; a plane off the origin): the first n unknowns of the
    unfolded problem reproduce NumCalc's reflected-element solve.  test_source_*: the point-
    source count -1, with which one plane wave reaches the grid but not the right-hand side.
    curves: frequency curves on admittances and source amplitudes, steps below, between and
    above the curve points.  interior: IFINTE 1 (tau -1, no Burton-Miller term)."""
    case = R.load_case(name)
    for s, f in enumerate(case["spec"]["freqs"]):
        pa = R.problem_arrays(case, f)
        pu = R.unfold_symmetry(pa, case["spec"])
        x = _solve(oracle, case, R.rhs_arrays(pu))
        pB, vB = R.surface_outputs(pa, x[:pa["n"]])
        _check(pB, case["pBoundary"][s], GOLDEN_TOL)
        _check(vB, case["vBoundary"][s], GOLDEN_TOL)
        phi, grad = oracle.field(pu, x, case["eval_nodes"])
        _check(1j * pa["rpfact"] * phi, case["pEvalGrid"][s], GOLDEN_TOL)
        _check(grad, case["vEvalGrid"][s], GOLDEN_TOL)


@pytest.mark.parametrize("name", ["sym_fmm_sl", "sym_fmm_ml"])
def test_symmetric_fmm_reference_within_fmm_accuracy(oracle, name):
    """The reference's SL-/ML-FMM runs with a symmetry plane against the dense unfolded
    solve: they differ by the FMM truncation error only (measured 1.4e-4 and 3.5e-4 on the
    surface, 3e-5 on the grid), so the unfolding also describes the reflected-cluster
    FMM of src/NC_Addresses.cpp:508-624,1317-1473."""
    case = R.load_case(name)
    f = case["spec"]["freqs"][0]
    pa = R.problem_arrays(case, f)
    pu = R.unfold_symmetry(pa, case["spec"])
    x = _solve(oracle, case, pu)
    pB, vB = R.surface_outputs(pa, x[:pa["n"]])
    _check(pB, case["pBoundary"][0], 1e-3)
    phi, grad = oracle.field(pu, x, case["eval_nodes"])
    _check(1j * pa["rpfact"] * phi, case["pEvalGrid"][0], 1e-4)
    _check(grad, case["vEvalGrid"][0], 1e-4)


@pytest.mark.slow
def test_oracle_matches_reference_numcalc_config1(oracle):
    case = R.load_case("c1")
    f = case["spec"]["freqs"][1]
    pa = R.problem_arrays(case, f)
    x = _solve(oracle, case, pa)
    pB, _ = R.surface_outputs(pa, x)
    _check(pB, case["pBoundary"][1], GOLDEN_TOL)
    phi, _ = oracle.field(pa, x, case["eval_nodes"][::7])
    _check(1j * pa["rpfact"] * phi, case["pEvalGrid"][1][::7], GOLDEN_TOL)


def rigid_sphere_point_source(a, r0, k, pts, terms=None):
    """Total pressure of a unit point source at (0,0,r0) outside a rigid sphere of radius a,
    e^{-i w t}, p_inc = e^{ikR}/(4 pi R)."""
    r = np.linalg.norm(pts, axis=1)
    cosg = pts[:, 2] / r
    n_max = terms or int(k * max(r.max(), r0) + 40)
    p = np.zeros(len(pts), complex)
    ka = k * a
    for n in range(n_max):
        jn = spherical_jn(n, k * r0)
        hn0 = jn + 1j * spherical_yn(n, k * r0)
        djn = spherical_jn(n, ka, derivative=True)
        dhn = djn + 1j * spherical_yn(n, ka, derivative=True)
        hr = spherical_jn(n, k * r) + 1j * spherical_yn(n, k * r)
        p += (2 * n + 1) * (jn - djn / dhn * hn0) * hr * eval_legendre(n, cosg)
    return 1j * k / (4 * np.pi) * p


def test_reference_golden_vs_rigid_sphere_series():
    """Discretisation-level sanity bound for config 1 (NumCalc vs the exact series)."""
    case = R.load_case("c1")
    for s, f in enumerate(case["spec"]["freqs"]):
        k = 2 * np.pi * f / case["spec"]["c"]
        exact = rigid_sphere_point_source(0.1, 0.2, k, case["eval_nodes"])
        assert R.rel_l2(case["pEvalGrid"][s], exact) < 5e-3


def test_gauss_legendre_against_numpy(oracle):
    import ctypes
    for n in range(1, 25):
        x = np.zeros(n)
        w = np.zeros(n)
        oracle.lib.ncp_gauss_legendre(n, ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(w.ctypes.data))
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xr).max() < 1e-14 and np.abs(w - wr).max() < 1e-14


@pytest.mark.parametrize("order,degree", [(1, 1), (2, 3), (3, 5), (4, 7)])
def test_triangle_rule_exactness(oracle, order, degree):
    import ctypes
    from math import factorial
    s, t, w = np.zeros(36), np.zeros(36), np.zeros(36)
    n = oracle.lib.ncp_gauss_points(3, order, *(ctypes.c_void_p(v.ctypes.data) for v in (s, t, w)))
    # the reference digits of the 7- and 13-point rules are 13 significant figures
    tol = 1e-15 if order <= 2 else 1e-12
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert abs(np.sum(w[:n] * s[:n] ** a * t[:n] ** b) - exact) < tol


def _tri_problem(X, k=1e-7):
    xyz = np.zeros((1, 4, 3))
    xyz[0, :3] = X
    return dict(n=1, nv=np.array([3], np.int32), xyz=xyz, kind=np.zeros(1, np.int32),
                bc=np.zeros((1, 4), complex), adm=np.zeros(1, complex),
                ps=np.zeros((0, 5)), pw=np.zeros((0, 5)), k=k, h=1.0, rpfact=1.0)


def _edge_terms(X, P):
    """(h_e, s1, s2, r1, r2) of each edge seen from an in-plane point P."""
    out = []
    n = np.cross(X[0] - X[2], X[1] - X[2])
    n /= np.linalg.norm(n)
    for e in range(3):
        A, B = X[e], X[(e + 1) % 3]
        t = (B - A) / np.linalg.norm(B - A)
        nu = np.cross(t, n)                       # in-plane edge normal
        if np.dot(nu, A - P) < 0:
            nu = -nu
        h = np.dot(A - P, nu)
        out.append((h, np.dot(A - P, t), np.dot(B - P, t), np.linalg.norm(A - P), np.linalg.norm(B - P)))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_singular_integrals_laplace_closed_form(oracle, seed):
    """k -> 0: G_ii = (1/4pi) sum h ln((s2+r2)/(s1+r1)); finite-part E_ii = -(1/4pi) sum (s2/r2-s1/r1)/h."""
    # well-shaped triangles (the reference's fixed 6x16-point rule loses accuracy on
    # badly distorted ones): a perturbed equilateral triangle, randomly rotated
    rng = np.random.default_rng(seed)
    X = np.array([[1.0, 0.0, 0.0], [-0.5, 0.866, 0.0], [-0.5, -0.866, 0.0]]) + 0.15 * rng.standard_normal((3, 3))
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    X = X @ Q.T * 0.01
    pa = _tri_problem(X)
    z = oracle.integrals(pa, 0, 0)
    P = X.mean(axis=0)
    terms = _edge_terms(X, P)
    g = sum(h * np.log((s2 + r2) / (s1 + r1)) for h, s1, s2, r1, r2 in terms) / (4 * np.pi)
    e = -sum((s2 / r2 - s1 / r1) / h for h, s1, s2, r1, r2 in terms) / (4 * np.pi)
    assert abs(z[0].real - g) / abs(g) < 5e-4
    assert abs(z[3].real - e) / abs(e) < 5e-4
    assert abs(z[1]) < 1e-9 * abs(z[0]) and abs(z[2]) < 1e-9 * abs(z[0])   # coplanar: no H


def test_regular_integral_against_brute_force(oracle):
    """An element seen from a nearby point: adaptive quadrature vs a 2^-7 uniform refinement."""
    X = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.0]])
    Y = np.array([[0.012, 0.012, 0.001], [0.02, 0.01, 0.0], [0.015, 0.02, 0.002]])
    k = 2 * np.pi * 3000 / 343.0
    xyz = np.zeros((2, 4, 3))
    xyz[0, :3] = Y
    xyz[1, :3] = X
    pa = dict(n=2, nv=np.array([3, 3], np.int32), xyz=xyz, kind=np.zeros(2, np.int32),
              bc=np.zeros((2, 4), complex), adm=np.zeros(2, complex), ps=np.zeros((0, 5)),
              pw=np.zeros((0, 5)), k=k, h=1.0, rpfact=1.0)
    z = oracle.integrals(pa, 0, 1)
    c, nrm, _ = oracle.geometry(pa)
    x, nx = c[0], nrm[0]
    ny = nrm[1]
    # brute force: midpoint rule on 4^7 sub-triangles via barycentric grid
    m = 256
    i, j = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    ok = i + j < m
    pts, wts = [], []
    for (u, v), flip in (((i + 1 / 3, j + 1 / 3), False), ((i + 2 / 3, j + 2 / 3), True)):
        sel = ok if not flip else (i + j < m - 1)
        s, t = u[sel] / m, v[sel] / m
        pts.append(np.outer(s, X[0]) + np.outer(t, X[1]) + np.outer(1 - s - t, X[2]))
        wts.append(np.full(sel.sum(), 0.5 / m ** 2))
    P, W = np.vstack(pts), np.concatenate(wts)
    jac = np.linalg.norm(np.cross(X[0] - X[2], X[1] - X[2]))
    d = P - x
    r = np.linalg.norm(d, axis=1)
    d /= r[:, None]
    G = np.exp(1j * k * r) / (4 * np.pi * r) * W * jac
    Gd = G * (-1 / r + 1j * k)
    rny, rnx = d @ ny, -(d @ nx)
    E = G * ((3 / r ** 2 - k * k) * rny * rnx + (nx @ ny) / r ** 2 - 1j * k / r * (3 * rny * rnx + nx @ ny))
    bf = np.array([G.sum(), (Gd * rny).sum(), (Gd * rnx).sum(), E.sum()])
    # NumCalc's adaptive rule targets 1e-3 per element
    assert np.all(np.abs(z[:4] - bf) / np.abs(bf) < 1e-3)


def test_subelements_tile_the_element(oracle):
    X = np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.0]])
    pa = _tri_problem(X)
    for x in ([0.004, 0.004, 0.0005], [0.011, 0.0, 0.0], [0.003, 0.003, 0.0002], [0.05, 0.05, 0.02]):
        s0, t0, fa, order = oracle.subelements(pa, 0, np.asarray(x))
        assert abs(np.sum(0.5 * fa ** 2) - 0.5) < 1e-14
        assert np.all((order >= 2) & (order <= 6))
        # sample the element; every point lies in exactly one subelement
        rng = np.random.default_rng(1)
        s, t = rng.random((2, 4000))
        keep = s + t < 1
        s, t = s[keep], t[keep]
        a = (s[:, None] - s0[None]) / fa[None]
        b = (t[:, None] - t0[None]) / fa[None]
        inside = (a > 0) & (b > 0) & (a + b < 1)
        assert np.all(inside.sum(axis=1) == 1)
    s0, t0, fa, order = oracle.subelements(pa, 0, np.array([0.05, 0.05, 0.02]))
    assert len(fa) == 1 and fa[0] == 1.0


def test_dense_solvers_against_numpy(oracle):
    case = R.load_case("adm_ps")
    pa = R.problem_arrays(case, 800.0)
    A, b = oracle.assemble(pa)
    ref = np.linalg.solve(A, b)
    x, it, err = oracle.cgs(A, b, precond=2)
    assert R.rel_l2(x, ref) < 1e-8
    x1, it1, err1 = oracle.cgs(A, b, precond=1)
    assert R.rel_l2(x1, ref) < 1e-8
    assert R.rel_l2(oracle.lu_solve(A, b), ref) < 1e-12


@pytest.mark.parametrize("name", ["soft_pw", "adm_ps", "velo_patch", "mixed_pres", "quad", "head8", "c1", "curves",
                                  "interior", "scan1_tbem", "scan2_tbem", "scan3_tbem"])
def test_cgs_ilu_matches_reference_iterations(oracle, name):
    """The incomplete-LU preconditioned CGS (IPRECOND 0, the value mesh2input writes, and
    the scanning degrees 1..3 of IPRECOND -1..-3, whose drop thresholds change the ILU
    pattern) takes the reference's iteration count at every step; its last relative
    residual agrees to the noise of a residual near 1e-10; the solution is the system's."""
    case = R.load_case(name)
    pc = case["spec"]["precond"]
    assert pc <= 0
    for s, f in enumerate(case["spec"]["freqs"]):
        pa = R.problem_arrays(case, f)
        A, b = oracle.assemble(pa)
        x, it, err = oracle.cgs(A, b, precond=pc)
        assert it == case["cgs_iterations"][s], (s, it)
        assert abs(err - case["cgs_relerr"][s]) < 0.2 * case["cgs_relerr"][s]
        assert R.rel_l2(x, np.linalg.solve(A, b)) < 1e-8
