/*
 * ORACLE — test infrastructure only.  A plain-C, FP64, single-threaded restatement
 * of NumCalc's traditional-BEM (TBEM) path, used to check the HIP library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.
 *
 * Every function cites the reference routine it restates (paths relative to the
 * reference repository, src/ = mesh2hrtf/NumCalc/src/).  How each one is pinned
 * (golden be.out from the compiled reference, closed forms, brute force) is listed
 * in the tests that exercise it (tests/test_oracle_*.py).
 *
 * Complex numbers are interleaved (re, im) doubles.  Element node coordinates are
 * stored 4 nodes x 3 components per element (the 4th node is unused for triangles).
 */
#ifndef NUMCALC_PORT_H
#define NUMCALC_PORT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int n;                  /* number of boundary elements (= unknowns)                */
    const int *nv;          /* nodes per element, 3 or 4                                */
    const double *xyz;      /* n*4*3 node coordinates                                   */
    const int *bc_kind;     /* 0 velocity, 1 pressure, 2 velocity + surface admittance  */
    const double *bc;       /* n*4 complex nodal BC values, frequency-scaled as in
                               NC_UpdateFreqCurves (src/NC_EquationSystem.cpp:1815-1961):
                               pressure already divided by i*rho*omega*h               */
    const double *adm;      /* n complex admittances, already multiplied by i*h*k       */
    int n_ps;               /* point sources: x y z re im (strength / (i rho omega h))  */
    const double *ps;
    int n_pw;               /* plane waves: unit direction, re im (/(i rho omega h))    */
    const double *pw;
    double k;               /* wave number                                              */
    double h;               /* harmonic time factor, +1 or -1                           */
    int interior;           /* 1: interior problem (tau = -1, beta = 0)                 */
} ncp_problem;

/* Element centre, unit normal at the centre, area (src/NC_Input.cpp:522-605). */
void ncp_element_geometry(const ncp_problem *p, double *centre, double *normal, double *area);

/* Adaptive subelements of element j seen from point x (src/NC_3dFunctions.cpp:575-783).
 * Writes at most 220 entries; returns the count, or -1 on the reference's error exits. */
int ncp_subelements(const ncp_problem *p, int j, const double *x,
                    double *s0, double *t0, double *fase, int *order);

/* Element integrals G, H, H^T, E and the RHS term [4] for collocation element i and
 * field element j (regular: src/NC_EquationSystem.cpp:1600-1728; i == j singular:
 * src/NC_EquationSystem.cpp:1408-1597).  out = 5 complex. */
int ncp_element_integrals(const ncp_problem *p, int i, int j, double *out);

/* Burton-Miller dense system, row-major A (n*n complex) and b (n complex)
 * (src/NC_EquationSystem.cpp:79-302, 1731-1812). */
int ncp_assemble(const ncp_problem *p, double *A, double *b);

/* Right-preconditioned CGS with the unconjugated bilinear form
 * (src/NC_CommonFunctions.cpp:1066-1208).  precond: 1 row-scaling, 2 none, 0..-3
 * incomplete LU with scanning degree -precond (the dense TBEM thresholds of
 * src/NC_CommonFunctions.cpp:1367-1376).
 * Returns iterations; *relerr receives the final relative residual. */
int ncp_cgs(int n, const double *A, const double *b, double *x, int precond,
            int maxit, double tol, double *relerr);

/* Unpivoted dense LU solve (src/NC_TypeDefinition.h:542-594); A is overwritten. */
void ncp_lu_solve(int n, double *A, double *b);

/* Velocity potential and its gradient at m field points, incident field included
 * (src/NC_PostProcessing.cpp:581-866, 869-975, 1217-1322).  sol = solved unknowns
 * (phi for velocity elements, v for pressure elements). */
int ncp_field(const ncp_problem *p, const double *sol, int m, const double *pts,
              double *phi, double *grad);

/* Gauss-Legendre nodes/weights on [-1,1] (the tables of
 * src/NC_IntegrationConstants.h:24-1900, computed by Newton iteration). */
void ncp_gauss_legendre(int n, double *x, double *w);

/* Quadrature points of one (sub)element: triangle rules or n x n Gauss-Legendre
 * (NC_ComputeGausseanPoints, src/NC_3dFunctions.cpp:455-530).  Returns the count. */
int ncp_gauss_points(int nv, int order, double *s, double *t, double *w);

#ifdef __cplusplus
}
#endif
#endif
