/*
 * ORACLE — test infrastructure only (see numcalc_port.h).  Plain C restatement of
 * NumCalc's TBEM path in FP64.  Built with -ffp-contract=off so that its rounding
 * follows the reference's x86-64 SSE2 build as closely as plain C allows.
 */
#include "numcalc_port.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_ 3.1415926535897932385
#define PI4_ 12.566370614359172944
#define EPSY_ 1.0e-14   /* src/NC_ConstantsVariables.h:60 */
#define MSBE_ 220       /* src/NC_ConstantsVariables.h:56 */

typedef struct { double re, im; } cplx;

static inline cplx cmk(double r, double i) { cplx z = {r, i}; return z; }
static inline cplx cadd(cplx a, cplx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cplx csub(cplx a, cplx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cplx cmul(cplx a, cplx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cplx cscal(cplx a, double s) { return cmk(a.re * s, a.im * s); }
static inline cplx cdiv(cplx a, cplx b)
{
    double d = b.re * b.re + b.im * b.im;
    return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
static inline double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
static inline cplx cld(const double *p) { return cmk(p[0], p[1]); }

/* ------------------------------------------------------------------ quadrature */

/* Gauss-Legendre nodes ascending on [-1,1] by Newton iteration on P_n; the reference
 * tabulates the same rule to 15 digits (src/NC_IntegrationConstants.h:24-1900,
 * looked up by BLGauCooWe, src/NC_CommonFunctions.cpp:412-1063). */
void ncp_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(PI_ * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; k++) {
                double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            pp = n * (z * p0 - p1) / (z * z - 1.0);
            double dz = p0 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        {   /* final derivative at the converged node */
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; k++) {
                double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            pp = n * (z * p0 - p1) / (z * z - 1.0);
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    if (n % 2) x[n / 2] = 0.0;
}

/* Triangle rules on the unit triangle (s, t >= 0, s + t <= 1), weights sum to 1/2.
 * Same published rules and digits as src/NC_IntegrationConstants.h:3729-3865:
 * order 1 centroid, order 2 Strang-Fix 4-point (negative centre weight), order 3
 * Bathe 7-point, orders 4..7 the 13-point degree-7 rule. */
static const double TRI1[1][3] = {{1.0 / 3.0, 1.0 / 3.0, 0.5}};
static const double TRI4[4][3] = {
    {1.0 / 3.0, 1.0 / 3.0, -27.0 / 96.0}, {0.2, 0.2, 25.0 / 96.0},
    {0.6, 0.2, 25.0 / 96.0}, {0.2, 0.6, 25.0 / 96.0}};
static const double TRI7[7][3] = {
    {0.1012865073235, 0.1012865073235, 0.1259391805448 / 2.0},
    {0.7974269853531, 0.1012865073235, 0.1259391805448 / 2.0},
    {0.1012865073235, 0.7974269853531, 0.1259391805448 / 2.0},
    {0.4701420641051, 0.0597158717898, 0.1323941527885 / 2.0},
    {0.4701420641051, 0.4701420641051, 0.1323941527885 / 2.0},
    {0.0597158717898, 0.4701420641051, 0.1323941527885 / 2.0},
    {0.3333333333333, 0.3333333333333, 0.225 / 2.0}};
static const double TRI13[13][3] = {
    {0.0651301029022, 0.0651301029022, 0.0533472356088 / 2.0},
    {0.8697397941956, 0.0651301029022, 0.0533472356088 / 2.0},
    {0.0651301029022, 0.8697397941956, 0.0533472356088 / 2.0},
    {0.3128654960049, 0.0486903154253, 0.0771137608903 / 2.0},
    {0.6384441885698, 0.3128654960049, 0.0771137608903 / 2.0},
    {0.0486903154253, 0.6384441885698, 0.0771137608903 / 2.0},
    {0.6384441885698, 0.0486903154253, 0.0771137608903 / 2.0},
    {0.3128654960049, 0.6384441885698, 0.0771137608903 / 2.0},
    {0.0486903154253, 0.3128654960049, 0.0771137608903 / 2.0},
    {0.2603459660790, 0.2603459660790, 0.1756152574332 / 2.0},
    {0.4793080678419, 0.2603459660790, 0.1756152574332 / 2.0},
    {0.2603459660790, 0.4793080678419, 0.1756152574332 / 2.0},
    {0.3333333333333, 0.3333333333333, -0.1495700444677 / 2.0}};

/* NC_ComputeGausseanPoints (src/NC_3dFunctions.cpp:455-530): returns the count. */
static int gauss_points(int nv, int order, double *s, double *t, double *w)
{
    if (nv == 3) {
        const double(*r)[3];
        int n;
        switch (order) {
        case 1: r = TRI1; n = 1; break;
        case 2: r = TRI4; n = 4; break;
        case 3: r = TRI7; n = 7; break;
        default: r = TRI13; n = 13; break;
        }
        for (int i = 0; i < n; i++) { s[i] = r[i][0]; t[i] = r[i][1]; w[i] = r[i][2]; }
        return n;
    }
    double x[8], wx[8];
    ncp_gauss_legendre(order, x, wx);
    int n = 0;
    for (int i = 0; i < order; i++)
        for (int j = 0; j < order; j++) { s[n] = x[i]; t[n] = x[j]; w[n] = wx[i] * wx[j]; n++; }
    return n;
}

int ncp_gauss_points(int nv, int order, double *s, double *t, double *w)
{
    return gauss_points(nv, order, s, t, w);
}

/* ------------------------------------------------------------------ geometry */

/* local node coordinates, src/NC_Arrays.h:126-129 */
static const double CSI6[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5}, ETA6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
static const double CSI8[8] = {1.0, -1.0, -1.0, 1.0, 0.0, -1.0, 0.0, 1.0};
static const double ETA8[8] = {1.0, 1.0, -1.0, -1.0, 1.0, 0.0, -1.0, 0.0};

/* shape functions (NC_ComputeGlobalCoordinates, src/NC_3dFunctions.cpp:95-129) */
static void shape(int nv, double s, double t, double *N)
{
    if (nv == 3) { N[0] = s; N[1] = t; N[2] = 1.0 - s - t; return; }
    double s1 = 0.25 * (s + 1.0), s2 = 0.25 * (s - 1.0), t1 = t + 1.0, t2 = t - 1.0;
    N[0] = s1 * t1; N[1] = -s2 * t1; N[2] = s2 * t2; N[3] = -s1 * t2;
}

static void global_point(const double *X, int nv, double s, double t, double *P, double *N)
{
    shape(nv, s, t, N);
    for (int d = 0; d < 3; d++) {
        P[d] = 0.0;
        for (int a = 0; a < nv; a++) P[d] += N[a] * X[3 * a + d];
    }
}

/* unit normal and Jacobian (NC_ComputeUnitVectorOnElement, src/NC_3dFunctions.cpp:132-178) */
static double unit_normal(const double *X, int nv, double s, double t, double *n)
{
    double ds[4], dt[4], xs[3], xt[3];
    if (nv == 3) {
        ds[0] = dt[1] = 1.0; ds[1] = dt[0] = 0.0; ds[2] = dt[2] = -1.0;
    } else {
        ds[0] = 0.25 * (t + 1.0); ds[2] = 0.25 * (t - 1.0); ds[1] = -ds[0]; ds[3] = -ds[2];
        dt[0] = 0.25 * (s + 1.0); dt[2] = 0.25 * (s - 1.0); dt[1] = -dt[2]; dt[3] = -dt[0];
    }
    for (int d = 0; d < 3; d++) {
        xs[d] = xt[d] = 0.0;
        for (int a = 0; a < nv; a++) { xs[d] += ds[a] * X[3 * a + d]; xt[d] += dt[a] * X[3 * a + d]; }
    }
    n[0] = xs[1] * xt[2] - xs[2] * xt[1];
    n[1] = xs[2] * xt[0] - xs[0] * xt[2];
    n[2] = xs[0] * xt[1] - xs[1] * xt[0];
    double L = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= L; n[1] /= L; n[2] /= L;
    return L;
}

void ncp_element_geometry(const ncp_problem *p, double *centre, double *normal, double *area)
{
    for (int e = 0; e < p->n; e++) {
        const double *X = p->xyz + 12 * e;
        int nv = p->nv[e];
        double s = nv == 3 ? 1.0 / 3.0 : 0.0, a0 = nv == 3 ? 0.5 : 4.0, N[4];
        global_point(X, nv, s, s, centre + 3 * e, N);
        area[e] = unit_normal(X, nv, s, s, normal + 3 * e) * a0;
    }
}

static void elem_geom(const ncp_problem *p, int e, double *c, double *n, double *area)
{
    const double *X = p->xyz + 12 * e;
    int nv = p->nv[e];
    double s = nv == 3 ? 1.0 / 3.0 : 0.0, a0 = nv == 3 ? 0.5 : 4.0, N[4];
    global_point(X, nv, s, s, c, N);
    *area = unit_normal(X, nv, s, s, n) * a0;
}

/* ------------------------------------------------------------------ subelements */

/* NC_GenerateSubelements (src/NC_3dFunctions.cpp:575-783): breadth-first quadtree.
 * A (sub)element is split in 4 while |centre - x| / sqrt(area) < 1.3; a stored one gets
 * the smallest Gauss order m in [2, 6] whose error estimates (src/NC_Macros.h:44-50) are
 * all below 1e-3 at disfac = 0.5 / ratdis. */
int ncp_subelements(const ncp_problem *p, int j, const double *x,
                    double *s0, double *t0, double *fase, int *order)
{
    enum { MNSE = 60, MVX = 4, NSE = 4 };
    const double *X = p->xyz + 12 * j;
    int nv = p->nv[j];
    double area, cj[3], nj[3];
    elem_geom(p, j, cj, nj, &area);

    double fs[MNSE][MVX], ft[MNSE][MVX], es[MNSE][MVX], et[MNSE][MVX];
    double ms[2 * MVX], mt[2 * MVX], P[3], N[4];
    const double *cs = nv == 3 ? CSI6 : CSI8, *ct = nv == 3 ? ETA6 : ETA8;
    for (int v = 0; v < nv; v++) { fs[0][v] = cs[v]; ft[0][v] = ct[v]; }

    int nsbe = 0, nsfl = 1, ndie;
    double faclin = 2.0;
    do {
        ndie = 0;
        faclin *= 0.5;
        double arels = area * faclin * faclin;
        int nsel = nsfl;
        for (int q = 0; q < nsel; q++)
            for (int v = 0; v < nv; v++) { es[q][v] = fs[q][v]; et[q][v] = ft[q][v]; }
        for (int q = 0; q < nsel; q++) {
            double sc = 0.0, tc = 0.0;
            for (int v = 0; v < nv; v++) { sc += es[q][v]; tc += et[q][v]; }
            sc /= (double)nv;
            tc /= (double)nv;
            global_point(X, nv, sc, tc, P, N);
            double dx = P[0] - x[0], dy = P[1] - x[1], dz = P[2] - x[2];
            double ratdis = sqrt(dx * dx + dy * dy + dz * dz) / sqrt(arels);
            if (ratdis < 1.3) {
                if (++ndie > 15) return -1;
                nsfl = ndie * NSE;
                int f0 = nsfl - NSE;
                for (int v = 0; v < nv; v++) {
                    int v1 = v + 1 == nv ? 0 : v + 1;
                    ms[v] = es[q][v]; mt[v] = et[q][v];
                    ms[v + nv] = (es[q][v] + es[q][v1]) / 2.0;
                    mt[v + nv] = (et[q][v] + et[q][v1]) / 2.0;
                }
                for (int v = 0; v < nv; v++) {
                    int u = f0 + v, j1 = v + nv, j2 = j1 - 1;
                    if (j2 < nv) j2 += nv;
                    fs[u][0] = ms[v]; fs[u][1] = ms[j1];
                    ft[u][0] = mt[v]; ft[u][1] = mt[j1];
                    if (nv == 4) {
                        fs[u][2] = sc; fs[u][3] = ms[j2];
                        ft[u][2] = tc; ft[u][3] = mt[j2];
                    } else {
                        fs[u][2] = ms[j2]; ft[u][2] = mt[j2];
                        fs[f0 + NSE - 1][v] = ms[j1];
                        ft[f0 + NSE - 1][v] = mt[j1];
                    }
                }
            } else {
                if (nsbe >= MSBE_) return -1;
                if (nv == 4) {
                    s0[nsbe] = (es[q][0] + es[q][1] + es[q][2] + es[q][3]) / 4.0;
                    t0[nsbe] = (et[q][0] + et[q][1] + et[q][2] + et[q][3]) / 4.0;
                    fase[nsbe] = faclin;
                } else {
                    double xi = (es[q][0] + es[q][1] + es[q][2]) / 3.0;
                    double eta = (et[q][0] + et[q][1] + et[q][2]) / 3.0;
                    double dxi = 0.0, deta;
                    int jr = 0;
                    for (int v = 0; v < 3; v++) {
                        dxi = xi - es[q][v];
                        deta = eta - et[q][v];
                        if (fabs(dxi - deta) < 1.0e-8) { jr = v; break; }
                    }
                    s0[nsbe] = es[q][jr];
                    t0[nsbe] = et[q][jr];
                    fase[nsbe] = dxi >= 0.0 ? faclin : -faclin;
                }
                double disfac = 0.5 / ratdis;
                int m = 6;
                for (int mm = 2; mm <= 6; mm++) {
                    double eg = 32.0, eh = 64.0 * (2.0 * mm + 1.0),
                           ee = 128.0 * (mm + 1.0) * (2.0 * mm + 1.0);
                    for (int i = 0; i < 2 * mm + 1; i++) eg *= disfac;
                    for (int i = 0; i < 2 * mm + 2; i++) eh *= disfac;
                    for (int i = 0; i < 2 * mm + 3; i++) ee *= disfac;
                    if (ee < 1e-3 && eh < 1e-3 && eg < 1e-3) { m = mm; break; }
                }
                order[nsbe] = m;
                nsbe++;
            }
        }
    } while (ndie > 0);
    return nsbe;
}

/* ------------------------------------------------------------------ BC helpers */

typedef struct {
    double gamma, tau;
    cplx beta;
    double hk;
} bm_consts;

static bm_consts bm(const ncp_problem *p)
{
    bm_consts c;
    c.gamma = 1.0;   /* src/NC_EquationSystem.cpp:85-86 */
    c.tau = p->interior ? -1.0 : 1.0;
    c.beta = p->interior ? cmk(0.0, 0.0) : cmk(0.0, p->h / p->k);
    c.hk = p->h * p->k;
    return c;
}

/* kind of the unknown (0 potential, 1 velocity) and the RHS flag (src/NC_EquationSystem.cpp:268-279) */
static int ibv_of(const ncp_problem *p, int e) { return p->bc_kind[e] == 1 ? 1 : 0; }
static int has_rhs(const ncp_problem *p, int e)
{
    for (int a = 0; a < p->nv[e]; a++)
        if (sqrt(cabs2(cld(p->bc + 8 * e + 2 * a))) > EPSY_) return 1;
    return 0;
}

/* ------------------------------------------------------------------ integrals */

/* NC_RegularIntegration (src/NC_EquationSystem.cpp:1600-1728) */
static int regular(const ncp_problem *p, const double *x, const double *nx, int j, cplx out[5])
{
    bm_consts c = bm(p);
    const double *X = p->xyz + 12 * j;
    int nv = p->nv[j], ibv = ibv_of(p, j), rhs = has_rhs(p, j);
    double s0[MSBE_], t0[MSBE_], fa[MSBE_];
    int ord[MSBE_];
    int nsb = ncp_subelements(p, j, x, s0, t0, fa, ord);
    if (nsb < 0) return -1;
    for (int i = 0; i < 5; i++) out[i] = cmk(0.0, 0.0);
    for (int sb = 0; sb < nsb; sb++) {
        double gs[36], gt[36], gw[36];
        int ng = gauss_points(nv, ord[sb], gs, gt, gw);
        int whole = fabs(fa[sb] - 1.0) < EPSY_;
        for (int g = 0; g < ng; g++) {
            double s, t, w;
            if (whole) { s = gs[g]; t = gt[g]; w = gw[g]; }
            else { s = s0[sb] + gs[g] * fa[sb]; t = t0[sb] + gt[g] * fa[sb]; w = gw[g] * fa[sb] * fa[sb]; }
            double P[3], N[4], ny[3];
            global_point(X, nv, s, t, P, N);
            double jac = unit_normal(X, nv, s, t, ny);
            cplx bcv = cmk(0.0, 0.0);
            if (rhs) for (int a = 0; a < nv; a++) bcv = cadd(bcv, cscal(cld(p->bc + 8 * j + 2 * a), N[a]));
            double wga = w * jac;
            double d[3] = {P[0] - x[0], P[1] - x[1], P[2] - x[2]};
            double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            d[0] /= r; d[1] /= r; d[2] /= r;
            double a1 = c.hk * r, a2 = wga / PI4_ / r;
            cplx G = cmk(cos(a1) * a2, sin(a1) * a2);
            cplx Gd = cmul(G, cmk(-1.0 / r, c.hk));
            double rny = d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2];
            double rnx = -(d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]);
            cplx H = cscal(Gd, rny), HT = cscal(Gd, rnx);
            double rq = rny * rnx, nn = nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2], dq = r * r;
            cplx E = cmul(G, cmk((3.0 / dq - p->k * p->k) * rq + nn / dq, -c.hk / r * (3.0 * rq + nn)));
            out[0] = cadd(out[0], G);
            out[1] = cadd(out[1], H);
            out[2] = cadd(out[2], HT);
            out[3] = cadd(out[3], E);
            if (rhs) {
                if (ibv == 0)
                    out[4] = cadd(out[4], cmul(cadd(cscal(G, c.gamma * c.tau), cmul(HT, c.beta)), bcv));
                else
                    out[4] = csub(out[4], cmul(cadd(cscal(H, c.gamma * c.tau), cmul(E, c.beta)), bcv));
            }
        }
    }
    return 0;
}

/* NC_SingularIntegration (src/NC_EquationSystem.cpp:1408-1597): G, H, H^T on two
 * collapsed (Duffy) sub-triangles per edge with 4x4 Gauss-Legendre; the hypersingular
 * E as an edge line integral (4 sections x 3 points per edge) plus k^2 * int G nx.ny. */
static void singular(const ncp_problem *p, int j, const double *x, const double *nx, cplx out[5])
{
    bm_consts c = bm(p);
    const double *X = p->xyz + 12 * j;
    int nv = p->nv[j], ibv = ibv_of(p, j), rhs = has_rhs(p, j);
    double k2 = p->k * p->k;
    double g3x[3], g3w[3], g4x[4], g4w[4];
    ncp_gauss_legendre(3, g3x, g3w);
    ncp_gauss_legendre(4, g4x, g4w);
    for (int i = 0; i < 5; i++) out[i] = cmk(0.0, 0.0);
    cplx bcv = cmk(0.0, 0.0);
    for (int eg = 0; eg < nv; eg++) {
        int e1 = eg + 1 == nv ? 0 : eg + 1, e2 = eg + nv;
        double dv[3], tv[3], len = 0.0;
        for (int d = 0; d < 3; d++) { dv[d] = X[3 * e1 + d] - X[3 * eg + d]; len += dv[d] * dv[d]; }
        len = sqrt(len);
        for (int d = 0; d < 3; d++) tv[d] = dv[d] / len;
        len /= 8.0;   /* 2 * 4 sections */
        cplx acc = cmk(0.0, 0.0);
        double mid = -1.25;
        for (int sec = 0; sec < 4; sec++) {
            mid += 0.5;
            for (int g = 0; g < 3; g++) {
                double sg = mid + g3x[g] / 4.0, wg = g3w[g] * len, d[3];
                for (int q = 0; q < 3; q++) d[q] = X[3 * eg + q] + dv[q] * (sg + 1.0) / 2.0 - x[q];
                double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                d[0] /= r; d[1] /= r; d[2] /= r;
                double a1 = c.hk * r, a2 = PI4_ * r;
                cplx G = cmul(cmk(cos(a1) / a2, sin(a1) / a2), cmk(-1.0 / r, c.hk));
                cplx gy[3] = {cscal(G, d[0]), cscal(G, d[1]), cscal(G, d[2])};
                cplx wk0 = csub(cscal(gy[1], tv[2]), cscal(gy[2], tv[1]));
                cplx wk1 = csub(cscal(gy[2], tv[0]), cscal(gy[0], tv[2]));
                cplx wk2 = csub(cscal(gy[0], tv[1]), cscal(gy[1], tv[0]));
                cplx s = cadd(cadd(cscal(wk0, nx[0]), cscal(wk1, nx[1])), cscal(wk2, nx[2]));
                acc = cadd(acc, cscal(s, wg));
            }
        }
        out[3] = cadd(out[3], acc);
        for (int sec = 0; sec < 2; sec++) {
            double ss[3], tt[3], ares;
            if (nv == 3) {
                ares = 1.0 / 24.0 / 2.0;
                ss[0] = tt[0] = 1.0 / 3.0;
                if (sec == 0) { ss[1] = CSI6[eg]; ss[2] = CSI6[e2]; tt[1] = ETA6[eg]; tt[2] = ETA6[e2]; }
                else { ss[1] = CSI6[e2]; ss[2] = CSI6[e1]; tt[1] = ETA6[e2]; tt[2] = ETA6[e1]; }
            } else {
                ares = 0.25 / 2.0;
                ss[0] = tt[0] = 0.0;
                if (sec == 0) { ss[1] = CSI8[eg]; ss[2] = CSI8[e2]; tt[1] = ETA8[eg]; tt[2] = ETA8[e2]; }
                else { ss[1] = CSI8[e2]; ss[2] = CSI8[e1]; tt[1] = ETA8[e2]; tt[2] = ETA8[e1]; }
            }
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) {
                    double sg = g4x[a], tg = g4x[b], wq = g4w[a] * g4w[b];
                    double s = 0.5 * (1.0 - sg) * ss[0] + 0.25 * (1.0 + sg) * ((1.0 - tg) * ss[1] + (1.0 + tg) * ss[2]);
                    double t = 0.5 * (1.0 - sg) * tt[0] + 0.25 * (1.0 + sg) * ((1.0 - tg) * tt[1] + (1.0 + tg) * tt[2]);
                    double P[3], N[4], ny[3];
                    global_point(X, nv, s, t, P, N);
                    double jac = unit_normal(X, nv, s, t, ny);
                    bcv = cmk(0.0, 0.0);
                    if (rhs) for (int q = 0; q < nv; q++) bcv = cadd(bcv, cscal(cld(p->bc + 8 * j + 2 * q), N[q]));
                    double wga = wq * (1.0 + sg) * ares * jac;
                    double d[3] = {P[0] - x[0], P[1] - x[1], P[2] - x[2]};
                    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                    d[0] /= r; d[1] /= r; d[2] /= r;
                    double a1 = c.hk * r, a2 = wga / PI4_ / r;
                    cplx G = cmk(cos(a1) * a2, sin(a1) * a2);
                    cplx Gd = cmul(G, cmk(-1.0 / r, c.hk));
                    double rny = d[0] * ny[0] + d[1] * ny[1] + d[2] * ny[2];
                    double rnx = -(d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]);
                    cplx H = cscal(Gd, rny), HT = cscal(Gd, rnx);
                    out[0] = cadd(out[0], G);
                    out[1] = cadd(out[1], H);
                    out[2] = cadd(out[2], HT);
                    out[3] = cadd(out[3], cscal(G, k2 * (nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2])));
                    if (rhs && ibv == 0)
                        out[4] = cadd(out[4], cmul(cadd(cscal(G, c.gamma * c.tau), cmul(HT, c.beta)), bcv));
                }
        }
    }
    if (rhs && ibv == 1)
        out[4] = cscal(cmul(cadd(cscal(out[1], c.gamma * c.tau), cmul(out[3], c.beta)), bcv), -1.0);
}

int ncp_element_integrals(const ncp_problem *p, int i, int j, double *out)
{
    double xi[3], ni[3], ai;
    cplx z[5];
    elem_geom(p, i, xi, ni, &ai);
    if (i == j) singular(p, j, xi, ni, z);
    else if (regular(p, xi, ni, j, z)) return -1;
    for (int q = 0; q < 5; q++) { out[2 * q] = z[q].re; out[2 * q + 1] = z[q].im; }
    return 0;
}

/* ------------------------------------------------------------------ assembly */

/* NC_IncidentWaveRHS (src/NC_EquationSystem.cpp:1731-1812), no symmetry planes */
static cplx incident_rhs(const ncp_problem *p, const double *x, const double *nx)
{
    bm_consts c = bm(p);
    cplx z = cmk(0.0, 0.0);
    for (int w = 0; w < p->n_pw; w++) {
        const double *d = p->pw + 5 * w;
        double a = (d[0] * x[0] + d[1] * x[1] + d[2] * x[2]) * c.hk;
        cplx g = cmul(cmk(cos(a), sin(a)), cld(d + 3));
        double b = c.hk * (d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]);
        cplx v = cmk(-g.im * b, g.re * b);
        z = csub(z, cadd(cscal(g, c.gamma), cmul(v, cscal(c.beta, c.tau))));
    }
    for (int s = 0; s < p->n_ps; s++) {
        const double *y = p->ps + 5 * s;
        double v2[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
        double R = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
        v2[0] /= R; v2[1] /= R; v2[2] /= R;
        double a = c.hk * R, b = 1.0 / PI4_ / R;
        cplx g = cmul(cmk(cos(a) * b, sin(a) * b), cld(y + 3));
        double rn = v2[0] * nx[0] + v2[1] * nx[1] + v2[2] * nx[2];
        cplx v = cmul(cmk(-rn / R, c.hk * rn), g);
        z = csub(z, cadd(cscal(g, c.gamma), cmul(v, cscal(c.beta, c.tau))));
    }
    return z;
}

/* NC_BuildSystemTBEM / NC_ComputeEntriesForTBEM / NC_AssembleTBEM
 * (src/NC_EquationSystem.cpp:100-302) */
int ncp_assemble(const ncp_problem *p, double *A, double *b)
{
    bm_consts c = bm(p);
    int n = p->n;
    memset(A, 0, sizeof(double) * 2 * (size_t)n * n);
    memset(b, 0, sizeof(double) * 2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        double xi[3], ni[3], ai;
        elem_geom(p, i, xi, ni, &ai);
        int ibv = ibv_of(p, i);
        cplx zb = cmk(0.0, 0.0);
        for (int a = 0; a < p->nv[i]; a++) zb = cadd(zb, cld(p->bc + 8 * i + 2 * a));
        zb = cmk(zb.re / (double)p->nv[i], zb.im / (double)p->nv[i]);
        cplx adm = p->bc_kind[i] == 2 ? cld(p->adm + 2 * i) : cmk(0.0, 0.0);
        double *Aii = A + 2 * ((size_t)i * n + i);
        cplx fr = ibv == 0 ? cscal(csub(cmul(adm, c.beta), cmk(c.gamma, 0.0)), 0.5)
                           : cscal(c.beta, -0.5 * c.tau);
        Aii[0] += fr.re; Aii[1] += fr.im;
        cplx z0 = cmk(0.0, 0.0);
        if (ibv == 0) z0 = cadd(z0, cscal(cmul(zb, c.beta), c.tau * 0.5));
        else z0 = cadd(z0, cscal(zb, c.tau * 0.5));
        if (p->n_ps + p->n_pw > 0) z0 = cadd(z0, incident_rhs(p, xi, ni));
        cplx bi = z0;
        for (int j = 0; j < n; j++) {
            cplx z[5];
            if (i == j) singular(p, j, xi, ni, z);
            else if (regular(p, xi, ni, j, z)) return -1;
            int jb = ibv_of(p, j), jr = has_rhs(p, j);
            cplx s;
            if (jb == 0) {
                s = cadd(cscal(z[1], c.gamma * c.tau), cmul(z[3], c.beta));
                if (p->bc_kind[j] == 2)
                    s = cadd(s, cmul(cadd(cscal(z[0], c.gamma), cmul(z[2], cscal(c.beta, c.tau))),
                                     cld(p->adm + 2 * j)));
            } else {
                s = csub(cscal(z[0], -c.gamma * c.tau), cmul(z[2], c.beta));
            }
            if (jr) bi = cadd(bi, z[4]);
            double *Aij = A + 2 * ((size_t)i * n + j);
            Aij[0] += s.re; Aij[1] += s.im;
        }
        b[2 * i] += bi.re;
        b[2 * i + 1] += bi.im;
    }
    return 0;
}

/* ------------------------------------------------------------------ solvers */

static void matvec(int n, const double *A, const cplx *x, cplx *y)
{
    for (int i = 0; i < n; i++) {
        cplx s = cmk(0.0, 0.0);
        const double *row = A + 2 * (size_t)i * n;
        for (int j = 0; j < n; j++) s = cadd(s, cmul(cld(row + 2 * j), x[j]));
        y[i] = s;
    }
}

static cplx dotu(int n, const cplx *a, const cplx *b)
{
    cplx s = cmk(0.0, 0.0);
    for (int i = 0; i < n; i++) s = cadd(s, cmul(a[i], b[i]));
    return s;
}

/* NC_IterativeSolverCGS (src/NC_CommonFunctions.cpp:1066-1208) with the scaling
 * preconditioner of NC_ComputeScalingVector (:1322-1354) applied on the right. */
/* Incomplete LU of the TBEM branch of NC_IncompleteLUDecomposition
 * (src/NC_CommonFunctions.cpp:1356-1659), dense storage: the matrix and right-hand side
 * are scaled row by row (d_i = sqrt(n / sum_j |a_ij|^2)), the pattern keeps
 * |a_ij| > threshold and the diagonal, and the Crout recurrences of the reference fill
 * L (lower, with the diagonal) column by column and U (unit upper) row by row. */
typedef struct {
    int n;
    unsigned char *mask;   /* [n][n] pattern */
    cplx *LU;              /* [n][n]: L on and below the diagonal, U above */
} ilu_t;

static void ilu_build(int n, cplx *As, cplx *bs, int degree, ilu_t *f)
{
    static const double thr_tbem[4] = {1.2, 1.0, 0.8, 0.6};
    const double thr = thr_tbem[degree < 0 ? 0 : degree > 3 ? 3 : degree];
    f->n = n;
    f->mask = calloc((size_t)n * n, 1);
    f->LU = calloc((size_t)n * n, sizeof(cplx));
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += cabs2(As[(size_t)i * n + j]);
        const double d = sqrt((double)n / s);
        for (int j = 0; j < n; j++) {
            cplx *a = As + (size_t)i * n + j;
            *a = cscal(*a, d);
            if (sqrt(cabs2(*a)) > thr || i == j) {
                f->mask[(size_t)i * n + j] = 1;
                f->LU[(size_t)i * n + j] = *a;
            }
        }
        bs[i] = cscal(bs[i], d);
    }
    const unsigned char *M = f->mask;
    cplx *F = f->LU;
    for (int i = 0; i < n; i++) {
        for (int k = i; k < n; k++) {          /* column i of L */
            if (!M[(size_t)k * n + i]) continue;
            cplx s = F[(size_t)k * n + i];
            for (int m = 0; m < i; m++)
                if (M[(size_t)k * n + m] && M[(size_t)m * n + i]) s = csub(s, cmul(F[(size_t)k * n + m], F[(size_t)m * n + i]));
            F[(size_t)k * n + i] = s;
        }
        for (int k = i + 1; k < n; k++) {      /* row i of U */
            if (!M[(size_t)i * n + k]) continue;
            cplx s = F[(size_t)i * n + k];
            for (int m = 0; m < i; m++)
                if (M[(size_t)i * n + m] && M[(size_t)m * n + k]) s = csub(s, cmul(F[(size_t)i * n + m], F[(size_t)m * n + k]));
            F[(size_t)i * n + k] = cdiv(s, F[(size_t)i * n + i]);
        }
    }
}

/* NC_IncompleteLUForBack (src/NC_CommonFunctions.cpp:1673-1721): y = (L U)^-1 x */
static void ilu_apply(const ilu_t *f, const cplx *x, cplx *y)
{
    const int n = f->n;
    const cplx *F = f->LU;
    const unsigned char *M = f->mask;
    for (int i = 0; i < n; i++) y[i] = x[i];
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < i; j++)
            if (M[(size_t)i * n + j]) y[i] = csub(y[i], cmul(F[(size_t)i * n + j], y[j]));
        y[i] = cdiv(y[i], F[(size_t)i * n + i]);
    }
    for (int i = n - 1; i > 0; i--)
        for (int j = 0; j < i; j++)
            if (M[(size_t)j * n + i]) y[j] = csub(y[j], cmul(F[(size_t)j * n + i], y[i]));
}

int ncp_cgs(int n, const double *A0, const double *b0, double *xout, int precond,
            int maxit, double tol, double *relerr)
{
    cplx *w = calloc((size_t)10 * n, sizeof(cplx));
    double *sc = malloc(sizeof(double) * n);
    cplx *X = w, *R0 = w + n, *R = w + 2 * n, *P = w + 3 * n, *V = w + 4 * n, *Q = w + 5 * n,
         *U = w + 6 * n, *UQ = w + 7 * n, *T = w + 8 * n, *B = w + 9 * n;
    const double *A = A0;
    cplx *As = NULL;
    ilu_t ilu = {0, NULL, NULL};
    for (int i = 0; i < n; i++) B[i] = cld(b0 + 2 * i);
    if (precond <= 0) {
        As = malloc(sizeof(cplx) * (size_t)n * n);
        memcpy(As, A0, sizeof(cplx) * (size_t)n * n);
        ilu_build(n, As, B, -precond, &ilu);
        A = (const double *)As;
    }
    for (int i = 0; i < n; i++) {
        if (precond == 1) {
            double s = 0.0;
            for (int j = 0; j < n; j++) s += cabs2(cld(A + 2 * ((size_t)i * n + j)));
            sc[i] = sqrt((double)n / s);
        } else sc[i] = 1.0;
    }
    /* M^-1 applied to the Krylov vector before the product */
#define PRECOND_IN(src)                                                 \
    do {                                                                \
        if (precond <= 0) ilu_apply(&ilu, (src), T);                    \
        else for (int i_ = 0; i_ < n; i_++) T[i_] = cscal((src)[i_], sc[i_]); \
    } while (0)
    /* x0 = 0, so A M^-1 x0 = 0 and r0 = b */
    for (int i = 0; i < n; i++) U[i] = R[i] = P[i] = R0[i] = B[i];
    double err0 = 0.0, err = 1.0;
    cplx rho0 = cmk(0, 0), rho1 = cmk(0, 0);
    int j;
    for (j = 0; j <= maxit; j++) {
        PRECOND_IN(P);
        matvec(n, A, T, V);
        rho0 = j == 0 ? dotu(n, R, R0) : rho1;
        cplx alpha = cdiv(rho0, dotu(n, V, R0));
        for (int i = 0; i < n; i++) { Q[i] = csub(U[i], cmul(alpha, V[i])); UQ[i] = cadd(U[i], Q[i]); }
        PRECOND_IN(UQ);
        matvec(n, A, T, V);
        double s = 0.0;
        for (int i = 0; i < n; i++) {
            X[i] = cadd(X[i], cmul(alpha, UQ[i]));
            R[i] = csub(R[i], cmul(alpha, V[i]));
            s += cabs2(R[i]);
        }
        if (j == 0) err0 = sqrt(s); else err = sqrt(s) / err0;
        if (err < tol || j == maxit) break;
        rho1 = dotu(n, R, R0);
        cplx beta = cdiv(rho1, rho0);
        for (int i = 0; i < n; i++) U[i] = cadd(R[i], cmul(beta, Q[i]));
        for (int i = 0; i < n; i++) P[i] = cadd(U[i], cmul(beta, cadd(Q[i], cmul(beta, P[i]))));
    }
#undef PRECOND_IN
    if (precond <= 0) {
        ilu_apply(&ilu, X, T);
        for (int i = 0; i < n; i++) { xout[2 * i] = T[i].re; xout[2 * i + 1] = T[i].im; }
        free(ilu.mask);
        free(ilu.LU);
        free(As);
    } else {
        for (int i = 0; i < n; i++) { cplx v = cscal(X[i], sc[i]); xout[2 * i] = v.re; xout[2 * i + 1] = v.im; }
    }
    if (relerr) *relerr = err;
    free(w);
    free(sc);
    return j;
}

/* Tfactor_usy + Tfbelim (src/NC_TypeDefinition.h:542-594) */
void ncp_lu_solve(int n, double *Ad, double *bd)
{
    cplx *A = (cplx *)Ad, *b = (cplx *)bd;
    for (int i = 0; i < n - 1; i++) {
        cplx dia = A[(size_t)i * n + i];
        for (int j = i + 1; j < n; j++) {
            cplx *Aj = A + (size_t)j * n;
            Aj[i] = cdiv(Aj[i], dia);
            cplx l = Aj[i];
            for (int k = i + 1; k < n; k++) Aj[k] = csub(Aj[k], cmul(A[(size_t)i * n + k], l));
        }
        for (int k = i + 1; k < n; k++) A[(size_t)i * n + k] = cdiv(A[(size_t)i * n + k], dia);
    }
    for (int i = 1; i < n; i++)
        for (int j = 0; j < i; j++) b[i] = csub(b[i], cmul(A[(size_t)i * n + j], b[j]));
    for (int i = 0; i < n; i++) b[i] = cdiv(b[i], A[(size_t)i * n + i]);
    for (int i = n - 1; i > 0; i--)
        for (int j = 0; j < i; j++) b[j] = csub(b[j], cmul(b[i], A[(size_t)j * n + i]));
}

/* ------------------------------------------------------------------ field */

/* NC_ContributionTBEM + NC_IntegrationTBEM + NC_ContributionIncidentWaves
 * (src/NC_PostProcessing.cpp:776-975, 1217-1322) */
int ncp_field(const ncp_problem *p, const double *sol, int m, const double *pts,
              double *phi_out, double *grad_out)
{
    bm_consts c = bm(p);
    double k2 = p->k * p->k;
    for (int q = 0; q < m; q++) {
        const double *x = pts + 3 * q;
        cplx phi = cmk(0, 0), grad[3] = {cmk(0, 0), cmk(0, 0), cmk(0, 0)};
        for (int e = 0; e < p->n; e++) {
            const double *X = p->xyz + 12 * e;
            int nv = p->nv[e];
            cplx pot, vel[4];
            if (p->bc_kind[e] == 1) {   /* pressure BC: potential given, velocity solved */
                pot = cld(p->bc + 8 * e);
                for (int a = 0; a < nv; a++) vel[a] = cld(sol + 2 * e);
            } else {
                pot = cld(sol + 2 * e);
                cplx corr = p->bc_kind[e] == 2 ? cscal(cmul(pot, cld(p->adm + 2 * e)), c.tau) : cmk(0, 0);
                for (int a = 0; a < nv; a++) vel[a] = csub(cld(p->bc + 8 * e + 2 * a), corr);
            }
            int ivd = 0;
            for (int a = 0; a < nv; a++) if (sqrt(cabs2(vel[a])) > EPSY_) ivd = 1;
            double s0[MSBE_], t0[MSBE_], fa[MSBE_];
            int ord[MSBE_];
            int nsb = ncp_subelements(p, e, x, s0, t0, fa, ord);
            if (nsb < 0) return -1;
            cplx ep = cmk(0, 0), eg[3] = {cmk(0, 0), cmk(0, 0), cmk(0, 0)};
            for (int sb = 0; sb < nsb; sb++) {
                double gs[36], gt[36], gw[36];
                int ng = gauss_points(nv, ord[sb], gs, gt, gw);
                int whole = fabs(fa[sb] - 1.0) < EPSY_;
                cplx sp = cmk(0, 0);
                for (int g = 0; g < ng; g++) {
                    double s, t, w;
                    if (whole) { s = gs[g]; t = gt[g]; w = gw[g]; }
                    else { s = s0[sb] + gs[g] * fa[sb]; t = t0[sb] + gt[g] * fa[sb]; w = gw[g] * fa[sb] * fa[sb]; }
                    double P[3], N[4], ny[3];
                    global_point(X, nv, s, t, P, N);
                    double jac = unit_normal(X, nv, s, t, ny);
                    double v[3] = {P[0] - x[0], P[1] - x[1], P[2] - x[2]};
                    double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    v[0] /= r; v[1] /= r; v[2] /= r;
                    double rk = r * c.hk, a1 = r * PI4_;
                    cplx G = cmk(cos(rk) / a1, sin(rk) / a1);
                    cplx Gi = cmul(G, cmk(-1.0 / r, c.hk));
                    double drdn = v[0] * ny[0] + v[1] * ny[1] + v[2] * ny[2];
                    cplx f1 = cmul(G, cmk(drdn * (k2 - 3.0 / r / r), c.hk * 3.0 * drdn / r));
                    cplx f2 = cmk(Gi.re / (-r), Gi.im / (-r));
                    cplx dGdn = cscal(Gi, drdn);
                    double wj = jac * w;
                    cplx z1 = cscal(pot, wj);
                    sp = cadd(sp, cmul(dGdn, z1));
                    for (int d = 0; d < 3; d++)
                        eg[d] = cadd(eg[d], cmul(cadd(cscal(f1, v[d]), cscal(f2, ny[d])), z1));
                    if (ivd)
                        for (int a = 0; a < nv; a++) {
                            cplx z2 = cmul(vel[a], cmk(wj * N[a], 0.0));
                            sp = csub(sp, cmul(G, z2));
                            for (int d = 0; d < 3; d++) eg[d] = csub(eg[d], cmul(cscal(Gi, -v[d]), z2));
                        }
                }
                ep = cadd(ep, sp);
            }
            if (c.tau != 1.0) {
                ep = cscal(ep, c.tau);
                for (int d = 0; d < 3; d++) eg[d] = cscal(eg[d], c.tau);
            }
            phi = cadd(phi, ep);
            for (int d = 0; d < 3; d++) grad[d] = cadd(grad[d], eg[d]);
        }
        /* incident field (src/NC_PostProcessing.cpp:776-866) */
        for (int w = 0; w < p->n_pw; w++) {
            const double *d = p->pw + 5 * w;
            double a = (d[0] * x[0] + d[1] * x[1] + d[2] * x[2]) * c.hk;
            cplx g = cmul(cmk(cos(a), sin(a)), cld(d + 3));
            phi = cadd(phi, g);
            cplx v = cmk(-g.im * c.hk, g.re * c.hk);
            for (int q2 = 0; q2 < 3; q2++) grad[q2] = cadd(grad[q2], cscal(v, d[q2]));
        }
        for (int s = 0; s < p->n_ps; s++) {
            const double *y = p->ps + 5 * s;
            double v2[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
            double R = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
            v2[0] /= R; v2[1] /= R; v2[2] /= R;
            double a = c.hk * R, b = 1.0 / PI4_ / R;
            cplx g = cmul(cmk(cos(a) * b, sin(a) * b), cld(y + 3));
            phi = cadd(phi, g);
            cplx v = cmul(cmk(-1.0 / R, c.hk), g);
            for (int q2 = 0; q2 < 3; q2++) grad[q2] = cadd(grad[q2], cscal(v, v2[q2]));
        }
        phi_out[2 * q] = phi.re;
        phi_out[2 * q + 1] = phi.im;
        for (int d = 0; d < 3; d++) { grad_out[6 * q + 2 * d] = grad[d].re; grad_out[6 * q + 2 * d + 1] = grad[d].im; }
    }
    return 0;
}
