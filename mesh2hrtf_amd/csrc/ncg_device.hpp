// Device building blocks of libncg (gfx950): complex arithmetic, quadrature rules,
// element geometry, NumCalc's adaptive subdivision, and the Burton–Miller kernels.
//
// Reference routines restated here (src/ = mesh2hrtf/NumCalc/src/):
//   NC_ComputeGausseanPoints      src/NC_3dFunctions.cpp:455-530
//   NC_GenerateSubelements        src/NC_3dFunctions.cpp:575-783  (error bounds src/NC_Macros.h:44-50)
//   NC_ComputeParameters          src/NC_3dFunctions.cpp:533-572
//   NC_RegularIntegration kernels src/NC_EquationSystem.cpp:1686-1726
//   NC_SingularIntegration        src/NC_EquationSystem.cpp:1408-1597
//   NC_IntegrationTBEM kernels    src/NC_PostProcessing.cpp:1264-1318
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ncg {

constexpr double kPi = 3.1415926535897932385;
constexpr double kPi4 = 12.566370614359172944;   // 4 pi, src/NC_ConstantsVariables.h:57
constexpr double kEpsy = 1.0e-14;                // src/NC_ConstantsVariables.h:60
constexpr double kTolf = 1.3;                    // subdivision ratio, src/NC_3dFunctions.cpp:598
constexpr int kMaxSubel = 220;                   // MSBE, src/NC_ConstantsVariables.h:56
constexpr int kMaxSplitPerLevel = 15;            // src/NC_3dFunctions.cpp:648
constexpr int kMaxDepth = 24;                    // DFS guard (the reference loops unbounded)

// error flags raised by device code (OR-ed into a device word)
enum : int { kErrSplits = 1, kErrSubel = 2, kErrDepth = 4 };

template <typename T>
struct cx {
    T re, im;
};
template <typename T>
__device__ __forceinline__ cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> operator*(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> &operator+=(cx<T> &a, cx<T> b) { a.re += b.re; a.im += b.im; return a; }
template <typename T>
__device__ __forceinline__ cx<T> &operator-=(cx<T> &a, cx<T> b) { a.re -= b.re; a.im -= b.im; return a; }
template <typename T>
__device__ __forceinline__ cx<T> cdiv(cx<T> a, cx<T> b)
{
    T d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

__device__ __forceinline__ void sincos_t(double x, double *s, double *c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_t(float x, float *s, float *c) { sincosf(x, s, c); }

// ---------------------------------------------------------------------------------- rules
// Triangle rules (s, t, weight), weights summing to 1/2: Strang-Fix 4-point (order 2),
// Bathe 7-point (order 3), 13-point degree-7 rule (orders 4..7) — the published rules
// with the digits NumCalc tabulates (src/NC_IntegrationConstants.h:3758-3865).
// Packed: [0,4) TR4, [4,11) TR7, [11,24) TR13.
__constant__ double c_tri[24][3] = {
    {1.0 / 3.0, 1.0 / 3.0, -27.0 / 96.0}, {0.2, 0.2, 25.0 / 96.0}, {0.6, 0.2, 25.0 / 96.0},
    {0.2, 0.6, 25.0 / 96.0},
    {0.1012865073235, 0.1012865073235, 0.1259391805448 / 2.0},
    {0.7974269853531, 0.1012865073235, 0.1259391805448 / 2.0},
    {0.1012865073235, 0.7974269853531, 0.1259391805448 / 2.0},
    {0.4701420641051, 0.0597158717898, 0.1323941527885 / 2.0},
    {0.4701420641051, 0.4701420641051, 0.1323941527885 / 2.0},
    {0.0597158717898, 0.4701420641051, 0.1323941527885 / 2.0},
    {0.3333333333333, 0.3333333333333, 0.225 / 2.0},
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

// Gauss-Legendre rules of 1..8 points packed: rule n starts at n(n-1)/2.  Filled by the
// host at library load from a Newton iteration (the reference tabulates them,
// src/NC_IntegrationConstants.h:24-1900).
__constant__ double c_gl_x[36];
__constant__ double c_gl_w[36];

__device__ __forceinline__ int tri_rule_offset(int order) { return order <= 1 ? -1 : order == 2 ? 0 : order == 3 ? 4 : 11; }
__device__ __forceinline__ int tri_rule_count(int order) { return order <= 1 ? 1 : order == 2 ? 4 : order == 3 ? 7 : 13; }

// Gauss order of a stored subelement: smallest m in [2, 6] with the G, H and E error
// estimates below 1e-3 at disfac = 0.5/ratdis (src/NC_3dFunctions.cpp:757-769,
// src/NC_Macros.h:44-50).
__device__ __forceinline__ int gauss_order(double ratdis)
{
    const double x = 0.5 / ratdis;
    const double x2 = x * x;
    double p = x * x2 * x2;    // x^5 = x^(2m+1) at m = 2
    for (int m = 2; m <= 6; ++m) {
        const double eg = 32.0 * p;
        const double eh = 64.0 * (2.0 * m + 1.0) * p * x;
        const double ee = 128.0 * (m + 1.0) * (2.0 * m + 1.0) * p * x2;
        if (ee < 1e-3 && eh < 1e-3 && eg < 1e-3) return m;
        p *= x2;
    }
    return 6;
}

// ---------------------------------------------------------------------------------- geometry
// Element data, structure of arrays of length n (frequency independent, resident in HBM).
struct Elements {
    int n;
    const double *X;        // [12][n]: node a, component d at X[(3a+d)*n + e]
    const int8_t *nv;       // [n] 3 or 4
    const double *cen;      // [3][n] collocation point (element centre)
    const double *nrm;      // [3][n] unit normal at the centre
    const double *area;     // [n]
    const double *jac;      // [n] |dx/ds x dx/dt| at the centre (constant on a flat triangle)
    const int8_t *kind;     // [n] NCG_BC_*
};

struct Vec3 {
    double x, y, z;
};

__device__ __forceinline__ Vec3 elem_node(const Elements &E, int e, int a)
{
    return {E.X[(3 * a + 0) * E.n + e], E.X[(3 * a + 1) * E.n + e], E.X[(3 * a + 2) * E.n + e]};
}

// Local element geometry relative to a field point x (FP64): node 2 minus x and the
// triangle edge vectors, or the four quad nodes minus x.
struct LocalElem {
    int nv;
    double p[4][3];   // nodes relative to x
    double ny[3];     // unit normal (triangle: constant)
    double jac;       // triangle Jacobian (constant)
    double area;
};

__device__ __forceinline__ void load_local(const Elements &E, int e, const double *x, LocalElem &L)
{
    L.nv = E.nv[e];
    for (int a = 0; a < L.nv; ++a)
        for (int d = 0; d < 3; ++d) L.p[a][d] = E.X[(3 * a + d) * E.n + e] - x[d];
    L.ny[0] = E.nrm[e];
    L.ny[1] = E.nrm[E.n + e];
    L.ny[2] = E.nrm[2 * E.n + e];
    L.jac = E.jac[e];
    L.area = E.area[e];
}

// shape functions, global point (relative to x), unit normal and Jacobian at (s, t)
// (NC_ComputeGlobalCoordinates / NC_ComputeUnitVectorOnElement, src/NC_3dFunctions.cpp:95-178)
__device__ __forceinline__ void point_on(const LocalElem &L, double s, double t, double *q, double *ny, double &jac)
{
    if (L.nv == 3) {
        const double u = 1.0 - s - t;
        for (int d = 0; d < 3; ++d) q[d] = s * L.p[0][d] + t * L.p[1][d] + u * L.p[2][d];
        ny[0] = L.ny[0]; ny[1] = L.ny[1]; ny[2] = L.ny[2];
        jac = L.jac;
        return;
    }
    const double s1 = 0.25 * (s + 1.0), s2 = 0.25 * (s - 1.0), t1 = t + 1.0, t2 = t - 1.0;
    const double N[4] = {s1 * t1, -s2 * t1, s2 * t2, -s1 * t2};
    const double ds[4] = {0.25 * t1, -0.25 * t1, 0.25 * t2, -0.25 * t2};
    const double dt[4] = {s1, -s2, s2, -s1};
    double xs[3], xt[3];
    for (int d = 0; d < 3; ++d) {
        q[d] = N[0] * L.p[0][d] + N[1] * L.p[1][d] + N[2] * L.p[2][d] + N[3] * L.p[3][d];
        xs[d] = ds[0] * L.p[0][d] + ds[1] * L.p[1][d] + ds[2] * L.p[2][d] + ds[3] * L.p[3][d];
        xt[d] = dt[0] * L.p[0][d] + dt[1] * L.p[1][d] + dt[2] * L.p[2][d] + dt[3] * L.p[3][d];
    }
    double n0 = xs[1] * xt[2] - xs[2] * xt[1], n1 = xs[2] * xt[0] - xs[0] * xt[2], n2 = xs[0] * xt[1] - xs[1] * xt[0];
    jac = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    ny[0] = n0 / jac; ny[1] = n1 / jac; ny[2] = n2 / jac;
}

// ratio distance(centre, x) / sqrt(area of the (sub)element) — the quantity that
// decides subdivision.  Evaluated without contraction so that every kernel that
// classifies a pair takes the same decision.
__device__ __forceinline__ double ratio_to_centre(double cx_, double cy_, double cz_, const double *x, double area)
{
#pragma clang fp contract(off)
    const double dx = cx_ - x[0], dy = cy_ - x[1], dz = cz_ - x[2];
    return sqrt(dx * dx + dy * dy + dz * dz) / sqrt(area);
}

__device__ __forceinline__ bool pair_is_near(const Elements &E, int j, const double *x)
{
    return ratio_to_centre(E.cen[j], E.cen[E.n + j], E.cen[2 * E.n + j], x, E.area[j]) < kTolf;
}

// ---------------------------------------------------------------------------------- quadrature visitor
// Visit every quadrature point (s, t, weight) of the adaptive subdivision of an element
// seen from x.  The reference builds the same quadtree breadth first
// (NC_GenerateSubelements); leaves and their Gauss orders depend only on each
// subelement, so a depth-first walk with an explicit stack visits the same points.
// Sub-triangles are kept as (origin, signed scale): the reference's "vertex with the
// right angle" and +-faclin.  Coordinates are dyadic, stored exactly as integers
// scaled by 2^-depth.  Returns 0 or a kErr* flag.
template <typename F>
__device__ int visit_adaptive(const LocalElem &L, F &&fn)
{
    struct Node {
        int32_t a, b;     // origin (triangle) or centre (quad) in units of 2^-(depth+1)
        int32_t depth;    // sign carries the triangle orientation
    };
    Node stack[3 * kMaxDepth + 4];
    int split_count[kMaxDepth + 1];
    for (int l = 0; l <= kMaxDepth; ++l) split_count[l] = 0;
    int sp = 0, leaves = 0;
    const bool tri = L.nv == 3;
    // root: triangle origin (0,0) scale +1; quad centre (0,0) half width 1
    stack[sp++] = {0, 0, 0};
    const double zero[3] = {0.0, 0.0, 0.0};
    while (sp > 0) {
        const Node nd = stack[--sp];
        const int depth = nd.depth < 0 ? -nd.depth - 1 : nd.depth;
        const double sgn = nd.depth < 0 ? -1.0 : 1.0;
        const double unit = ldexp(1.0, -(depth + 1));    // coordinates are a*unit
        const double f = ldexp(1.0, -depth);              // faclin
        const double s0 = nd.a * unit, t0 = nd.b * unit;
        // centre of the subelement
        double sc, tc;
        if (tri) { sc = s0 + sgn * f / 3.0; tc = t0 + sgn * f / 3.0; }
        else { sc = s0; tc = t0; }
        double q[3], ny[3], jac;
        point_on(L, sc, tc, q, ny, jac);
        const double ratdis = ratio_to_centre(q[0], q[1], q[2], zero, L.area * f * f);
        if (ratdis < kTolf) {
            if (depth + 1 > kMaxDepth) return kErrDepth;
            if (++split_count[depth] > kMaxSplitPerLevel) return kErrSplits;
            const int cd = depth + 1;
            // children in units of 2^-(cd+1): parent coordinates double
            const int32_t A = 2 * nd.a, B = 2 * nd.b;
            if (tri) {
                // parent (s0,t0,g) with g = sgn*f: children (s0+g/2,t0,g/2), (s0,t0+g/2,g/2),
                // (s0,t0,g/2), (s0+g/2,t0+g/2,-g/2); g/2 = sgn * 2^-(cd) = sgn*2 units
                const int32_t h = nd.depth < 0 ? -2 : 2;
                const int32_t dp = nd.depth < 0 ? -cd - 1 : cd, dm = nd.depth < 0 ? cd : -cd - 1;
                stack[sp++] = {A + h, B, dp};
                stack[sp++] = {A, B + h, dp};
                stack[sp++] = {A, B, dp};
                stack[sp++] = {A + h, B + h, dm};
            } else {
                // quad centre (sc,tc), half width f: children centres +-f/2 = +-2 units
                stack[sp++] = {A + 2, B + 2, cd};
                stack[sp++] = {A - 2, B + 2, cd};
                stack[sp++] = {A - 2, B - 2, cd};
                stack[sp++] = {A + 2, B - 2, cd};
            }
            continue;
        }
        if (++leaves > kMaxSubel) return kErrSubel;
        const int m = gauss_order(ratdis);
        const bool whole = depth == 0;
        if (tri) {
            const int off = tri_rule_offset(m), cnt = tri_rule_count(m);
            const double g = sgn * f;
            for (int k = 0; k < cnt; ++k) {
                const double gs = c_tri[off + k][0], gt = c_tri[off + k][1], gw = c_tri[off + k][2];
                if (whole) fn(gs, gt, gw);
                else fn(s0 + gs * g, t0 + gt * g, gw * g * g);
            }
        } else {
            const int base = m * (m - 1) / 2;
            for (int u = 0; u < m; ++u)
                for (int v = 0; v < m; ++v) {
                    const double gs = c_gl_x[base + u], gt = c_gl_x[base + v], gw = c_gl_w[base + u] * c_gl_w[base + v];
                    if (whole) fn(gs, gt, gw);
                    else fn(s0 + gs * f, t0 + gt * f, gw * f * f);
                }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------- kernels
// Element integrals of the Burton–Miller collocation at x with normal nx: G, H (dG/dn_y),
// HT (dG/dn_x), E (d2G/dn_x dn_y), G = e^{ihkr}/(4 pi r)
// (NC_RegularIntegration, src/NC_EquationSystem.cpp:1686-1717).
template <typename T>
struct BMAcc {
    cx<T> G, H, HT, E;
};

// q = Gauss point minus x, ny its unit normal, wga = weight * Jacobian
template <typename T>
__device__ __forceinline__ void bm_point(BMAcc<T> &acc, T qx, T qy, T qz, T nyx, T nyy, T nyz, T nxx, T nxy, T nxz,
                                         T wga, T hk, T k2)
{
    T r, ir, sn, cs;
    if constexpr (sizeof(T) == 4) {
        // FP32 (the mixed form's regular quadrature): the hardware reciprocal square root
        // and sine/cosine (v_rsq_f32, v_sin_f32, v_cos_f32; the latter accurate to about
        // 1e-6 absolute for arguments within 256 periods, while h k r stays below ~120 at
        // 24 kHz over a 0.25 m mesh); errors of the order of the FP32 quadrature's own
        const T r2 = qx * qx + qy * qy + qz * qz;
        ir = __builtin_amdgcn_rsqf(r2);
        r = r2 * ir;
        __sincosf(hk * r, &sn, &cs);
    } else {
        r = sqrt(qx * qx + qy * qy + qz * qz);
        ir = T(1) / r;
        sincos_t(hk * r, &sn, &cs);
    }
    const T dx = qx * ir, dy = qy * ir, dz = qz * ir;
    const T a2 = wga * T(1.0 / kPi4) * ir;
    const cx<T> g = {cs * a2, sn * a2};
    const cx<T> gd = {-ir * g.re - hk * g.im, -ir * g.im + hk * g.re};
    const T rny = dx * nyx + dy * nyy + dz * nyz;
    const T rnx = -(dx * nxx + dy * nxy + dz * nxz);
    const T rq = rny * rnx, nn = nxx * nyx + nxy * nyy + nxz * nyz;
    const T ir2 = ir * ir;
    const cx<T> z1 = {(T(3) * ir2 - k2) * rq + nn * ir2, -hk * ir * (T(3) * rq + nn)};
    acc.G += g;
    acc.H += gd * rny;
    acc.HT += gd * rnx;
    acc.E += g * z1;
}

// Singular (self) integrals of element j at its own centre, FP64
// (NC_SingularIntegration, src/NC_EquationSystem.cpp:1408-1597).
__device__ inline void bm_singular(const LocalElem &L, const double *nx, double hk, double k2, BMAcc<double> &acc,
                                   unsigned long long *nev = nullptr)
{
    acc.G = acc.H = acc.HT = acc.E = mk(0.0, 0.0);
    const int nv = L.nv;
    // local node coordinates (src/NC_Arrays.h:126-129): corners then edge midpoints
    const double CS6[6] = {1.0, 0.0, 0.0, 0.5, 0.0, 0.5}, ET6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
    const double CS8[8] = {1.0, -1.0, -1.0, 1.0, 0.0, -1.0, 0.0, 1.0}, ET8[8] = {1.0, 1.0, -1.0, -1.0, 1.0, 0.0, -1.0, 0.0};
    for (int eg = 0; eg < nv; ++eg) {
        const int e1 = eg + 1 == nv ? 0 : eg + 1, e2 = eg + nv;
        double dv[3], tv[3], len = 0.0;
        for (int d = 0; d < 3; ++d) { dv[d] = L.p[e1][d] - L.p[eg][d]; len += dv[d] * dv[d]; }
        len = sqrt(len);
        for (int d = 0; d < 3; ++d) tv[d] = dv[d] / len;
        len /= 8.0;   // 4 sections, parameter range 2
        // hypersingular part: line integral of (grad_y G x t) . n_x along the edge
        cx<double> accE = mk(0.0, 0.0);
        double mid = -1.25;
        for (int sec = 0; sec < 4; ++sec) {
            mid += 0.5;
            for (int g = 0; g < 3; ++g) {
                const double sg = mid + c_gl_x[3 + g] / 4.0, wg = c_gl_w[3 + g] * len;
                double d[3];
                for (int q = 0; q < 3; ++q) d[q] = L.p[eg][q] + dv[q] * (sg + 1.0) / 2.0;
                const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                d[0] /= r; d[1] /= r; d[2] /= r;
                double sn, cs;
                sincos(hk * r, &sn, &cs);
                const double a2 = kPi4 * r;
                const cx<double> G = mk(cs / a2, sn / a2) * mk(-1.0 / r, hk);
                const double w0 = d[1] * tv[2] - d[2] * tv[1], w1 = d[2] * tv[0] - d[0] * tv[2],
                             w2 = d[0] * tv[1] - d[1] * tv[0];
                accE += G * ((w0 * nx[0] + w1 * nx[1] + w2 * nx[2]) * wg);
            }
        }
        if (nev) *nev += 12;   // 4 sections x 3 points
        acc.E += accE;
        // weakly singular parts: two collapsed (Duffy) sub-triangles per edge, 4x4 Gauss
        for (int sec = 0; sec < 2; ++sec) {
            double ss[3], tt[3], ares;
            if (nv == 3) {
                ares = 1.0 / 48.0;
                ss[0] = tt[0] = 1.0 / 3.0;
                if (sec == 0) { ss[1] = CS6[eg]; ss[2] = CS6[e2]; tt[1] = ET6[eg]; tt[2] = ET6[e2]; }
                else { ss[1] = CS6[e2]; ss[2] = CS6[e1]; tt[1] = ET6[e2]; tt[2] = ET6[e1]; }
            } else {
                ares = 0.125;
                ss[0] = tt[0] = 0.0;
                if (sec == 0) { ss[1] = CS8[eg]; ss[2] = CS8[e2]; tt[1] = ET8[eg]; tt[2] = ET8[e2]; }
                else { ss[1] = CS8[e2]; ss[2] = CS8[e1]; tt[1] = ET8[e2]; tt[2] = ET8[e1]; }
            }
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    const double sg = c_gl_x[6 + a], tg = c_gl_x[6 + b], wq = c_gl_w[6 + a] * c_gl_w[6 + b];
                    const double s = 0.5 * (1.0 - sg) * ss[0] + 0.25 * (1.0 + sg) * ((1.0 - tg) * ss[1] + (1.0 + tg) * ss[2]);
                    const double t = 0.5 * (1.0 - sg) * tt[0] + 0.25 * (1.0 + sg) * ((1.0 - tg) * tt[1] + (1.0 + tg) * tt[2]);
                    double q[3], ny[3], jac;
                    point_on(L, s, t, q, ny, jac);
                    const double wga = wq * (1.0 + sg) * ares * jac;
                    const double r = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                    const double dx = q[0] / r, dy = q[1] / r, dz = q[2] / r;
                    double sn, cs;
                    sincos(hk * r, &sn, &cs);
                    const double a2 = wga / kPi4 / r;
                    const cx<double> g = mk(cs * a2, sn * a2);
                    const cx<double> gd = g * mk(-1.0 / r, hk);
                    const double rny = dx * ny[0] + dy * ny[1] + dz * ny[2];
                    const double rnx = -(dx * nx[0] + dy * nx[1] + dz * nx[2]);
                    acc.G += g;
                    acc.H += gd * rny;
                    acc.HT += gd * rnx;
                    acc.E += g * (k2 * (nx[0] * ny[0] + nx[1] * ny[1] + nx[2] * ny[2]));
                }
            if (nev) *nev += 16;   // 4 x 4 points
        }
    }
}

// FP64 integrals of any pair (regular, subdivided or singular) — used for the near
// pairs, the right-hand-side source columns and the diagonal.  Returns a kErr* flag.
__device__ inline int bm_pair(const Elements &E, int i, int j, double hk, double k2, BMAcc<double> &acc,
                              unsigned long long *nev = nullptr)
{
    const double x[3] = {E.cen[i], E.cen[E.n + i], E.cen[2 * E.n + i]};
    const double nx[3] = {E.nrm[i], E.nrm[E.n + i], E.nrm[2 * E.n + i]};
    LocalElem L;
    load_local(E, j, x, L);
    if (i == j) {
        bm_singular(L, nx, hk, k2, acc, nev);
        return 0;
    }
    acc.G = acc.H = acc.HT = acc.E = mk(0.0, 0.0);
    return visit_adaptive(L, [&](double s, double t, double w) {
        double q[3], ny[3], jac;
        point_on(L, s, t, q, ny, jac);
        bm_point<double>(acc, q[0], q[1], q[2], ny[0], ny[1], ny[2], nx[0], nx[1], nx[2], w * jac, hk, k2);
        if (nev) ++*nev;
    });
}

// Burton–Miller coefficients (src/NC_EquationSystem.cpp:85-86): gamma = 1,
// beta = i h / k exterior (0 interior), tau = +1 exterior (-1 interior).
struct BMConst {
    double tau;
    double beta_im;   // beta = (0, beta_im)
    double hk, k2, k;
};

// Matrix entry of column j from its integrals (NC_AssembleTBEM, src/NC_EquationSystem.cpp:165-207)
template <typename T>
__device__ __forceinline__ cx<T> bm_entry(const BMAcc<T> &a, int kind, cx<T> adm, T tau, T beta_im)
{
    const cx<T> beta = {T(0), beta_im};
    if (kind == 1) return a.G * (-tau) - a.HT * beta;
    cx<T> s = a.H * tau + a.E * beta;
    if (kind == 2) s += (a.G + a.HT * beta * tau) * adm;
    return s;
}

// RHS coefficient of column j per unit prescribed value (the zrsintel[4] terms,
// src/NC_EquationSystem.cpp:1570-1596, 1719-1725)
template <typename T>
__device__ __forceinline__ cx<T> bm_rhs_coef(const BMAcc<T> &a, int kind, T tau, T beta_im)
{
    const cx<T> beta = {T(0), beta_im};
    if (kind == 1) return mk(T(0), T(0)) - (a.H * tau + a.E * beta);
    return a.G * tau + a.HT * beta;
}

}  // namespace ncg
