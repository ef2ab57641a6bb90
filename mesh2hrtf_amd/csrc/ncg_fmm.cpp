// Host side of the FMM plan: NumCalc's clustering rules restated (see ncg_fmm.hpp).
#include "ncg_fmm.hpp"

#include <math.h>

#include <algorithm>
#include <array>

namespace ncgf {

namespace {

constexpr double kEps = 1.0e-6;    // epsilon_, src/NC_Input.cpp:342
constexpr double kPi = 3.1415926535897932385;
constexpr double kPi4 = 12.566370614359172944;

double dist(const double *a, const double *b)
{
    return sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
}

// unique node ids of a list of elements, first-appearance order (a per-thread stamp
// array marks the ids already seen)
void unique_nodes(const Mesh &M, const std::vector<int> &elems, std::vector<int> &nodes, std::vector<const double *> &xyz)
{
    thread_local std::vector<unsigned> seen;
    thread_local unsigned stamp = 0;
    if (++stamp == 0) {
        std::fill(seen.begin(), seen.end(), 0u);
        stamp = 1;
    }
    nodes.clear();
    xyz.clear();
    for (int e : elems)
        for (int a = 0; a < M.nv[e]; ++a) {
            const int id = M.node[4 * (size_t)e + a];
            if (id >= (int)seen.size()) seen.resize((size_t)id + 1 + id / 2, 0u);
            if (seen[id] == stamp) continue;
            seen[id] = stamp;
            nodes.push_back(id);
            xyz.push_back(&M.X[12 * (size_t)e + 3 * a]);
        }
}

// centre = mean of the unique nodes, radius = largest node distance
// (src/NC_Addresses.cpp:1007-1031)
void centre_radius(const Mesh &M, const std::vector<int> &elems, double *c, double &r)
{
    std::vector<int> nodes;
    std::vector<const double *> xyz;
    unique_nodes(M, elems, nodes, xyz);
    c[0] = c[1] = c[2] = 0.0;
    for (const double *x : xyz)
        for (int d = 0; d < 3; ++d) c[d] += x[d];
    for (int d = 0; d < 3; ++d) c[d] /= (double)xyz.size();
    r = 0.0;
    for (const double *x : xyz) r = std::max(r, dist(c, x));
}

// NC_CancelSmallClusters (src/NC_Addresses.cpp:2077-2244): clusters whose radius is
// below a third of the mean radius are merged into the nearest regular cluster.
void cancel_small(const Mesh &M, std::vector<std::vector<int>> &cl)
{
    const int nc = (int)cl.size();
    std::vector<double> rad(nc, 0.0);
    std::vector<std::array<double, 3>> cen(nc, {0.0, 0.0, 0.0});
    for (int i = 0; i < nc; ++i) {
        if (cl[i].empty()) continue;
        double c[3], r;
        centre_radius(M, cl[i], c, r);
        cen[i] = {c[0], c[1], c[2]};
        rad[i] = r;
    }
    double rmin = 1.0e20, ravg = 0.0;
    int cnt = 0;
    for (int i = 0; i < nc; ++i) {
        if (cl[i].empty()) continue;
        ravg += rad[i];
        cnt++;
        rmin = std::min(rmin, rad[i]);
    }
    ravg /= (double)cnt;
    const double f = 3.0;
    if (!(ravg > f * rmin)) return;
    std::vector<int> target(nc, -1);
    for (int i = 0; i < nc; ++i) {
        if (cl[i].empty() || f * rad[i] >= ravg) continue;
        double dmin = 1.0e20;
        for (int j = 0; j < nc; ++j) {
            if (i == j || cl[j].empty() || f * rad[j] < ravg) continue;
            const double d = dist(cen[i].data(), cen[j].data());
            if (d < dmin) { dmin = d; target[i] = j; }
        }
    }
    for (int i = 0; i < nc; ++i) {
        if (target[i] < 0) continue;
        std::vector<int> &dst = cl[target[i]];
        dst.insert(dst.end(), cl[i].begin(), cl[i].end());
        cl[i].clear();
    }
}

// number of grid divisions along one axis (src/NC_Addresses.cpp:697-726)
int divisions(double extent, double edge, bool capped)
{
    const double d0 = extent / edge;
    int i = (int)d0;
    if (i == 0) return 1;
    if (d0 - (double)i > 0.5) i++;
    if (capped && i > 2) i = 2;
    return i;
}

// NC_GenerateClustersFMM for one father group: grid boxes over the group's bounding box
std::vector<std::vector<int>> split_group(const Mesh &M, const std::vector<int> &elems, double edge, bool capped,
                                          std::string &err)
{
    std::vector<int> nodes;
    std::vector<const double *> xyz;
    unique_nodes(M, elems, nodes, xyz);
    double mx[3] = {-1.0e20, -1.0e20, -1.0e20}, mn[3] = {1.0e20, 1.0e20, 1.0e20};
    for (const double *x : xyz)
        for (int d = 0; d < 3; ++d) {
            if (x[d] > mx[d]) mx[d] = x[d];
            if (x[d] < mn[d]) mn[d] = x[d];
        }
    int nd[3];
    double dif[3];
    for (int d = 0; d < 3; ++d) {
        mx[d] += M.delta;
        mn[d] -= M.delta;
        dif[d] = mx[d] - mn[d];
        nd[d] = divisions(dif[d], edge, capped);
        dif[d] /= (double)nd[d];
    }
    const int ne = (int)elems.size();
    std::vector<int> idx(3 * (size_t)ne, -1);
    for (int d = 0; d < 3; ++d) {
        double d1 = mn[d];
        for (int k = 0; k < nd[d]; ++k) {
            const double d0 = d1;
            d1 += dif[d];
            for (int i = 0; i < ne; ++i) {
                const double c = M.cen[3 * (size_t)elems[i] + d];
                if (idx[3 * (size_t)i + d] == -1 && c > d0 - kEps && c < d1 + kEps) idx[3 * (size_t)i + d] = k;
            }
        }
    }
    for (int i = 0; i < ne; ++i)
        for (int d = 0; d < 3; ++d)
            if (idx[3 * (size_t)i + d] == -1) err = "The interval number of an element is not found!";
    std::vector<std::vector<int>> out;
    for (int a = 0; a < nd[0]; ++a)
        for (int b = 0; b < nd[1]; ++b)
            for (int c = 0; c < nd[2]; ++c) {
                std::vector<int> cl;
                for (int i = 0; i < ne; ++i)
                    if (idx[3 * (size_t)i] == a && idx[3 * (size_t)i + 1] == b && idx[3 * (size_t)i + 2] == c)
                        cl.push_back(elems[i]);
                out.push_back(cl);
            }
    return out;
}

void finish_level(const Mesh &M, Level &L)
{
    L.rmax = L.ravg = 0.0;
    L.rmin = 1.0e40;
    for (Cluster &c : L.cl) {
        centre_radius(M, c.elems, c.cen, c.radius);
        L.rmax = std::max(L.rmax, c.radius);
        L.rmin = std::min(L.rmin, c.radius);
        L.ravg += c.radius;
    }
    L.ravg /= (double)L.cl.size();
}

// Near lists, interaction lists and the cancelled-pair counts of one level
// (src/NC_Addresses.cpp:1046-1083; interaction lists NC_AllocateDmtxMLFMM,
// src/NC_Addresses.cpp:1699-1767).  Clusters i, j are far when |c_i - c_j| > (2/sqrt 3)
// (R_i + R_j) and > min_dist; geometrically far pairs closer than min_dist are cancelled
// and stay near.  Only pairs closer than h = max(2 (2/sqrt 3) R_max, min_dist) (with a
// margin) can be near or cancelled, so each cluster tests the clusters in the 27 grid
// cells of side h around its centre; every other pair is far.  The interaction list of
// a cluster is its far clusters (level 0 and single level) or its far clusters whose
// fathers are near its father (the children of the father's near clusters, ascending).
void lists(Level &L, double min_dist, const Level *F, double *dfar = nullptr)
{
    const int nc = (int)L.cl.size();
    const double fac = 2.0 / sqrt(3.0);
    double rmax = 0.0, mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (const Cluster &c : L.cl) {
        rmax = std::max(rmax, c.radius);
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], c.cen[d]);
            mx[d] = std::max(mx[d], c.cen[d]);
        }
    }
    const double h = 1.001 * std::max(2.0 * fac * rmax, min_dist) + 1e-300;
    long long dims[3], cells = 1;
    for (int d = 0; d < 3; ++d) {
        dims[d] = std::min<long long>((long long)((mx[d] - mn[d]) / h) + 1, 1 << 20);
        cells *= dims[d];
    }
    const bool grid = cells <= 8LL * nc + 64;
    auto cell_of = [&](const double *c, int d) {
        return grid ? std::min<long long>((long long)((c[d] - mn[d]) / h), dims[d] - 1) : 0;
    };
    // clusters bucketed by cell (one cell holding every cluster without the grid)
    std::vector<int> start((size_t)(grid ? cells : 1) + 1, 0), item(nc);
    std::vector<long long> cid(nc);
    for (int i = 0; i < nc; ++i) {
        const double *c = L.cl[i].cen;
        cid[i] = grid ? (cell_of(c, 0) * dims[1] + cell_of(c, 1)) * dims[2] + cell_of(c, 2) : 0;
        start[cid[i] + 1]++;
    }
    for (size_t q = 1; q < start.size(); ++q) start[q] += start[q - 1];
    {
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int i = 0; i < nc; ++i) item[fill[cid[i]]++] = i;
    }
    long long geom_near = 0, canceled = 0;   // pairs j < i
    unsigned long long cancel_hash = 0;
    // pairs outside the searched cells are farther apart than h
    double dfar_seen = h;
    for (int i = 0; i < nc; ++i) {
        std::vector<int> &nr = L.cl[i].near;
        nr.clear();
        const double *ci = L.cl[i].cen;
        long long lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            const long long k = cell_of(ci, d);
            lo[d] = grid ? std::max(0LL, k - 1) : 0;
            hi[d] = grid ? std::min(dims[d] - 1, k + 1) : 0;
        }
        for (long long a = lo[0]; a <= hi[0]; ++a)
            for (long long b = lo[1]; b <= hi[1]; ++b)
                for (long long c = lo[2]; c <= hi[2]; ++c) {
                    const long long q = grid ? (a * dims[1] + b) * dims[2] + c : 0;
                    for (int t = start[q]; t < start[q + 1]; ++t) {
                        const int j = item[t];
                        if (j == i) {
                            nr.push_back(j);
                            continue;
                        }
                        const double dd = dist(ci, L.cl[j].cen);
                        const double w = L.cl[i].radius + L.cl[j].radius;
                        const bool far_geom = dd > fac * w;
                        const bool far = far_geom && dd > min_dist;
                        if (far_geom) dfar_seen = std::min(dfar_seen, dd);
                        if (!far) nr.push_back(j);
                        if (j < i) {
                            if (!far_geom) geom_near++;
                            else if (!far) {
                                canceled++;
                                // order-independent digest of the cancelled pairs
                                unsigned long long v = ((unsigned long long)i << 32) ^ (unsigned long long)j;
                                v *= 0x9E3779B97F4A7C15ull;
                                cancel_hash += v ^ (v >> 31);
                            }
                        }
                    }
                }
        std::sort(nr.begin(), nr.end());
    }
    if (dfar) *dfar = dfar_seen;
    L.n_far_geom = (long long)nc * (nc - 1) / 2 - geom_near;
    L.n_canceled = canceled;
    L.cancel_hash = cancel_hash;
    // interaction lists
    std::vector<int> mark(nc, -1);
    std::vector<int> child_begin, child_end;
    if (F) {
        child_begin.assign(F->cl.size(), 0);
        child_end.assign(F->cl.size(), 0);
        for (int j = nc - 1; j >= 0; --j) child_begin[L.cl[j].father] = j;
        for (int j = 0; j < nc; ++j) child_end[L.cl[j].father] = j + 1;
    }
    for (int i = 0; i < nc; ++i) {
        Cluster &c = L.cl[i];
        for (int j : c.near) mark[j] = i;
        c.fan.clear();
        if (!F) {
            for (int j = 0; j < nc; ++j)
                if (mark[j] != i) c.fan.push_back(j);
        } else {
            for (int f : F->cl[c.father].near)
                for (int j = child_begin[f]; j < child_end[f]; ++j)
                    if (mark[j] != i) c.fan.push_back(j);
        }
    }
}

// expansion length and unit-sphere quadrature of a level (src/NC_Main.cpp:447-489,
// NC_ComputeUnitVectorOnUnitSphere src/NC_3dFunctions.cpp:786-818)
bool expansion(Level &L, double k, int min_terms, std::string &err)
{
    const double rw = 2.0 * L.rmax * k + 1.8 * log10(2.0 * L.rmax * k + kPi);
    int t = (int)rw;
    if (rw - (double)t >= 0.5) t++;
    if (t < min_terms) t = min_terms;
    if (t > 80) {
        err = "Too many integral points in the theta-direction!";
        return false;
    }
    L.L = L.ntheta = t;
    L.nphi = 2 * t;
    L.P = t * 2 * t;
    std::vector<double> x(t), w(t);
    gauss_legendre(t, x.data(), w.data());
    const double dphi = 2.0 * kPi / (double)L.nphi;
    L.u.assign(3 * (size_t)L.P, 0.0);
    L.w.assign(L.P, 0.0);
    int q = 0;
    for (int i = 0; i < t; ++i) {
        const double ct = x[i], st = sqrt(1.0 - ct * ct);
        for (int j = 0; j < L.nphi; ++j, ++q) {
            L.w[q] = w[i] * dphi / kPi4;
            const double phi = dphi * (double)j;
            L.u[3 * q] = st * cos(phi);
            L.u[3 * q + 1] = st * sin(phi);
            L.u[3 * q + 2] = ct;
        }
    }
    return true;
}

// NC_GenerateClustersEvalGrid (src/NC_Addresses.cpp:1134-1314)
bool eval_clusters(const Mesh &M, Plan &P, std::string &err)
{
    P.ev.clear();
    if (M.m == 0) return true;
    double mx[3] = {-1.0e20, -1.0e20, -1.0e20}, mn[3] = {1.0e20, 1.0e20, 1.0e20}, dif[3];
    int nd[3];
    for (int p = 0; p < M.m; ++p)
        for (int d = 0; d < 3; ++d) {
            mx[d] = std::max(mx[d], M.eval[3 * (size_t)p + d]);
            mn[d] = std::min(mn[d], M.eval[3 * (size_t)p + d]);
        }
    for (int d = 0; d < 3; ++d) {
        mx[d] += M.delta;
        mn[d] -= M.delta;
        dif[d] = mx[d] - mn[d];
        nd[d] = divisions(dif[d], P.clus_edge, false);
        dif[d] /= (double)nd[d];
    }
    const long long nclus = (long long)nd[0] * nd[1] * nd[2];
    if (nclus > M.m) {
        for (int p = 0; p < M.m; ++p) {
            EvalCluster c;
            c.pts.push_back(p);
            for (int d = 0; d < 3; ++d) c.cen[d] = M.eval[3 * (size_t)p + d];
            c.radius = 0.0;
            P.ev.push_back(c);
        }
        P.ev_rmax = P.ev_rmin = 0.0;
        return true;
    }
    std::vector<int> idx(3 * (size_t)M.m, -1);
    for (int d = 0; d < 3; ++d) {
        double d0 = mn[d];
        for (int k = 0; k < nd[d]; ++k) {
            const double d1 = d0 + dif[d];
            for (int p = 0; p < M.m; ++p) {
                const double c = M.eval[3 * (size_t)p + d];
                if (idx[3 * (size_t)p + d] == -1 && c >= d0 && c < d1) idx[3 * (size_t)p + d] = k;
            }
            d0 = d1;
        }
    }
    for (int p = 0; p < M.m; ++p)
        for (int d = 0; d < 3; ++d)
            if (idx[3 * (size_t)p + d] == -1) {
                err = "The interval number of a internal point is not found!";
                return false;
            }
    P.ev_rmax = 0.0;
    P.ev_rmin = 1.0e15;
    for (int a = 0; a < nd[0]; ++a)
        for (int b = 0; b < nd[1]; ++b)
            for (int c = 0; c < nd[2]; ++c) {
                EvalCluster e;
                for (int p = 0; p < M.m; ++p)
                    if (idx[3 * (size_t)p] == a && idx[3 * (size_t)p + 1] == b && idx[3 * (size_t)p + 2] == c) e.pts.push_back(p);
                if (e.pts.empty()) continue;
                double s[3] = {0, 0, 0};
                for (int p : e.pts)
                    for (int d = 0; d < 3; ++d) s[d] += M.eval[3 * (size_t)p + d];
                for (int d = 0; d < 3; ++d) e.cen[d] = s[d] / (double)e.pts.size();
                double r = 0.0;
                for (int p : e.pts) {
                    double q = 0.0;
                    for (int d = 0; d < 3; ++d) {
                        const double t = M.eval[3 * (size_t)p + d] - e.cen[d];
                        q += t * t;
                    }
                    r = std::max(r, sqrt(q));
                }
                e.radius = r;
                P.ev_rmax = std::max(P.ev_rmax, r);
                P.ev_rmin = std::min(P.ev_rmin, r);
                P.ev.push_back(e);
            }
    return true;
}

}  // namespace

void gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            pp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = 0.0;
        for (int k = 1; k <= n; ++k) {
            const double p2 = p1;
            p1 = p0;
            p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
        }
        pp = n * (z * p0 - p1) / (z * z - 1.0);
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    if (n % 2) x[n / 2] = 0.0;
}

// centres and areas as NC_ReadMesh computes them, minimum edge (delta_ of
// NC_WriteParameters) and the bounding-box diagonal of the BE nodes
void prepare_mesh(Mesh &M)
{
    M.cen.assign(3 * (size_t)M.n, 0.0);
    M.area.assign(M.n, 0.0);
    double area_sum = 0.0, dmin = 1.0e20;
    double mx[3] = {-9.9e20, -9.9e20, -9.9e20}, mn[3] = {9.9e20, 9.9e20, 9.9e20};
    for (int e = 0; e < M.n; ++e) {
        const int nv = M.nv[e];
        const double *X = &M.X[12 * (size_t)e];
        double N[4], ds[4], dt[4];
        const double s = nv == 3 ? 1.0 / 3.0 : 0.0;
        if (nv == 3) {
            N[0] = s; N[1] = s; N[2] = 1.0 - s - s;
            ds[0] = dt[1] = 1.0; ds[1] = dt[0] = 0.0; ds[2] = dt[2] = -1.0;
        } else {
            N[0] = N[1] = N[2] = N[3] = 0.25;
            ds[0] = 0.25; ds[1] = -0.25; ds[2] = -0.25; ds[3] = 0.25;
            dt[0] = 0.25; dt[1] = 0.25; dt[2] = -0.25; dt[3] = -0.25;
        }
        double xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
        for (int d = 0; d < 3; ++d) {
            double c = 0.0;
            for (int a = 0; a < nv; ++a) {
                c += N[a] * X[3 * a + d];
                xs[d] += ds[a] * X[3 * a + d];
                xt[d] += dt[a] * X[3 * a + d];
            }
            M.cen[3 * (size_t)e + d] = c;
        }
        const double n0 = xs[1] * xt[2] - xs[2] * xt[1], n1 = xs[2] * xt[0] - xs[0] * xt[2], n2 = xs[0] * xt[1] - xs[1] * xt[0];
        M.area[e] = sqrt(n0 * n0 + n1 * n1 + n2 * n2) * (nv == 3 ? 0.5 : 4.0);
        area_sum += M.area[e];
        for (int a = 0; a < nv; ++a) {
            const double *p = X + 3 * a, *q = X + 3 * ((a + 1) % nv);
            dmin = std::min(dmin, dist(p, q));
            for (int d = 0; d < 3; ++d) {
                mx[d] = std::max(mx[d], p[d]);
                mn[d] = std::min(mn[d], p[d]);
            }
        }
    }
    M.avg_area = area_sum / (double)M.n;
    M.delta = dmin / 100.0;
    M.diameter = sqrt((mx[0] - mn[0]) * (mx[0] - mn[0]) + (mx[1] - mn[1]) * (mx[1] - mn[1]) + (mx[2] - mn[2]) * (mx[2] - mn[2]));
    if (M.group_order.empty()) {
        for (int e = 0; e < M.n; ++e)
            if (std::find(M.group_order.begin(), M.group_order.end(), M.group[e]) == M.group_order.end())
                M.group_order.push_back(M.group[e]);
    }
}

bool make_plan(const Mesh &M, const Options &O, double k, Plan &P)
{
    P = Plan();
    const double lambda = 2.0 * kPi / k;
    const double ratio = lambda / M.diameter;
    static const double hmin[11] = {0, 0, 0, 0, 0, 0.0051, 0.0177, 0.0445, 0.0905, 0.1597, 0.2547};
    P.min_dist = (O.min_terms >= 5 && O.min_terms <= 10 ? hmin[O.min_terms] : 0.0) / k;
    // method choice (NC_DeclareArrays, src/NC_Addresses.cpp:78-122)
    int method = 0, nlev = 1;
    double edge_ml = 0.0;
    if (O.method_bem == 0 || ratio > 80.0) method = 0;
    else if (O.method_bem == 1) method = 1;
    else if (ratio > 1.0) method = 1;
    else {
        // NC_EstimateNumLevelsMLFMM (src/NC_Addresses.cpp:804-845)
        const double ele = sqrt(M.avg_area), dmin = 25.0;
        double root = sqrt((double)M.n) * 0.9;
        if (root < dmin) root = dmin;
        const double lmin = sqrt(dmin) * ele;
        if (O.clus_edge <= 0.0) edge_ml = sqrt((double)M.n / root) * ele;
        else {
            edge_ml = O.clus_edge;
            if (edge_ml < lmin) { P.error = "Inputed value of cluster diameter for level 0 is too small!"; return false; }
        }
        const double dl = log10(edge_ml / lmin) / log10(2.0);
        nlev = (int)dl;
        if (dl - (double)nlev > 0.5) nlev++;
        nlev++;
        if (nlev <= 0) nlev = 1;
        method = nlev == 1 ? 1 : 3;
        if (nlev > 10) { P.error = "numClusterLevels_ must <= 10!"; return false; }
    }
    for (;;) {
        P.method = method;
        P.lev.clear();
        if (method == 0) return true;
        std::string err;
        if (method == 1) {
            // NC_GenerateClustersSLFMM (src/NC_Addresses.cpp:290-350)
            const double nels = sqrt((double)M.n);
            if (O.clus_edge <= 0.0) P.clus_edge = sqrt(nels * M.avg_area);
            else {
                P.clus_edge = O.clus_edge;
                if (P.clus_edge < sqrt(4.0 * M.avg_area)) { P.error = "The read cluster diameter is too small!"; return false; }
            }
        } else {
            P.clus_edge = edge_ml;
        }
        // the cluster tree depends on the mesh, the method and the number of levels only
        const int tlev = method == 1 ? 1 : nlev;
        const Tree *tree = nullptr;
        for (const Tree &t : M.trees)
            if (t.method == method && t.nlev == tlev) tree = &t;
        if (!tree) {
            Tree t;
            t.method = method;
            t.nlev = tlev;
            for (int lv = 0; lv < tlev; ++lv) {
                Level L;
                const double edge = P.clus_edge / pow(2.0, lv);
                if (lv == 0) {
                    for (int g : M.group_order) {
                        std::vector<int> elems;
                        for (int e = 0; e < M.n; ++e)
                            if (M.group[e] == g) elems.push_back(e);
                        if (elems.empty()) continue;
                        auto parts = split_group(M, elems, P.clus_edge, false, err);
                        cancel_small(M, parts);
                        for (auto &c : parts)
                            if (!c.empty()) { Cluster cl; cl.elems = c; L.cl.push_back(cl); }
                    }
                } else {
                    const Level &F = t.lev[lv - 1];
                    for (int f = 0; f < (int)F.cl.size(); ++f) {
                        auto parts = split_group(M, F.cl[f].elems, edge, true, err);
                        cancel_small(M, parts);
                        for (auto &c : parts)
                            if (!c.empty()) { Cluster cl; cl.elems = c; cl.father = f; L.cl.push_back(cl); }
                    }
                }
                if (!err.empty()) { P.error = err; return false; }
                finish_level(M, L);
                t.lev.push_back(L);
            }
            M.trees.push_back(std::move(t));
            tree = &M.trees.back();
        }
        // lists without cancelled pairs, once per tree
        Tree *tw = const_cast<Tree *>(tree);
        if (!tw->geo_ready) {
            tw->geo = tw->lev;
            tw->dfar.assign(tw->geo.size(), 0.0);
            for (size_t lv = 0; lv < tw->geo.size(); ++lv)
                lists(tw->geo[lv], 0.0, method == 3 && lv > 0 ? &tw->geo[lv - 1] : nullptr, &tw->dfar[lv]);
            tw->geo_ready = true;
        }
        bool geo = true;
        for (size_t lv = 0; lv < tw->geo.size(); ++lv)
            if (!(P.min_dist < tw->dfar[lv])) geo = false;
        if (geo) {
            P.lev = tw->geo;
        } else {
            P.lev = tree->lev;
            for (size_t lv = 0; lv < P.lev.size(); ++lv)
                lists(P.lev[lv], P.min_dist, method == 3 && lv > 0 ? &P.lev[lv - 1] : nullptr);
        }
        // canceled-pair ratios (src/NC_Addresses.cpp:224-283)
        P.cancel_ratio.clear();
        bool fallback = false;
        for (const Level &L : P.lev) {
            const double r = L.n_canceled > 0 ? (double)L.n_canceled / (double)L.n_far_geom : 0.0;
            P.cancel_ratio.push_back(r);
            if (r > 0.75) fallback = true;
        }
        if (!fallback) break;
        if (method == 1) {
            if (M.n >= 20000)
                P.warning = "Frquency is too low for the single level FMBEM; the dense method is used "
                            "(the reference stops above 20000 elements)";
            method = 0;
            P.lev.clear();
            P.method = 0;
            return true;
        }
        if (M.n >= 50000) { P.error = "Frquency is too low, the multilevel FMBEM can not be used!"; return false; }
        method = 1;
        nlev = 1;
    }
    for (Level &L : P.lev) {
        std::string err;
        if (!expansion(L, k, O.min_terms, err)) { P.error = err; return false; }
    }
    // the evaluation clusters depend on the mesh and the cluster edge only
    Tree *tree = nullptr;
    for (Tree &t : M.trees)
        if (t.method == P.method && t.nlev == (int)P.lev.size()) tree = &t;
    if (tree && tree->ev_ready && tree->ev_edge == P.clus_edge) {
        P.ev = tree->ev;
        P.ev_rmax = tree->ev_rmax;
        P.ev_rmin = tree->ev_rmin;
        return true;
    }
    std::string err;
    if (!eval_clusters(M, P, err)) { P.error = err; return false; }
    if (tree) {
        tree->ev = P.ev;
        tree->ev_rmax = P.ev_rmax;
        tree->ev_rmin = P.ev_rmin;
        tree->ev_edge = P.clus_edge;
        tree->ev_ready = true;
    }
    return true;
}

}  // namespace ncgf
