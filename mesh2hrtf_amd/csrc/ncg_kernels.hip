// libncg kernels for gfx950 (wave64).  One translation unit holds every kernel so the
// quadrature tables in constant memory exist once per code object.
//
// Layouts in HBM (all row-major, complex interleaved):
//   A      [n][n]  cx<S>, S = double (NCG_FP64) or float (NCG_MIXED)
//   vectors [nrhs][n] cx<double>
//   element data: structure of arrays (see ncg::Elements)
// the descriptors of the grouped launches hold global-memory pointers in this unit's device pass
#define NCG_KERNELS_TU 1
#include "ncg_kernels.hpp"

#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "ncg_device.hpp"

// Every launch goes through ncg_launch: queued on its stream, or, while the calling thread
// has a graph recorder set (ncgk::set_graph_recorder), added to the recorder's graph as a
// kernel node after the previous one.  Recording needs no stream capture, so other threads
// may allocate, free and synchronise meanwhile.
namespace {
thread_local ncgk::GraphRecorder *t_recorder = nullptr;

template <typename Tup, size_t... I>
void arg_pointers(Tup &t, void **a, std::index_sequence<I...>)
{
    ((a[I] = (void *)&std::get<I>(t)), ...);
}

template <typename... P, typename... A>
void ncg_launch(void (*k)(P...), dim3 grid, dim3 block, size_t lds, hipStream_t s, A &&...a)
{
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    std::tuple<std::decay_t<P>...> v(static_cast<std::decay_t<P>>(a)...);
    void *args[sizeof...(P) + 1];
    arg_pointers(v, args, std::index_sequence_for<P...>{});
    if (ncgk::GraphRecorder *r = t_recorder) {
        if (r->err != hipSuccess) return;
        hipKernelNodeParams prm{};
        prm.func = (void *)k;
        prm.gridDim = grid;
        prm.blockDim = block;
        prm.sharedMemBytes = (unsigned)lds;
        prm.kernelParams = args;
        prm.extra = nullptr;
        hipGraphNode_t node = nullptr;
        r->err = hipGraphAddKernelNode(&node, r->graph, r->last ? &r->last : nullptr, r->last ? 1 : 0, &prm);
        if (r->err == hipSuccess) {
            r->last = node;
            r->nodes++;
        }
        return;
    }
    (void)hipLaunchKernel((const void *)k, grid, block, args, lds, s);
}
}  // namespace

namespace ncgk {
void set_graph_recorder(GraphRecorder *r) { t_recorder = r; }
}  // namespace ncgk

namespace ncg {

// a pointer into global memory (NCG_GLOBAL, ncg_kernels.hpp) from a generic one
template <typename T>
__device__ __forceinline__ NCG_GLOBAL T *to_global(T *p)
{
    return (NCG_GLOBAL T *)p;
}

// PL adjacent complex values q[0 .. PL) of one array; two FP32 complex values move as one
// 16-byte load or store (the caller's index is even and the array 16-byte aligned)
typedef float ncg_f4 __attribute__((ext_vector_type(4)));
template <typename T, int PL>
__device__ __forceinline__ void load_pts(const NCG_GLOBAL cx<T> *q, cx<T> (&o)[PL])
{
    if constexpr (PL == 2 && sizeof(T) == 4) {
        const ncg_f4 v = *(const NCG_GLOBAL ncg_f4 *)q;
        o[0] = cx<T>{v.x, v.y};
        o[1] = cx<T>{v.z, v.w};
    } else {
#pragma unroll
        for (int h = 0; h < PL; ++h) o[h] = q[h];
    }
}
template <typename T, int PL>
__device__ __forceinline__ void store_pts(NCG_GLOBAL cx<T> *q, const cx<T> (&o)[PL])
{
    if constexpr (PL == 2 && sizeof(T) == 4) {
        *(NCG_GLOBAL ncg_f4 *)q = ncg_f4{o[0].re, o[0].im, o[1].re, o[1].im};
    } else {
#pragma unroll
        for (int h = 0; h < PL; ++h) q[h] = o[h];
    }
}

// ======================================================================================
// geometry: centre, unit normal, area, Jacobian per element (NC_ReadMesh,
// src/NC_Input.cpp:522-605; triangle centre s = t = 1/3 and area 0.5*J, quad centre
// (0, 0) and area 4*J)
// ======================================================================================
__global__ void k_element_geometry(int n, const double *X, const int8_t *nv, double *cen, double *nrm,
                                   double *area, double *jac)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    Elements E{};
    E.n = n;
    E.X = X;
    LocalElem L;
    L.nv = nv[e];
    for (int a = 0; a < L.nv; ++a)
        for (int d = 0; d < 3; ++d) L.p[a][d] = X[(3 * a + d) * n + e];
    const double s = L.nv == 3 ? 1.0 / 3.0 : 0.0;
    double q[3], ny[3], J;
    if (L.nv == 3) {
        // NC_ComputeUnitVectorOnElement for a flat triangle: dx/ds = x0 - x2, dx/dt = x1 - x2
        double xs[3], xt[3];
        for (int d = 0; d < 3; ++d) { xs[d] = L.p[0][d] - L.p[2][d]; xt[d] = L.p[1][d] - L.p[2][d]; }
        double n0 = xs[1] * xt[2] - xs[2] * xt[1], n1 = xs[2] * xt[0] - xs[0] * xt[2], n2 = xs[0] * xt[1] - xs[1] * xt[0];
        J = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        ny[0] = n0 / J; ny[1] = n1 / J; ny[2] = n2 / J;
        for (int d = 0; d < 3; ++d) q[d] = s * L.p[0][d] + s * L.p[1][d] + (1.0 - s - s) * L.p[2][d];
    } else {
        L.jac = 0.0;
        point_on(L, s, s, q, ny, J);
    }
    for (int d = 0; d < 3; ++d) { cen[d * n + e] = q[d]; nrm[d * n + e] = ny[d]; }
    area[e] = J * (L.nv == 3 ? 0.5 : 4.0);
    jac[e] = J;
}

// ======================================================================================
// near-pair table: pairs (i, j != i) whose element j must be subdivided as seen from
// the collocation point of i (ratio < 1.3 at level 0).  Frequency independent.
// ======================================================================================
// One wave per collocation row i, lanes over the columns j; fill keeps the columns of a
// row in ascending order (ballot prefix per step of 64 columns).
__global__ void __launch_bounds__(256) k_near_count(Elements E, int *count)
{
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= E.n) return;
    const double x[3] = {E.cen[i], E.cen[E.n + i], E.cen[2 * E.n + i]};
    int c = 0;
    for (int j0 = 0; j0 < E.n; j0 += 64) {
        const int j = j0 + lane;
        c += __popcll(__ballot(j < E.n && j != i && pair_is_near(E, j, x)));
    }
    if (lane == 0) count[i] = c;
}

__global__ void __launch_bounds__(256) k_near_fill(Elements E, const long long *offset, int *rows, int *cols)
{
    const int lane = threadIdx.x & 63, i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= E.n) return;
    const double x[3] = {E.cen[i], E.cen[E.n + i], E.cen[2 * E.n + i]};
    long long o = offset[i];
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int j0 = 0; j0 < E.n; j0 += 64) {
        const int j = j0 + lane;
        const bool near = j < E.n && j != i && pair_is_near(E, j, x);
        const unsigned long long m = __ballot(near);
        if (near) {
            const long long a = o + __popcll(m & below);
            rows[a] = i;
            cols[a] = j;
        }
        o += __popcll(m);
    }
}

// ======================================================================================
// far pairs: the bulk of the dense operator.  Block = 256 threads = 4 waves, one column
// element per thread (held in registers), rows_per_block collocation rows walked by the
// whole block (row data is wave-uniform -> scalar loads).  Pairs that are near or on
// the diagonal are left to k_assemble_near / k_assemble_diag.  Writes are coalesced:
// consecutive lanes write consecutive columns of one row.
// ======================================================================================
// Quadrature evaluation counts (the measurement pass of a plan only, evals != nullptr): the
// wave's count added to *evals with one atomic.  Every lane of the wave calls it.
__device__ __forceinline__ void wave_count(unsigned long long *evals, unsigned long long nev)
{
    for (int o = 32; o > 0; o >>= 1) nev += __shfl_xor(nev, o);
    if ((threadIdx.x & 63) == 0 && nev) atomicAdd(evals, nev);
}

template <typename T, typename S>
__global__ void __launch_bounds__(256) k_assemble_far(Elements E, BMConst bm, const double *adm,
                                                       cx<S> *A, int rows_per_block, unsigned long long *evals)
{
    const int n = E.n;
    const int j = blockIdx.x * 256 + threadIdx.x;
    const int i0 = blockIdx.y * rows_per_block;
    const int i1 = min(i0 + rows_per_block, n);
    unsigned long long nev = 0;
    if (j < n) {
    const int nv = E.nv[j];
    const int kind = E.kind[j];
    const cx<T> admj = {T(adm[2 * j]), T(adm[2 * j + 1])};
    const double cj[3] = {E.cen[j], E.cen[n + j], E.cen[2 * n + j]};
    const double areaj = E.area[j];
    const T nyx = T(E.nrm[j]), nyy = T(E.nrm[n + j]), nyz = T(E.nrm[2 * n + j]);
    const T jacj = T(E.jac[j]);
    double P[4][3];
    for (int a = 0; a < nv; ++a)
        for (int d = 0; d < 3; ++d) P[a][d] = E.X[(3 * a + d) * n + j];
    // triangle: x(s,t) = x2 + s (x0 - x2) + t (x1 - x2)
    const T e1x = T(P[0][0] - P[2][0]), e1y = T(P[0][1] - P[2][1]), e1z = T(P[0][2] - P[2][2]);
    const T e2x = T(P[1][0] - P[2][0]), e2y = T(P[1][1] - P[2][1]), e2z = T(P[1][2] - P[2][2]);
    const T hk = T(bm.hk), k2 = T(bm.k2), tau = T(bm.tau), bim = T(bm.beta_im);
    for (int i = i0; i < i1; ++i) {
        if (i == j) continue;
        const double x[3] = {E.cen[i], E.cen[n + i], E.cen[2 * n + i]};
        const double ratdis = ratio_to_centre(cj[0], cj[1], cj[2], x, areaj);
        if (ratdis < kTolf) continue;
        const T nxx = T(E.nrm[i]), nxy = T(E.nrm[n + i]), nxz = T(E.nrm[2 * n + i]);
        const int m = gauss_order(ratdis);
        BMAcc<T> acc;
        acc.G = acc.H = acc.HT = acc.E = mk(T(0), T(0));
        if (nv == 3) {
            const T px = T(P[2][0] - x[0]), py = T(P[2][1] - x[1]), pz = T(P[2][2] - x[2]);
            const int off = tri_rule_offset(m), cnt = tri_rule_count(m);
            for (int g = 0; g < cnt; ++g) {
                const T s = T(c_tri[off + g][0]), t = T(c_tri[off + g][1]), w = T(c_tri[off + g][2]);
                const T qx = px + s * e1x + t * e2x, qy = py + s * e1y + t * e2y, qz = pz + s * e1z + t * e2z;
                bm_point<T>(acc, qx, qy, qz, nyx, nyy, nyz, nxx, nxy, nxz, w * jacj, hk, k2);
            }
            nev += cnt;
        } else {
            LocalElem L;
            L.nv = 4;
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) L.p[a][d] = P[a][d] - x[d];
            const int base = m * (m - 1) / 2;
            for (int u = 0; u < m; ++u)
                for (int v = 0; v < m; ++v) {
                    double q[3], ny[3], J;
                    point_on(L, c_gl_x[base + u], c_gl_x[base + v], q, ny, J);
                    const double w = c_gl_w[base + u] * c_gl_w[base + v] * J;
                    bm_point<T>(acc, T(q[0]), T(q[1]), T(q[2]), T(ny[0]), T(ny[1]), T(ny[2]), nxx, nxy, nxz, T(w), hk, k2);
                }
            nev += m * m;
        }
        const cx<T> a = bm_entry<T>(acc, kind, admj, tau, bim);
        A[(size_t)i * n + j] = cx<S>{S(a.re), S(a.im)};
    }
    }
    if (evals) wave_count(evals, nev);
}

// ======================================================================================
// near pairs: adaptive subdivision in FP64, one thread per pair
// ======================================================================================
template <typename S>
__global__ void __launch_bounds__(128) k_assemble_near(Elements E, BMConst bm, const double *adm, const int *rows,
                                                        const int *cols, long long npairs, cx<S> *A, int *err,
                                                        unsigned long long *evals)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nev = 0;
    if (p < npairs) {
        const int i = rows[p], j = cols[p];
        BMAcc<double> acc;
        const int e = bm_pair(E, i, j, bm.hk, bm.k2, acc, &nev);
        if (e) atomicOr(err, e);
        const cx<double> a = bm_entry<double>(acc, E.kind[j], mk(adm[2 * j], adm[2 * j + 1]), bm.tau, bm.beta_im);
        A[(size_t)i * E.n + j] = cx<S>{S(a.re), S(a.im)};
    }
    if (evals) wave_count(evals, nev);
}

// ======================================================================================
// right-hand-side contributions of prescribed boundary values: b_r[i] += sum over the
// source columns j (elements with a nonzero prescribed value in some load case) of
// c_ij * bc_r[j].  One wave per row, lanes over the source columns; FP64.
// ======================================================================================
constexpr int kRhsChunk = 8;

__global__ void __launch_bounds__(256) k_rhs_sources(Elements E, BMConst bm, const int *src, int nsrc,
                                                      const cx<double> *bcval, int nrhs, int rhs0,
                                                      cx<double> *rhs_src, int *err, const int *elem_cluster,
                                                      const unsigned char *cluster_near, int nclusters)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= E.n) return;
    const int nr = min(kRhsChunk, nrhs - rhs0);
    cx<double> part[kRhsChunk];
    for (int r = 0; r < kRhsChunk; ++r) part[r] = mk(0.0, 0.0);
    for (int q = lane; q < nsrc; q += 64) {
        const int j = src[q];
        // FMM: only the near field is integrated here; far sources enter through the T-vector
        if (cluster_near && !cluster_near[(size_t)elem_cluster[i] * nclusters + elem_cluster[j]]) continue;
        BMAcc<double> acc;
        const int e = bm_pair(E, i, j, bm.hk, bm.k2, acc);
        if (e) atomicOr(err, e);
        const cx<double> c = bm_rhs_coef<double>(acc, E.kind[j], bm.tau, bm.beta_im);
        for (int r = 0; r < nr; ++r) part[r] += c * bcval[(size_t)(rhs0 + r) * E.n + j];
    }
    for (int r = 0; r < nr; ++r) {
        double re = part[r].re, im = part[r].im;
        for (int o = 32; o > 0; o >>= 1) {
            re += __shfl_xor(re, o);
            im += __shfl_xor(im, o);
        }
        if (lane == 0) rhs_src[(size_t)(rhs0 + r) * E.n + i] = mk(re, im);
    }
}

// ======================================================================================
// diagonal: singular integrals + free terms, and the right-hand sides
// (NC_ComputeEntriesForTBEM free terms src/NC_EquationSystem.cpp:218-237 and
// NC_IncidentWaveRHS src/NC_EquationSystem.cpp:1731-1812)
// ======================================================================================
__global__ void __launch_bounds__(64) k_assemble_diag(Elements E, BMConst bm, const double *adm, const cx<double> *bcval,
                                                      int nrhs, const double *ps, int nps, const double *pw, int npw,
                                                      const cx<double> *rhs_src, int has_src, void *Aout,
                                                      int mixed, cx<double> *b, const long long *diag_index,
                                                      unsigned long long *evals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (evals) {   // singular rule: the same count for every element of a kind
        unsigned long long nev = 0;
        if (i < E.n) nev = E.nv[i] * 44ull;   // per edge 4 x 3 line points, 2 x 16 surface points
        wave_count(evals, nev);
    }
    if (i >= E.n) return;
    const int n = E.n;
    BMAcc<double> acc;
    bm_pair(E, i, i, bm.hk, bm.k2, acc);
    const int kind = E.kind[i];
    const cx<double> admi = mk(adm[2 * i], adm[2 * i + 1]);
    const cx<double> beta = mk(0.0, bm.beta_im);
    cx<double> a = bm_entry<double>(acc, kind, admi, bm.tau, bm.beta_im);
    if (kind == 1) a -= beta * (0.5 * bm.tau);
    else a += (admi * beta - mk(1.0, 0.0)) * 0.5;
    // dense operator: entry (i, i); near-field CSR: its diagonal position
    const size_t at = diag_index ? (size_t)diag_index[i] : (size_t)i * n + i;
    if (mixed) ((cx<float> *)Aout)[at] = cx<float>{float(a.re), float(a.im)};
    else ((cx<double> *)Aout)[at] = a;

    const double x[3] = {E.cen[i], E.cen[n + i], E.cen[2 * n + i]};
    const double nx[3] = {E.nrm[i], E.nrm[n + i], E.nrm[2 * n + i]};
    const double hk = bm.hk;
    const cx<double> btau = beta * bm.tau;
    for (int r = 0; r < nrhs; ++r) {
        const cx<double> v = bcval[(size_t)r * n + i];
        cx<double> z = kind == 1 ? v * (bm.tau * 0.5) : v * beta * (bm.tau * 0.5);
        for (int w = 0; w < npw; ++w) {
            const double *d = pw + 6 * w;
            if ((int)d[5] != r) continue;
            const double ph = (d[0] * x[0] + d[1] * x[1] + d[2] * x[2]) * hk;
            double sn, cs;
            sincos(ph, &sn, &cs);
            const cx<double> g = mk(cs, sn) * mk(d[3], d[4]);
            const double c = hk * (d[0] * nx[0] + d[1] * nx[1] + d[2] * nx[2]);
            const cx<double> vv = mk(-g.im * c, g.re * c);
            z -= g + vv * btau;
        }
        for (int s = 0; s < nps; ++s) {
            const double *y = ps + 6 * s;
            if ((int)y[5] != r) continue;
            double v2[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
            const double R = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
            v2[0] /= R; v2[1] /= R; v2[2] /= R;
            double sn, cs;
            sincos(hk * R, &sn, &cs);
            const double f = 1.0 / kPi4 / R;
            const cx<double> g = mk(cs * f, sn * f) * mk(y[3], y[4]);
            const double rn = v2[0] * nx[0] + v2[1] * nx[1] + v2[2] * nx[2];
            const cx<double> vv = mk(-rn / R, hk * rn) * g;
            z -= g + vv * btau;
        }
        if (has_src) z += rhs_src[(size_t)r * n + i];
        b[(size_t)r * n + i] = z;
    }
}

// ======================================================================================
// Krylov primitives
// ======================================================================================
// y[r] = A (d .* x[r]) for r < nr (<= 4), wave = 4 rows, block = 4 waves = 16 rows.
// Lanes stride the columns: each load instruction reads 64 consecutive entries of a row
// (1 KiB FP64, 512 B FP32), so HBM traffic is the operator once per call plus the
// vectors once per 16 rows.
template <typename S, typename Acc, int NR>
__global__ void __launch_bounds__(256) k_matvec(const cx<S> *__restrict__ A, int n, const cx<double> *__restrict__ x,
                                                 const double *__restrict__ d, cx<double> *__restrict__ y)
{
    constexpr int ROWS = 4;
    const int lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * ROWS;
    if (row0 >= n) return;
    cx<Acc> acc[ROWS][NR];
#pragma unroll
    for (int a = 0; a < ROWS; ++a)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[a][r] = mk(Acc(0), Acc(0));
    const cx<S> *rowp[ROWS];
#pragma unroll
    for (int a = 0; a < ROWS; ++a) rowp[a] = A + (size_t)min(row0 + a, n - 1) * n;
    for (int j = lane; j < n; j += 64) {
        cx<Acc> xv[NR];
        const double dj = d ? d[j] : 1.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const cx<double> v = x[(size_t)r * n + j];
            xv[r] = mk(Acc(v.re * dj), Acc(v.im * dj));
        }
#pragma unroll
        for (int a = 0; a < ROWS; ++a) {
            const cx<S> e = rowp[a][j];
            const cx<Acc> ea = mk(Acc(e.re), Acc(e.im));
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[a][r] += ea * xv[r];
        }
    }
#pragma unroll
    for (int a = 0; a < ROWS; ++a)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double re = double(acc[a][r].re), im = double(acc[a][r].im);
            for (int o = 32; o > 0; o >>= 1) {
                re += __shfl_xor(re, o);
                im += __shfl_xor(im, o);
            }
            if (lane == 0 && row0 + a < n) y[(size_t)r * n + row0 + a] = mk(re, im);
        }
}

// Block partial sums of up to 3 bilinear forms per load case:
//   out[b][r][q] = sum_i u_q[r][i] * v_q[r][i]   (bit q of conj_mask: conjugate u_q)
// Deterministic: fixed partition, reduced by k_reduce_partials.
constexpr int kDotBlocks = 256;

__global__ void __launch_bounds__(256) k_dots(int n, int nrhs, int ndots, const cx<double> *u0, const cx<double> *v0,
                                              const cx<double> *u1, const cx<double> *v1, const cx<double> *u2,
                                              const cx<double> *v2, int conj_mask, cx<double> *partial)
{
    __shared__ double sre[256], sim[256];
    const NCG_GLOBAL cx<double> *U[3] = {to_global(u0), to_global(u1), to_global(u2)};
    const NCG_GLOBAL cx<double> *V[3] = {to_global(v0), to_global(v1), to_global(v2)};
    for (int r = 0; r < nrhs; ++r)
        for (int q = 0; q < ndots; ++q) {
            double re = 0.0, im = 0.0;
            const NCG_GLOBAL cx<double> *u = U[q] + (size_t)r * n, *v = V[q] + (size_t)r * n;
            for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
                cx<double> a = u[i];
                if ((conj_mask >> q) & 1) a.im = -a.im;
                const cx<double> p = a * v[i];
                re += p.re;
                im += p.im;
            }
            sre[threadIdx.x] = re;
            sim[threadIdx.x] = im;
            __syncthreads();
            for (int s = 128; s > 0; s >>= 1) {
                if (threadIdx.x < s) {
                    sre[threadIdx.x] += sre[threadIdx.x + s];
                    sim[threadIdx.x] += sim[threadIdx.x + s];
                }
                __syncthreads();
            }
            if (threadIdx.x == 0) partial[((size_t)blockIdx.x * nrhs + r) * 3 + q] = mk(sre[0], sim[0]);
            __syncthreads();
        }
}

// out[r][q] = sum over blocks of partial; one block of 256 threads, nblocks <= 256.  Each
// sum is a fixed tree over the block partials (deterministic).
__global__ void __launch_bounds__(256) k_reduce_partials(int nblocks, int nrhs, int ndots, const cx<double> *partial,
                                                         cx<double> *out)
{
    __shared__ double sre[256], sim[256];
    const int t = threadIdx.x;
    for (int rq = 0; rq < nrhs * ndots; ++rq) {
        const int r = rq / ndots, q = rq % ndots;
        const cx<double> v = t < nblocks ? partial[((size_t)t * nrhs + r) * 3 + q] : mk(0.0, 0.0);
        sre[t] = v.re;
        sim[t] = v.im;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (t < s) {
                sre[t] += sre[t + s];
                sim[t] += sim[t + s];
            }
            __syncthreads();
        }
        if (t == 0) out[r * 3 + q] = mk(sre[0], sim[0]);
        __syncthreads();
    }
}

// CGS scalars, one thread per load case.  sc layout per r (complex):
// [0] rho0 [1] rho1 [2] sigma [3] alpha [4] beta [5] (rr, -) [6] err0 [7] flags
__global__ void k_cgs_alpha(int nrhs, int first, const cx<double> *dots, cx<double> *sc, const int *active)
{
    const int r = threadIdx.x;
    if (r >= nrhs || !active[r]) return;
    cx<double> *s = sc + 8 * r;
    // dots: [r][0] = <R,R0> (only valid in the first iteration), [r][1] = <V,R0>
    if (first) s[0] = dots[r * 3 + 0];
    else s[0] = s[1];
    s[2] = dots[r * 3 + 1];
    s[3] = cdiv(s[0], s[2]);
}

__global__ void k_cgs_beta(int nrhs, const cx<double> *dots, cx<double> *sc, const int *active)
{
    const int r = threadIdx.x;
    if (r >= nrhs || !active[r]) return;
    cx<double> *s = sc + 8 * r;
    // dots: [r][0] = <R,R> (real), [r][1] = <R,R0>
    s[5] = dots[r * 3 + 0];
    s[1] = dots[r * 3 + 1];
    s[4] = cdiv(s[1], s[0]);
}

// q = u - alpha v ; uq = u + q
__global__ void k_cgs_update1(int n, int nrhs, const cx<double> *sc, const int *active, const cx<double> *u,
                              const cx<double> *v, cx<double> *q, cx<double> *uq)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    if (!active[r]) return;
    const cx<double> alpha = sc[8 * r + 3];
    const cx<double> qq = u[t] - alpha * v[t];
    q[t] = qq;
    uq[t] = u[t] + qq;
}

// x += alpha uq ; r -= alpha auq
__global__ void k_cgs_update2(int n, int nrhs, const cx<double> *sc, const int *active, const cx<double> *uq,
                              const cx<double> *auq, cx<double> *x, cx<double> *rr)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    if (!active[r]) return;
    const cx<double> alpha = sc[8 * r + 3];
    x[t] += alpha * uq[t];
    rr[t] -= alpha * auq[t];
}

// u = r + beta q ; p = u + beta (q + beta p)
__global__ void k_cgs_update3(int n, int nrhs, const cx<double> *sc, const int *active, const cx<double> *rr,
                              const cx<double> *q, cx<double> *u, cx<double> *p)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    if (!active[r]) return;
    const cx<double> beta = sc[8 * r + 4];
    const cx<double> uu = rr[t] + beta * q[t];
    u[t] = uu;
    p[t] = uu + beta * (q[t] + beta * p[t]);
}

// y = alpha x + beta y elementwise per load case with per-case coefficients
// (coef[r] = (alpha, beta) complex pairs); used by GMRES / refinement.
__global__ void k_axpby(int n, int nrhs, const cx<double> *coef, const cx<double> *x, size_t ldx, cx<double> *y,
                        size_t ldy)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    const size_t i = t % n;
    const cx<double> a = coef[2 * r], b = coef[2 * r + 1];
    y[r * ldy + i] = a * x[r * ldx + i] + b * y[r * ldy + i];
}

__global__ void k_scale_vec(int n, int nrhs, const double *d, cx<double> *x)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    x[t] = x[t] * d[t % n];
}

// row scaling of NC_ComputeScalingVector (src/NC_CommonFunctions.cpp:1322-1354):
// d_i = sqrt(n / sum_j |a_ij|^2), one wave per row
template <typename S>
__global__ void k_row_scaling(const cx<S> *A, int n, double *d)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) {
        const cx<S> a = A[(size_t)i * n + j];
        s += double(a.re) * double(a.re) + double(a.im) * double(a.im);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) d[i] = sqrt((double)n / s);
}

// the same scaling for the FMM operator, over each row's near-field entries
// (src/NC_CommonFunctions.cpp:1344-1352)
template <typename S>
__global__ void k_row_scaling_csr(int n, const long long *ptr, const cx<S> *val, double *d)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    double s = 0.0;
    for (long long k = ptr[i] + lane; k < ptr[i + 1]; k += 64) {
        const cx<S> a = val[k];
        s += double(a.re) * double(a.re) + double(a.im) * double(a.im);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) d[i] = sqrt((double)(ptr[i + 1] - ptr[i]) / s);
}

// GMRES: h[r][k] = <V_k, w_r> for k < kcount (Hermitian).  One block per (k, r, slice of
// n) writes a partial sum; k_multidot_sum adds the slices in a fixed order.
// V layout: [basis k][rhs r][n]; w: [r][n]; part: [(k * nrhs + r) * nslice + slice].
__device__ __forceinline__ void multidot_item(int k, int r, int sl, int nsl, int n, int nrhs, int slice,
                                              const cx<double> *V, const cx<double> *w, cx<double> *part)
{
    __shared__ double sre[256], sim[256];
    const cx<double> *v = V + ((size_t)k * nrhs + r) * n, *ww = w + (size_t)r * n;
    const int i0 = sl * slice, i1 = min(n, i0 + slice);
    double re = 0.0, im = 0.0;
    // four elements per thread and step, every load first (past the slice: weight zero)
    for (int i = i0 + threadIdx.x; i < i1; i += 1024) {
        cx<double> a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ii = min(i + 256 * q, i1 - 1);
            a[q] = v[ii];
            b[q] = ww[ii];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (i + 256 * q < i1) {
                re += a[q].re * b[q].re + a[q].im * b[q].im;
                im += a[q].re * b[q].im - a[q].im * b[q].re;
            }
    }
    sre[threadIdx.x] = re;
    sim[threadIdx.x] = im;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sre[threadIdx.x] += sre[threadIdx.x + s];
            sim[threadIdx.x] += sim[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[((size_t)k * nrhs + r) * nsl + sl] = mk(sre[0], sim[0]);
}

__global__ void __launch_bounds__(256) k_multidot(int n, int nrhs, int kcount, int slice, const cx<double> *V,
                                                  const cx<double> *w, cx<double> *part)
{
    multidot_item(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, n, nrhs, slice, V, w, part);
}

__global__ void k_multidot_sum(int nrhs, int kcount, int nsl, const cx<double> *part, cx<double> *h)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrhs * kcount) return;
    const int k = t / nrhs, r = t % nrhs;
    cx<double> s = mk(0.0, 0.0);
    for (int q = 0; q < nsl; ++q) s += part[(size_t)t * nsl + q];
    h[(size_t)r * kcount + k] = s;
}

// w[r] -= sum_k V_k[r] h[r][k]
__global__ void k_multiaxpy(int n, int nrhs, int kcount, const cx<double> *V, const cx<double> *h, cx<double> *w)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    const size_t i = t % n;
    cx<double> s = w[t];
    for (int k = 0; k < kcount; ++k) s -= V[((size_t)k * nrhs + r) * n + i] * h[(size_t)r * kcount + k];
    w[t] = s;
}

// One Gram-Schmidt pass of GMRES after k_multidot: h[r][k] = sum over the slices of
// part (fixed order), w[r] -= sum_k V_k[r] h[r][k].  Block (x, r); block x = 0 stores
// h[r][:] for the host.  With npart, block x also writes its partial sum of |w_new|^2
// to npart[r][x] (gridDim.x <= 256).
constexpr int kGsMaxBasis = 256;
__device__ __forceinline__ void gs_update_item(int bx, int r, int gx, int n, int nrhs, int kcount, int nsl,
                                               const cx<double> *part, const cx<double> *V, cx<double> *w,
                                               cx<double> *hout, double *npart)
{
    __shared__ cx<double> hs[kGsMaxBasis];
    __shared__ double red[256];
    const int t = threadIdx.x;
    for (int k = t; k < kcount; k += 256) {
        cx<double> s = mk(0.0, 0.0);
        for (int q = 0; q < nsl; ++q) s += part[((size_t)k * nrhs + r) * nsl + q];
        hs[k] = s;
        if (bx == 0) hout[(size_t)r * kcount + k] = s;
    }
    __syncthreads();
    double nn = 0.0;
    for (int i = bx * 256 + t; i < n; i += gx * 256) {
        cx<double> s = w[(size_t)r * n + i];
        for (int k = 0; k < kcount; ++k) s -= V[((size_t)k * nrhs + r) * n + i] * hs[k];
        w[(size_t)r * n + i] = s;
        nn += s.re * s.re + s.im * s.im;
    }
    if (!npart) return;
    red[t] = nn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) npart[(size_t)r * gx + bx] = red[0];
}

__global__ void __launch_bounds__(256) k_gs_update(int n, int nrhs, int kcount, int nsl, const cx<double> *part,
                                                   const cx<double> *V, cx<double> *w, cx<double> *hout, double *npart)
{
    gs_update_item(blockIdx.x, blockIdx.y, gridDim.x, n, nrhs, kcount, nsl, part, V, w, hout, npart);
}

// v[r] = w[r] / sqrt(sum_x npart[r][x]) after k_gs_update; block x = 0 stores the
// squared norm as nout[r][0] (the layout of a dots() result).
__device__ __forceinline__ void gs_normalize_item(int bx, int r, int gx, int n, int nb, const double *npart,
                                                  const cx<double> *w, cx<double> *v, cx<double> *nout)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    red[t] = t < nb ? npart[(size_t)r * nb + t] : 0.0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double nrm2 = red[0];
    if (bx == 0 && t == 0) nout[3 * r] = mk(nrm2, 0.0);
    const double sc = sqrt(nrm2);
    const double inv = sc > 1e-300 ? 1.0 / sc : 0.0;
    for (int i = bx * 256 + t; i < n; i += gx * 256) v[(size_t)r * n + i] = w[(size_t)r * n + i] * inv;
}

__global__ void __launch_bounds__(256) k_gs_normalize(int n, int nb, const double *npart, const cx<double> *w,
                                                      cx<double> *v, cx<double> *nout)
{
    gs_normalize_item(blockIdx.x, blockIdx.y, gridDim.x, n, nb, npart, w, v, nout);
}

// x[r] += sum_k V_k[r] y[r][k]
__global__ void k_multiaxpy_add(int n, int nrhs, int kcount, const cx<double> *V, const cx<double> *y, cx<double> *x)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = t / n;
    const size_t i = t % n;
    cx<double> s = x[t];
    for (int k = 0; k < kcount; ++k) s += V[((size_t)k * nrhs + r) * n + i] * y[(size_t)r * kcount + k];
    x[t] = s;
}

// ======================================================================================
// unpivoted LU (Tfactor_usy / Tfbelim, src/NC_TypeDefinition.h:542-594) — FP64
// ======================================================================================
// trailing update of step i: A[j][k] -= A[i][k] * A[j][i] / A[i][i], j, k > i
__global__ void k_lu_update(cx<double> *A, int n, int i)
{
    const int k = i + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    const int j = i + 1 + blockIdx.y;
    if (k >= n || j >= n) return;
    const cx<double> dia = A[(size_t)i * n + i];
    const cx<double> l = cdiv(A[(size_t)j * n + i], dia);
    A[(size_t)j * n + k] -= A[(size_t)i * n + k] * l;
}

// after the update: column i below and row i right of the diagonal divided by A[i][i]
__global__ void k_lu_scale(cx<double> *A, int n, int i)
{
    const int t = i + 1 + blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const cx<double> dia = A[(size_t)i * n + i];
    A[(size_t)t * n + i] = cdiv(A[(size_t)t * n + i], dia);
    A[(size_t)i * n + t] = cdiv(A[(size_t)i * n + t], dia);
}

// forward elimination, diagonal division and back substitution, one block per load case
__global__ void __launch_bounds__(1024) k_lu_solve(const cx<double> *A, int n, cx<double> *bb)
{
    __shared__ double sre[1024], sim[1024];
    cx<double> *b = bb + (size_t)blockIdx.x * n;
    for (int i = 1; i < n; ++i) {
        double re = 0.0, im = 0.0;
        for (int j = threadIdx.x; j < i; j += blockDim.x) {
            const cx<double> p = A[(size_t)i * n + j] * b[j];
            re += p.re;
            im += p.im;
        }
        sre[threadIdx.x] = re;
        sim[threadIdx.x] = im;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                sre[threadIdx.x] += sre[threadIdx.x + s];
                sim[threadIdx.x] += sim[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) b[i] -= mk(sre[0], sim[0]);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) b[i] = cdiv(b[i], A[(size_t)i * n + i]);
    __syncthreads();
    for (int i = n - 1; i > 0; --i) {
        const cx<double> bi = b[i];
        for (int j = threadIdx.x; j < i; j += blockDim.x) b[j] -= bi * A[(size_t)j * n + i];
        __syncthreads();
    }
}

// ======================================================================================
// incomplete LU preconditioner (NC_IncompleteLUDecomposition / NC_IncompleteLUForBack,
// src/NC_CommonFunctions.cpp:1356-1721) — FP64.
// Rows are scaled by d_i (the scaling vector); the pattern keeps entries with
// |d_i a_ij| > threshold and the diagonal.  ILU(0) on that pattern is unique, so the
// factors are computed row by row (IKJ, unit lower L, U with the diagonal) in level
// order: rows of one level depend only on rows of earlier levels.
// Pattern CSR: rp[n+1] (int), col ascending per row, val = scaled entries.
// ======================================================================================
__device__ __forceinline__ bool ilu_keep(cx<double> a, double d, double thr)
{
#pragma clang fp contract(off)
    const double re = a.re * d, im = a.im * d;
    return sqrt(re * re + im * im) > thr;
}

// kept entries per row of the dense operator, one wave per row
__global__ void k_ilu_count_dense(const cx<double> *A, int n, const double *d, double thr, int *cnt)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double di = d[i];
    int c = 0;
    for (int j = lane; j < n; j += 64) c += (j == i || ilu_keep(A[(size_t)i * n + j], di, thr)) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) cnt[i] = c;
}

// kept entries written in column order: ballot over 64 columns at a time
__global__ void k_ilu_fill_dense(const cx<double> *A, int n, const double *d, double thr, const int *rp, int *col,
                                 cx<double> *val)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double di = d[i];
    int base = rp[i];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        cx<double> a = mk(0.0, 0.0);
        bool keep = false;
        if (j < n) {
            a = A[(size_t)i * n + j];
            keep = j == i || ilu_keep(a, di, thr);
        }
        const unsigned long long m = __ballot(keep);
        if (keep) {
            const int q = base + __popcll(m & below);
            col[q] = j;
            val[q] = a * di;
        }
        base += __popcll(m);
    }
}

// the same over the near-field CSR rows (SL/ML-FMM)
__global__ void k_ilu_count_csr(int n, const long long *ptr, const int *ncol, const cx<double> *nval, const double *d,
                                double thr, int *cnt)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const double di = d[i];
    int c = 0;
    for (long long k = ptr[i] + lane; k < ptr[i + 1]; k += 64) c += (ncol[k] == i || ilu_keep(nval[k], di, thr)) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) cnt[i] = c;
}

// one thread per row: copy the kept entries, then insertion-sort them by column
// (near-field rows are ordered by cluster, not by column)
__global__ void k_ilu_fill_csr(int n, const long long *ptr, const int *ncol, const cx<double> *nval, const double *d,
                               double thr, const int *rp, int *col, cx<double> *val)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double di = d[i];
    const int b = rp[i];
    int q = b;
    for (long long k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int j = ncol[k];
        const cx<double> a = nval[k];
        if (j == i || ilu_keep(a, di, thr)) {
            int t = q++;
            while (t > b && col[t - 1] > j) {
                col[t] = col[t - 1];
                val[t] = val[t - 1];
                --t;
            }
            col[t] = j;
            val[t] = a * di;
        }
    }
}

// IKJ factorization in place, one work-group walking the levels of the L dependency graph
// ILU factorization level by level, a wave per row of the level, a lane per entry of
// the row held in registers while the row's lower entries k (in order) eliminate:
// l_ik = a_ik / u_kk, then a_ij -= l_ik u_kj for the later entries j of the row that row
// k's upper part holds.  From uoff[i] (computed on the host from the pattern) upd holds
// the positions of the pivots u_kk of row i's L lower entries, then for lower entry a and
// entry e (R entries) the position of u_kj or -1 at L + a R + e; a row's pivots and u_kj
// of a chunk of kFacChunk lower entries thus come in one round of loads, all from rows of
// earlier levels.  Rows longer than a wave walk row k.
constexpr int kFacChunk = 8;
__global__ void __launch_bounds__(1024) k_ilu_factor(int nlev, const int *lev_ptr, const int *lev_rows, const int *rp,
                                                     const int *col, const int *diag, const long long *uoff,
                                                     const int *upd, cx<double> *val, int *err)
{
    constexpr int NW = 1024 / 64;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int l = 0; l < nlev; ++l) {
        for (int t = lev_ptr[l] + wv; t < lev_ptr[l + 1]; t += NW) {
            const int i = lev_rows[t], r0 = rp[i], r1 = rp[i + 1], dg = diag[i];
            const int R = r1 - r0;
            if (R <= 64) {
                const bool on = lane < R;
                cx<double> v = on ? val[r0 + lane] : mk(0.0, 0.0);
                const int *um = upd + uoff[i];
                const int Lw = dg - r0;
                for (int a0 = 0; a0 < Lw; a0 += kFacChunk) {
                    cx<double> piv[kFacChunk], u[kFacChunk];
                    bool has[kFacChunk];
#pragma unroll
                    for (int c = 0; c < kFacChunk; ++c) {
                        const int a = a0 + c;
                        const bool in = a < Lw;
                        piv[c] = in ? val[um[a]] : mk(1.0, 0.0);
                        const int q = in && on ? um[Lw + (size_t)a * R + lane] : -1;
                        has[c] = q >= 0;
                        u[c] = has[c] ? val[q] : mk(0.0, 0.0);
                    }
#pragma unroll
                    for (int c = 0; c < kFacChunk; ++c) {
                        const int a = a0 + c;
                        if (r0 + a >= dg) break;
                        if (piv[c].re == 0.0 && piv[c].im == 0.0) {
                            if (lane == 0) atomicOr(err, 8);
                            continue;
                        }
                        const cx<double> lik = cdiv(mk(__shfl(v.re, a), __shfl(v.im, a)), piv[c]);
                        if (lane == a) v = lik;
                        if (lane > a && has[c]) v -= lik * u[c];
                    }
                }
                if (on) val[r0 + lane] = v;
            } else if (lane == 0) {
                for (int p = r0; p < dg; ++p) {
                    const int k = col[p];
                    const cx<double> piv = val[diag[k]];
                    if (piv.re == 0.0 && piv.im == 0.0) {
                        atomicOr(err, 8);
                        continue;
                    }
                    const cx<double> lik = cdiv(val[p], piv);
                    val[p] = lik;
                    int q = diag[k] + 1;
                    const int qe = rp[k + 1];
                    for (int p2 = p + 1; p2 < r1 && q < qe; ++p2) {
                        const int jj = col[p2];
                        while (q < qe && col[q] < jj) ++q;
                        if (q < qe && col[q] == jj) val[p2] -= lik * val[q];
                    }
                }
            }
        }
        __syncthreads();
    }
}

// Level-ordered copies of the factors for the triangular solves: position u of a sweep's
// schedule is row row[u]; its strictly lower (upper) entries are c/v[off[u] .. off[u+1])
// in the CSR order, the first kIluB of them also in cd/vd[u][kIluB] (padded with column
// 0 and value 0); d[u] is U's diagonal.  Map entries -1 pack as zero.
__global__ void k_ilu_pack(int cnt, const int *map, const cx<double> *val, cx<double> *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cnt) out[t] = map[t] >= 0 ? val[map[t]] : mk(0.0, 0.0);
}

// one triangular row: s -= sum_k v[k] y[c[k]] over [k0, k1) in increasing k, eight entries
// per step with every load issued first (the indices past k1 are clamped and skipped)
__device__ __forceinline__ cx<double> ilu_row(cx<double> s, int k0, int k1, const int *c, const cx<double> *v,
                                              const cx<double> *yr)
{
    constexpr int B = 8;
    for (int k = k0; k < k1; k += B) {
        int cc[B];
        cx<double> vv[B], yy[B];
#pragma unroll
        for (int q = 0; q < B; ++q) {
            const int kk = min(k + q, k1 - 1);
            cc[q] = c[kk];
            vv[q] = v[kk];
        }
#pragma unroll
        for (int q = 0; q < B; ++q) yy[q] = yr[cc[q]];
#pragma unroll
        for (int q = 0; q < B; ++q)
            if (k + q < k1) s -= vv[q] * yy[q];
    }
    return s;
}

// One row of a triangular sweep with everything that does not depend on the solution
// loaded ahead: entry range, the first kIluB columns and values (from the padded copies
// cd/vd, [position][kIluB]; a negative column -1 - l names lane l of the same batch), the
// sub-step of its level inside the batch, the right-hand side (forward sweep), the
// diagonal (backward).
constexpr int kIluB = ncgk::kIluB;
struct IluItem {
    int i, k0, k1, ss;
    bool on;
    int cc[kIluB];
    cx<double> vv[kIluB];
    cx<double> xv, dv;
};

// thread t's row i of the batch [b, b + w) (w <= kIluRows, i already loaded); every load
// here depends only on the position and i; d (backward sweep only) may be null.
__device__ __forceinline__ void ilu_item_load(IluItem &it, int t, int b, int w, int i, const int *off,
                                              const unsigned char *sub, const int *cd, const cx<double> *vd,
                                              const cx<double> *xr, const cx<double> *d)
{
    it.on = t < w;
    const int u = b + (it.on ? t : 0);
    it.i = i;
    it.k0 = off[u];
    it.k1 = off[u + 1];
    it.ss = sub[u];
#pragma unroll
    for (int q = 0; q < kIluB; ++q) {
        it.cc[q] = cd[(size_t)u * kIluB + q];
        it.vv[q] = vd[(size_t)u * kIluB + q];
    }
    it.xv = xr[i];
    it.dv = d ? d[u] : mk(1.0, 0.0);
}

constexpr int kIluSched = 1024;   // batches whose schedule is kept in LDS
// One triangular sweep of one load case by one work-group of kIluRows threads, a batch of
// consecutive levels (at most kIluRows rows, thread t taking position t of the batch) per
// step; its rows run level by level (sub-steps, an LDS barrier after each).  Every solved
// value goes to memory and to an LDS ring (slot = position mod kIluRing); an entry among a
// row's first kIluB reads its value from the ring while the ring still holds it (the
// host-built schedule encodes which).  The loads of the next batch's rows are requested
// at the start of a batch and stay in flight while its sub-steps run on LDS and registers
// only: the schedule sits in LDS (LDS_SCHED), the few values no longer in the ring are read
// before that request (the batch is flagged), and every select is done per component.
// Stores become visible to the work-group at the ends of the batches the schedule flags,
// before any later batch reads them from memory.  Rows past their kIluB-th entry and levels
// wider than the work-group (a batch of their own, a plain loop) read memory.  Every row
// sums its entries in the CSR order.  xr: the right-hand side (forward sweep) or the
// forward sweep's result (backward, the same vector as yr).
template <bool UPPER, bool LDS_SCHED>
__device__ __forceinline__ void ilu_sweep(const ncgk::IluSweep &S, const cx<double> *v, const cx<double> *vd,
                                          const cx<double> *d, const cx<double> *xr, cx<double> *yr,
                                          cx<double> *ring)
{
    constexpr int W = ncgk::kIluRows;
    const int t = threadIdx.x, nb = S.nbat;
    const int *row = S.row, *off = S.off, *c = S.c, *cd = S.cd;
    // the schedule (batch starts, levels and flags) in LDS for the whole sweep (LDS_SCHED),
    // so that no memory load besides the prefetch of the next batch is in flight
    __shared__ int sched[2 * kIluSched + 1];
    if (LDS_SCHED) {
        for (int q = t; q <= nb; q += W) sched[q] = S.bat_ptr[q];
        for (int q = t; q < nb; q += W) sched[kIluSched + 1 + q] = S.bat_m[q] | (S.bat_flag[q] << 16);
        __syncthreads();
    }
    auto bpv = [&](int b) { return LDS_SCHED ? sched[b] : S.bat_ptr[b]; };
    auto mfl = [&](int b) { return LDS_SCHED ? sched[kIluSched + 1 + b] : S.bat_m[b] | (S.bat_flag[b] << 16); };
    auto width = [&](int b) { return bpv(b + 1) - bpv(b); };
    auto row_of = [&](int b) { return b < nb && t < width(b) ? row[bpv(b) + t] : 0; };
    IluItem cur, nxt;
    cur.on = false;
    int rnext = row_of(1);
    if (nb > 0 && width(0) <= W)
        ilu_item_load(cur, t, bpv(0), width(0), row_of(0), off, S.sub, cd, vd, xr,
                      UPPER ? d : nullptr);
    for (int b = 0; b < nb; ++b) {
        const int b0 = bpv(b), w = bpv(b + 1) - b0;
        // this batch's rows (requested while the batch before ran) have arrived: every vector
        // memory operation but the last (this thread's store of the batch before) is done,
        // so the loads issued below are the only ones in flight during the sub-steps
        __builtin_amdgcn_s_waitcnt(0x0F71);
        // the batch's entries whose values are no longer in the ring, read from memory before
        // the next batch is requested (rare; the batch is flagged)
        cx<double> yy[kIluB];
#pragma unroll
        for (int q = 0; q < kIluB; ++q) yy[q] = mk(0.0, 0.0);
        if ((mfl(b) >> 16) & 2) {
#pragma unroll
            for (int q = 0; q < kIluB; ++q) yy[q] = yr[cur.cc[q] >= 0 ? cur.cc[q] : 0];
            __builtin_amdgcn_s_waitcnt(0x0F70);   // arrived before anything else is requested
        }
        nxt.on = false;
        const int rafter = row_of(b + 2);
        if (b + 1 < nb && width(b + 1) <= W)
            ilu_item_load(nxt, t, bpv(b + 1), width(b + 1), rnext, off, S.sub, cd, vd, xr,
                          UPPER ? d : nullptr);
        if (w <= W) {
            const cx<double> s0 = cur.xv;
            const int m = mfl(b) & 0xffff;
                for (int st = 0; st < m; ++st) {
                const bool mine = cur.on && cur.ss == st;
                if (mine) {
                    cx<double> acc = s0;
                    // every ring read first (one wait), then the entries in order; an
                    // entry past the row's end leaves acc as it is
                    cx<double> rv[kIluB];
#pragma unroll
                    for (int q = 0; q < kIluB; ++q) rv[q] = ring[cur.cc[q] < 0 ? -1 - cur.cc[q] : 0];
#pragma unroll
                    for (int q = 0; q < kIluB; ++q) {
                        // component selects (a select of the struct goes through scratch)
                        const bool ring_q = cur.cc[q] < 0, in_q = cur.k0 + q < cur.k1;
                        const cx<double> y = mk(ring_q ? rv[q].re : yy[q].re, ring_q ? rv[q].im : yy[q].im);
                        const cx<double> a = acc - cur.vv[q] * y;
                        acc.re = in_q ? a.re : acc.re;
                        acc.im = in_q ? a.im : acc.im;
                    }
                    if (cur.k1 - cur.k0 > kIluB) acc = ilu_row(acc, cur.k0 + kIluB, cur.k1, c, v, yr);
                    if (UPPER) acc = cdiv(acc, cur.dv);
                    ring[(b0 + t) % ncgk::kIluRing] = acc;
                    yr[cur.i] = acc;
                }
                if (st + 1 < m) {
                    // this level's LDS slots before the next level reads them (LDS only: the
                    // global stores complete in the background until the batch ends)
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                    __builtin_amdgcn_s_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                }
            }
        } else {
                for (int q = t; q < w; q += W) {
                const int u = b0 + q, i = row[u];
                cx<double> acc = ilu_row(xr[i], off[u], off[u + 1], c, v, yr);
                if (UPPER) acc = cdiv(acc, d[u]);
                ring[u % ncgk::kIluRing] = acc;
                yr[i] = acc;
            }
        }
        if ((mfl(b) >> 16) & 1) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }
        cur = nxt;
        rnext = rafter;
    }
}

// y = (L U)^-1 x: one work-group of kIluRows threads per load case (block r), forward
// over the L levels, backward over the U levels.  The sums run in the CSR order of each
// row, as the factorization stored them.  x and y may be the same vector (a row's
// right-hand side is read before the row is written).
__global__ void __launch_bounds__(ncgk::kIluRows) k_ilu_solve(int n, ncgk::IluSweep L, ncgk::IluSweep U,
                                                             const cx<double> *x, cx<double> *y)
{
    __shared__ cx<double> ring[ncgk::kIluRing];
    const size_t r = blockIdx.x;
    if (L.nbat <= kIluSched)
        ilu_sweep<false, true>(L, (const cx<double> *)L.v, (const cx<double> *)L.vd, nullptr, x + r * n, y + r * n, ring);
    else
        ilu_sweep<false, false>(L, (const cx<double> *)L.v, (const cx<double> *)L.vd, nullptr, x + r * n, y + r * n, ring);
    __syncthreads();
    if (U.nbat <= kIluSched)
        ilu_sweep<true, true>(U, (const cx<double> *)U.v, (const cx<double> *)U.vd, (const cx<double> *)U.d, y + r * n,
                              y + r * n, ring);
    else
        ilu_sweep<true, false>(U, (const cx<double> *)U.v, (const cx<double> *)U.vd, (const cx<double> *)U.d, y + r * n,
                               y + r * n, ring);
}

// ======================================================================================
// evaluation grid: potential and gradient (NC_IntegrationTBEM, src/NC_PostProcessing.cpp:1264-1318)
// one wave per evaluation point, lanes over the boundary elements.
// ======================================================================================
// pot/vel of every element for every load case from the solution
// (NC_WriteResultsObjectMesh, src/NC_PostProcessing.cpp:399-421)
__global__ void k_field_prep(int n, int nrhs, const int8_t *kind, const double *adm, double tau,
                             const cx<double> *bcval, const cx<double> *x, cx<double> *pot, cx<double> *vel)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int j = t % n;
    if (kind[j] == 1) {
        pot[t] = bcval[t];
        vel[t] = x[t];
    } else {
        pot[t] = x[t];
        cx<double> v = bcval[t];
        if (kind[j] == 2) v -= x[t] * mk(adm[2 * j], adm[2 * j + 1]) * tau;
        vel[t] = v;
    }
}

struct FieldAcc {
    cx<double> K1, K2, K3[3], K4[3];
};

template <typename T>
__device__ __forceinline__ void field_point(FieldAcc &f, T qx, T qy, T qz, T nyx, T nyy, T nyz, T wj, T hk, T k2)
{
    const T r = sqrt(qx * qx + qy * qy + qz * qz);
    const T ir = T(1) / r;
    const T vx = qx * ir, vy = qy * ir, vz = qz * ir;
    T sn, cs;
    sincos_t(hk * r, &sn, &cs);
    const T a = T(1.0 / kPi4) * ir * wj;
    const cx<T> G = {cs * a, sn * a};                               // G * w J
    const cx<T> Gi = {-ir * G.re - hk * G.im, -ir * G.im + hk * G.re};
    const T drdn = vx * nyx + vy * nyy + vz * nyz;
    const cx<T> f1 = G * mk(drdn * (k2 - T(3) * ir * ir), hk * T(3) * drdn * ir);
    const cx<T> f2 = Gi * (-ir);
    const cx<T> dgdn = Gi * drdn;
    f.K1 += mk(double(dgdn.re), double(dgdn.im));
    f.K2 += mk(double(G.re), double(G.im));
    const T v[3] = {vx, vy, vz}, ny[3] = {nyx, nyy, nyz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const cx<T> t3 = f1 * v[d] + f2 * ny[d];
        const cx<T> t4 = Gi * (-v[d]);
        f.K3[d] += mk(double(t3.re), double(t3.im));
        f.K4[d] += mk(double(t4.re), double(t4.im));
    }
}

constexpr int kFieldChunk = 4;

template <typename T>
__global__ void __launch_bounds__(256) k_field(Elements E, BMConst bm, const double *pts, int m, int nrhs, int rhs0,
                                               const cx<double> *pot, const cx<double> *vel, const double *ps, int nps,
                                               const double *pw, int npw, cx<double> *phi, cx<double> *grad, int *err,
                                               const int *pt_ev, const int *el_ptr, const int *el_idx,
                                               const cx<double> *fphi, const cx<double> *fgrad,
                                               unsigned long long *evals)
{
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (p >= m) return;
    const int n = E.n;
    const double x[3] = {pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
    const int nr = min(kFieldChunk, nrhs - rhs0);
    cx<double> aphi[kFieldChunk], agr[kFieldChunk][3];
    for (int r = 0; r < kFieldChunk; ++r) {
        aphi[r] = mk(0.0, 0.0);
        agr[r][0] = agr[r][1] = agr[r][2] = mk(0.0, 0.0);
    }
    const T hk = T(bm.hk), k2 = T(bm.k2);
    unsigned long long nev = 0;
    // all elements (dense method) or the elements of the near clusters of the point's
    // evaluation cluster (FMM)
    const int k0 = el_ptr ? el_ptr[pt_ev[p]] : 0, k1 = el_ptr ? el_ptr[pt_ev[p] + 1] : n;
    for (int kk = k0 + lane; kk < k1; kk += 64) {
        const int j = el_ptr ? el_idx[kk] : kk;
        FieldAcc f;
        f.K1 = f.K2 = mk(0.0, 0.0);
        for (int d = 0; d < 3; ++d) f.K3[d] = f.K4[d] = mk(0.0, 0.0);
        LocalElem L;
        load_local(E, j, x, L);
        const double ratdis = ratio_to_centre(E.cen[j], E.cen[n + j], E.cen[2 * n + j], x, L.area);
        if (ratdis < kTolf) {
            const int e = visit_adaptive(L, [&](double s, double t, double w) {
                double q[3], ny[3], J;
                point_on(L, s, t, q, ny, J);
                field_point<double>(f, q[0], q[1], q[2], ny[0], ny[1], ny[2], w * J, bm.hk, bm.k2);
                ++nev;
            });
            if (e) atomicOr(err, e);
        } else {
            const int mo = gauss_order(ratdis);
            if (L.nv == 3) {
                const int off = tri_rule_offset(mo), cnt = tri_rule_count(mo);
                for (int g = 0; g < cnt; ++g) {
                    double q[3], ny[3], J;
                    point_on(L, c_tri[off + g][0], c_tri[off + g][1], q, ny, J);
                    field_point<T>(f, T(q[0]), T(q[1]), T(q[2]), T(ny[0]), T(ny[1]), T(ny[2]),
                                   T(c_tri[off + g][2] * J), hk, k2);
                }
                nev += cnt;
            } else {
                const int base = mo * (mo - 1) / 2;
                for (int u = 0; u < mo; ++u)
                    for (int v = 0; v < mo; ++v) {
                        double q[3], ny[3], J;
                        point_on(L, c_gl_x[base + u], c_gl_x[base + v], q, ny, J);
                        field_point<T>(f, T(q[0]), T(q[1]), T(q[2]), T(ny[0]), T(ny[1]), T(ny[2]),
                                       T(c_gl_w[base + u] * c_gl_w[base + v] * J), hk, k2);
                    }
                nev += mo * mo;
            }
        }
        for (int r = 0; r < nr; ++r) {
            const cx<double> pj = pot[(size_t)(rhs0 + r) * n + j], vj = vel[(size_t)(rhs0 + r) * n + j];
            aphi[r] += f.K1 * pj - f.K2 * vj;
            for (int d = 0; d < 3; ++d) agr[r][d] += f.K3[d] * pj - f.K4[d] * vj;
        }
    }
    if (evals) wave_count(evals, nev);
    for (int r = 0; r < nr; ++r) {
        double v[8] = {aphi[r].re, aphi[r].im, agr[r][0].re, agr[r][0].im, agr[r][1].re, agr[r][1].im,
                       agr[r][2].re, agr[r][2].im};
        for (int c = 0; c < 8; ++c)
            for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
        if (lane != 0) continue;
        cx<double> ph = mk(v[0], v[1]) * bm.tau;
        cx<double> gr[3] = {mk(v[2], v[3]) * bm.tau, mk(v[4], v[5]) * bm.tau, mk(v[6], v[7]) * bm.tau};
        const int rr = rhs0 + r;
        if (fphi) {   // far clusters through the multipole expansion
            ph += fphi[(size_t)rr * m + p];
            for (int q = 0; q < 3; ++q) gr[q] += fgrad[((size_t)rr * m + p) * 3 + q];
        }
        // incident field (NC_ContributionIncidentWaves, src/NC_PostProcessing.cpp:776-866)
        for (int w = 0; w < npw; ++w) {
            const double *d = pw + 6 * w;
            if ((int)d[5] != rr) continue;
            double sn, cs;
            sincos((d[0] * x[0] + d[1] * x[1] + d[2] * x[2]) * bm.hk, &sn, &cs);
            const cx<double> g = mk(cs, sn) * mk(d[3], d[4]);
            ph += g;
            const cx<double> gv = mk(-g.im * bm.hk, g.re * bm.hk);
            for (int q = 0; q < 3; ++q) gr[q] += gv * d[q];
        }
        for (int s = 0; s < nps; ++s) {
            const double *y = ps + 6 * s;
            if ((int)y[5] != rr) continue;
            double v2[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
            const double R = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
            v2[0] /= R; v2[1] /= R; v2[2] /= R;
            double sn, cs;
            sincos(bm.hk * R, &sn, &cs);
            const double f = 1.0 / kPi4 / R;
            const cx<double> g = mk(cs * f, sn * f) * mk(y[3], y[4]);
            ph += g;
            const cx<double> gv = mk(-1.0 / R, bm.hk) * g;
            for (int q = 0; q < 3; ++q) gr[q] += gv * v2[q];
        }
        phi[(size_t)rr * m + p] = ph;
        if (grad)
            for (int q = 0; q < 3; ++q) grad[((size_t)rr * m + p) * 3 + q] = gr[q];
    }
}

// ======================================================================================
// Fast multipole operators (NumCalc's SL-FMM and direct ML-FMM: per level an aggregation
// T, a diagonal translation D between interacting clusters and a disaggregation S;
// src/NC_EquationSystem.cpp:305-1240, src/NC_CommonFunctions.cpp:1247-1316).
// Layouts (p, the unit-sphere point, fastest so that lanes coalesce):
//   T[(cl_ptr[c] + local element) * P + p]   S[cl_ptr[c] * P + p * m_c + local element]
//   D[q * P + p], q = interaction
//   t, s work vectors [nrhs][ncl * P]
// ======================================================================================
constexpr int kMaxTerms = 80;

// Series coefficients a_n = (i h)^n (2n + 1) h_n(x), n < N, into o[2n], o[2n + 1], with
// h_n = j_n + i h y_n (NC_SphericalHankel, src/NC_3dFunctions.cpp:821-886): y_n by upward
// recurrence, j_n by downward recurrence from the continued fraction of j_N/j_{N-1}
// (modified Lentz), normalised with j_0 = sin x / x (or j_0' when j_0 ~ 0).  When
// x > N + 10 every order is well below the argument and j_n also comes from the upward
// recurrence (the continued fraction would take O(x) steps there).  The recurrences run in
// registers (per-thread arrays took 2.5 KB of scratch per thread): the downward one twice,
// first for the normalisation, then storing j_n in o[2n]; one upward pass forms y_n and a_n.
__device__ void hankel_series(int N, double x, double h, double *o)
{
    const double sx = sin(x), cxs = cos(x);
    const bool up = x > (double)N + 10.0;
    if (!up) {
        const double nu = (double)(N - 1);
        double di = (2.0 * (nu + 1.0) + 1.0) / x, C = di, D = 0.0, err = 1.0;
        for (int j = 1; err > 1.0e-9 && j < 100000; ++j) {
            const double b = (2.0 * (nu + j + 1.0) + 1.0) / x;
            D = b - D;
            if (D == 0.0) D = 1.0e-30;
            D = 1.0 / D;
            C = b - 1.0 / C;
            if (C == 0.0) C = 1.0e-30;
            di = di * C * D;
            err = fabs(C * D - 1.0);
        }
        const double gnu = nu / x - 1.0 / di;
        // downward recurrence twice: first for g_0 (the normalisation), then to store j_n
        double g = 1.0, dg = gnu;
        for (int i = N - 2; i >= 0; --i) {
            const double gi = ((double)i + 2.0) / x * g + dg;
            dg = (double)i / x * gi - g;
            g = gi;
        }
        const double dp = fabs(g) > 1.0e-5 ? sx / x / g : (cxs - sx / x) / x / dg;
        g = 1.0;
        dg = gnu;
        o[2 * (N - 1)] = dp * g;
        for (int i = N - 2; i >= 0; --i) {
            const double gi = ((double)i + 2.0) / x * g + dg;
            dg = (double)i / x * gi - g;
            g = gi;
            o[2 * i] = dp * g;
        }
    }
    double jm2 = 0.0, jm1 = 0.0, ym2 = 0.0, ym1 = 0.0;
    cx<double> zi = mk(1.0, 0.0);
    for (int n = 0; n < N; ++n) {
        double yn, jn;
        if (n == 0) yn = -cxs / x;
        else if (n == 1) yn = -(cxs / x + sx) / x;
        else yn = ym1 * (double)(2 * n - 1) / x - ym2;
        if (up) {
            if (n == 0) jn = sx / x;
            else if (n == 1) jn = (sx / x - cxs) / x;
            else jn = jm1 * (double)(2 * n - 1) / x - jm2;
            jm2 = jm1;
            jm1 = jn;
        } else {
            jn = o[2 * n];
        }
        ym2 = ym1;
        ym1 = yn;
        const cx<double> an = zi * mk(jn, h * yn) * double(2 * n + 1);
        o[2 * n] = an.re;
        o[2 * n + 1] = an.im;
        zi = mk(-zi.im * h, zi.re * h);
    }
}

// translation function mu(s) = sum_{n<L} (i h)^n (2n+1) h_n P_n(dhat . s)
// (NC_BuildDforSLFMM, src/NC_EquationSystem.cpp:517-537)
__device__ __forceinline__ cx<double> translation(int L, double c, double h, const double *hre, const double *him)
{
    cx<double> mu = mk(0.0, 0.0), zi = mk(1.0, 0.0);
    double p0 = 1.0, p1 = c;
    for (int n = 0; n < L; ++n) {
        double pn;
        if (n == 0) pn = 1.0;
        else if (n == 1) pn = c;
        else {
            pn = (p1 * c * (double)(2 * n - 1) - p0 * (double)(n - 1)) / (double)n;
            p0 = p1;
            p1 = pn;
        }
        mu += zi * mk(hre[n], him[n]) * (pn * (double)(2 * n + 1));
        zi = mk(-zi.im * h, zi.re * h);
    }
    return mu;
}

// D values of one level (NC_BuildD*FMM, src/NC_EquationSystem.cpp:488-547, 838-900):
// D_q(s_p) = sum_{n<L} (i h)^n (2n + 1) h_n(k |C_i - C_j|) P_n(d_q . s_p) per interaction
// q = (cluster i, far cluster j).  First pass, a thread per interaction: the direction
// d_q and the series coefficients a_n = (i h)^n (2n + 1) h_n into coef[q][3 + 2L].
__global__ void __launch_bounds__(128) k_fmm_D_coef(int nfan, int L, double k, double h, const double *ccen,
                                                    const int *fan_row, const int *fan_idx, double *coef)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nfan) return;
    const int i = fan_row[q], j = fan_idx[q];
    const double d[3] = {ccen[3 * i] - ccen[3 * j], ccen[3 * i + 1] - ccen[3 * j + 1], ccen[3 * i + 2] - ccen[3 * j + 2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double *o = coef + (size_t)q * (3 + 2 * L);
    o[0] = d[0] / r;
    o[1] = d[1] / r;
    o[2] = d[2] / r;
    hankel_series(L, r * k, h, o + 3);
}

// Second pass: one block per interaction, threads over the sphere points, the series
// summed in FP64 with the Legendre recurrence P_n = ((2n-1) c P_{n-1} - (n-1) P_{n-2}) / n
// (the two ratios tabulated per block); stored as S.
template <typename S>
__global__ void __launch_bounds__(128) k_fmm_D(int L, int P, const double *coef, const double *u, cx<S> *Dv)
{
    __shared__ double cf[3 + 2 * kMaxTerms], ca[kMaxTerms], cb[kMaxTerms];
    const int q = blockIdx.x, w = 3 + 2 * L;
    for (int t = threadIdx.x; t < w; t += blockDim.x) cf[t] = coef[(size_t)q * w + t];
    for (int t = threadIdx.x; t < L; t += blockDim.x) {
        ca[t] = t >= 2 ? (double)(2 * t - 1) / (double)t : 0.0;
        cb[t] = t >= 2 ? (double)(t - 1) / (double)t : 0.0;
    }
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        const double c = cf[0] * u[3 * p] + cf[1] * u[3 * p + 1] + cf[2] * u[3 * p + 2];
        double mre = cf[3], mim = cf[4];
        double p0 = 1.0, p1 = c;
        if (L > 1) {
            mre += cf[5] * c;
            mim += cf[6] * c;
        }
        for (int n = 2; n < L; ++n) {
            const double pn = ca[n] * c * p1 - cb[n] * p0;
            p0 = p1;
            p1 = pn;
            mre += cf[3 + 2 * n] * pn;
            mim += cf[4 + 2 * n] * pn;
        }
        Dv[(size_t)q * P + p] = cx<S>{S(mre), S(mim)};
    }
}

// Gauss points of the FMM aggregation rule (order 2: TR4 on triangles, 2x2 on quads,
// GauFMBOrd3 = 2 of src/NC_EquationSystem.cpp:73)
__device__ __forceinline__ int fmm_gauss(int nv, int g, double &s, double &t, double &w)
{
    if (nv == 3) {
        s = c_tri[g][0]; t = c_tri[g][1]; w = c_tri[g][2];
        return 4;
    }
    s = c_gl_x[1 + g / 2]; t = c_gl_x[1 + g % 2]; w = c_gl_w[1 + g / 2] * c_gl_w[1 + g % 2];
    return 4;
}

// T values of one level (NC_ComputeEntriesTforFMM, src/NC_EquationSystem.cpp:1150-1240).
// A work-group takes 4 elements x 64 sphere points: the 4 aggregation Gauss points of
// each element (point relative to the cluster centre, normal, weight x Jacobian) are
// formed once in FP64 into LDS, then each thread forms one (element, point) entry.
template <typename S>
__global__ void __launch_bounds__(256) k_fmm_T(Elements E, BMConst bm, const double *adm, int P, const double *ccen,
                                               const int *elem_cl, const int *elem_pos, const double *u, cx<S> *Tv)
{
    __shared__ double gp[4][4][7];   // [element][gauss] q(3), ny(3), w*J
    const int le = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + le, p = blockIdx.y * 64 + lane;
    if (lane < 4 && e < E.n) {
        const int c = elem_cl[e];
        const double C[3] = {ccen[3 * c], ccen[3 * c + 1], ccen[3 * c + 2]};
        LocalElem L;
        load_local(E, e, C, L);   // nodes relative to the cluster centre
        double s, tt, w;
        fmm_gauss(L.nv, lane, s, tt, w);
        double q[3], ny[3], J;
        point_on(L, s, tt, q, ny, J);
        for (int d = 0; d < 3; ++d) {
            gp[le][lane][d] = q[d];
            gp[le][lane][3 + d] = ny[d];
        }
        gp[le][lane][6] = w * J;
    }
    __syncthreads();
    if (e >= E.n || p >= P) return;
    const double su[3] = {u[3 * p], u[3 * p + 1], u[3 * p + 2]};
    const int kind = E.kind[e];
    const cx<double> ad = mk(adm[2 * e], adm[2 * e + 1]);
    const bool has_adm = kind == 2 && (ad.re != 0.0 || ad.im != 0.0);
    cx<double> acc = mk(0.0, 0.0);
    for (int g = 0; g < 4; ++g) {
        const double *G = gp[le][g];
        const double wga = G[6];
        const double dw = -(G[0] * su[0] + G[1] * su[1] + G[2] * su[2]) * bm.hk;
        // the phase in the storage precision (FP32 in the mixed form), formed in FP64
        S sn, cs;
        sincos_t(S(dw), &sn, &cs);
        cx<double> z = mk(double(cs), double(sn));
        if (kind == 1) {
            z = z * wga;
            z = mk(z.im, -z.re);   // times -i
        } else {
            const double d1 = (G[3] * su[0] + G[4] * su[1] + G[5] * su[2]) * bm.k * bm.tau;
            if (!has_adm) z = z * (d1 * wga);
            else z = z * (mk(d1 - bm.hk / bm.k * ad.im, bm.hk / bm.k * ad.re) * wga);
        }
        acc += z;
    }
    acc = acc * (bm.k / kPi4);
    Tv[(size_t)elem_pos[e] * P + p] = cx<S>{S(acc.re), S(acc.im)};
}

// T-vector of prescribed values of one level (the right-hand side's far part), thread per
// (cluster, sphere point), summed over the cluster's elements
__global__ void __launch_bounds__(256) k_fmm_Tvec_bc(Elements E, BMConst bm, int P, int ncl, const double *ccen,
                                                     const int *cl_ptr, const int *cl_elem, const double *u,
                                                     const cx<double> *bcval, int nrhs, cx<double> *tv)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncl * P) return;
    const int c = (int)(t / P), p = (int)(t % P);
    const double C[3] = {ccen[3 * c], ccen[3 * c + 1], ccen[3 * c + 2]};
    const double su[3] = {u[3 * p], u[3 * p + 1], u[3 * p + 2]};
    for (int r = 0; r < nrhs; ++r) {
        cx<double> acc = mk(0.0, 0.0);
        for (int k = cl_ptr[c]; k < cl_ptr[c + 1]; ++k) {
            const int e = cl_elem[k];
            const cx<double> bv = bcval[(size_t)r * E.n + e];
            if (bv.re == 0.0 && bv.im == 0.0) continue;
            LocalElem L;
            load_local(E, e, C, L);
            const int kind = E.kind[e];
            for (int g = 0; g < 4; ++g) {
                double s, tt, w;
                fmm_gauss(L.nv, g, s, tt, w);
                double q[3], ny[3], J;
                point_on(L, s, tt, q, ny, J);
                const double wga = w * J;
                const double dw = -(q[0] * su[0] + q[1] * su[1] + q[2] * su[2]) * bm.hk;
                double sn, cs;
                sincos(dw, &sn, &cs);
                const cx<double> z = mk(cs, sn);
                if (kind == 1) {
                    const cx<double> zr = mk((z * wga).im, -(z * wga).re);
                    const double d1 = (ny[0] * su[0] + ny[1] * su[1] + ny[2] * su[2]) * bm.k * bm.tau;
                    acc -= zr * bv * d1;
                } else {
                    acc += z * (bv * wga);
                }
            }
        }
        const double f = bm.k / kPi4 * bm.tau * (bm.hk / bm.k);
        tv[(size_t)r * ncl * P + t] = mk(-acc.im * f, acc.re * f);
    }
}

// S values of one level (NC_BuildS*FMM, src/NC_EquationSystem.cpp:550-609, 903-977),
// stored per cluster column-major: the block of cluster c starts at cl_ptr[c] * P and
// holds [point p][local element a] with a running fastest (thread per stored entry).
template <typename S>
__global__ void __launch_bounds__(256) k_fmm_S(Elements E, BMConst bm, int P, const double *ccen, const int *elem_cl,
                                               const int *cl_ptr, const int *cl_elem, const double *u,
                                               const double *wsph, cx<S> *Sv)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)E.n * P) return;
    const int c = elem_cl[cl_elem[t / P]];
    const int b = cl_ptr[c], mc = cl_ptr[c + 1] - b;
    const size_t off = t - (size_t)b * P;
    const int p = (int)(off / mc), e = cl_elem[b + (int)(off % mc)];
    const double d[3] = {E.cen[e] - ccen[3 * c], E.cen[E.n + e] - ccen[3 * c + 1], E.cen[2 * E.n + e] - ccen[3 * c + 2]};
    const double su[3] = {u[3 * p], u[3 * p + 1], u[3 * p + 2]};
    const double d0 = (d[0] * su[0] + d[1] * su[1] + d[2] * su[2]) * bm.hk;
    // the phase in the storage precision (FP32 in the mixed form), formed in FP64
    S sn, cs;
    sincos_t(S(d0), &sn, &cs);
    const cx<double> z0 = mk(double(cs), double(sn)) * wsph[p];
    const double dn = (su[0] * E.nrm[e] + su[1] * E.nrm[E.n + e] + su[2] * E.nrm[2 * E.n + e]) * bm.hk;
    const cx<double> zd = mk(-z0.im * dn, z0.re * dn);
    const cx<double> beta_tau = mk(0.0, bm.beta_im * bm.tau);
    const cx<double> v = z0 + zd * beta_tau;
    Sv[t] = cx<S>{S(v.re), S(v.im)};
}

// CSR position of every subdivided pair (i, j) of the dense near-pair table inside the
// FMM near field: the near-list slot of j's cluster in i's cluster, its element offset
// there, and j's place in its cluster; -1 when the clusters are far from each other.
__global__ void k_adj_positions(long long npairs, const int *rows, const int *cols, const int *elem_cl,
                                const int *elem_pos, const int *cl_ptr, const int *nn_ptr, const int *nn_idx,
                                const int *nn_off, const long long *near_ptr, long long *pos)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    const int i = rows[q], j = cols[q];
    const int ci = elem_cl[i], cj = elem_cl[j];
    long long r = -1;
    for (int t = nn_ptr[ci]; t < nn_ptr[ci + 1]; ++t)
        if (nn_idx[t] == cj) {
            r = near_ptr[i] + nn_off[t] + (elem_pos[j] - cl_ptr[cj]);
            break;
        }
    pos[q] = r;
}

// near-field entries NumCalc subdivides: adaptive quadrature in FP64, one thread per pair
template <typename S>
__global__ void __launch_bounds__(128) k_fmm_near_adj(Elements E, BMConst bm, const double *adm, const int *rows,
                                                      const int *cols, const long long *pos, long long npairs,
                                                      cx<S> *vals, int *err, unsigned long long *evals)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long nev = 0;
    if (q < npairs && pos[q] >= 0) {
    const int i = rows[q], j = cols[q];
    const double x[3] = {E.cen[i], E.cen[E.n + i], E.cen[2 * E.n + i]};
    const double nx[3] = {E.nrm[i], E.nrm[E.n + i], E.nrm[2 * E.n + i]};
    LocalElem L;
    load_local(E, j, x, L);
    BMAcc<double> acc;
    acc.G = acc.H = acc.HT = acc.E = mk(0.0, 0.0);
    const int e = visit_adaptive(L, [&](double s, double t, double w) {
        double q2[3], ny[3], J;
        point_on(L, s, t, q2, ny, J);
        bm_point<double>(acc, q2[0], q2[1], q2[2], ny[0], ny[1], ny[2], nx[0], nx[1], nx[2], w * J, bm.hk, bm.k2);
        ++nev;
    });
    if (e) atomicOr(err, e);
    const cx<double> a = bm_entry<double>(acc, E.kind[j], mk(adm[2 * j], adm[2 * j + 1]), bm.tau, bm.beta_im);
    vals[pos[q]] = cx<S>{S(a.re), S(a.im)};
    }
    if (evals) wave_count(evals, nev);
}

// One column element of the regular near-field quadrature: its corners (FP64), the
// triangle edges from corner 2, normal and Jacobian in the quadrature's arithmetic T.
template <typename T>
struct NearCol {
    int nv, kind;
    cx<T> adm;
    double cj[3], area;
    double P[4][3];
    T ny[3], jac, e1[3], e2[3];
};

template <typename T>
__device__ __forceinline__ void near_col_load(const Elements &E, const double *adm, int j, NearCol<T> &C)
{
    const int n = E.n;
    C.nv = E.nv[j];
    C.kind = E.kind[j];
    C.adm = cx<T>{T(adm[2 * j]), T(adm[2 * j + 1])};
    C.cj[0] = E.cen[j];
    C.cj[1] = E.cen[n + j];
    C.cj[2] = E.cen[2 * n + j];
    C.area = E.area[j];
    for (int d = 0; d < 3; ++d) C.ny[d] = T(E.nrm[d * n + j]);
    C.jac = T(E.jac[j]);
    for (int a = 0; a < 4; ++a)
        for (int d = 0; d < 3; ++d) C.P[a][d] = a < C.nv ? E.X[(3 * a + d) * n + j] : 0.0;
    // triangle: x(s,t) = x2 + s (x0 - x2) + t (x1 - x2)
    for (int d = 0; d < 3; ++d) {
        C.e1[d] = T(C.P[0][d] - C.P[2][d]);
        C.e2[d] = T(C.P[1][d] - C.P[2][d]);
    }
}

// Entry (row i, column C) of the regular near field: false for the diagonal and for a
// subdivided pair (those come from k_assemble_diag and k_fmm_near_adj).  nev counts the
// Green's-function evaluations.
template <typename T>
__device__ __forceinline__ bool near_entry(const Elements &E, const BMConst &bm, const NearCol<T> &C, int i, int j,
                                           cx<T> &out, unsigned long long &nev)
{
    const int n = E.n;
    if (i == j) return false;
    const double x[3] = {E.cen[i], E.cen[n + i], E.cen[2 * n + i]};
    const double ratdis = ratio_to_centre(C.cj[0], C.cj[1], C.cj[2], x, C.area);
    if (ratdis < kTolf) return false;
    const T nxx = T(E.nrm[i]), nxy = T(E.nrm[n + i]), nxz = T(E.nrm[2 * n + i]);
    const T hk = T(bm.hk), k2 = T(bm.k2), tau = T(bm.tau), bim = T(bm.beta_im);
    const int mg = gauss_order(ratdis);
    BMAcc<T> acc;
    acc.G = acc.H = acc.HT = acc.E = mk(T(0), T(0));
    if (C.nv == 3) {
        const T px = T(C.P[2][0] - x[0]), py = T(C.P[2][1] - x[1]), pz = T(C.P[2][2] - x[2]);
        const int off = tri_rule_offset(mg), cnt = tri_rule_count(mg);
        for (int g = 0; g < cnt; ++g) {
            const T s = T(c_tri[off + g][0]), t = T(c_tri[off + g][1]), wg = T(c_tri[off + g][2]);
            bm_point<T>(acc, px + s * C.e1[0] + t * C.e2[0], py + s * C.e1[1] + t * C.e2[1],
                        pz + s * C.e1[2] + t * C.e2[2], C.ny[0], C.ny[1], C.ny[2], nxx, nxy, nxz, wg * C.jac, hk, k2);
        }
        nev += cnt;
    } else {
        LocalElem L;
        L.nv = 4;
        for (int b = 0; b < 4; ++b)
            for (int d = 0; d < 3; ++d) L.p[b][d] = C.P[b][d] - x[d];
        const int base = mg * (mg - 1) / 2;
        for (int s2 = 0; s2 < mg; ++s2)
            for (int t2 = 0; t2 < mg; ++t2) {
                double q2[3], ny[3], J;
                point_on(L, c_gl_x[base + s2], c_gl_x[base + t2], q2, ny, J);
                bm_point<T>(acc, T(q2[0]), T(q2[1]), T(q2[2]), T(ny[0]), T(ny[1]), T(ny[2]), nxx, nxy, nxz,
                            T(c_gl_w[base + s2] * c_gl_w[base + t2] * J), hk, k2);
            }
        nev += mg * mg;
    }
    out = bm_entry<T>(acc, C.kind, C.adm, tau, bim);
    return true;
}

// Regular near-field entries (NC_ComputeEntriesNforFMM, src/NC_EquationSystem.cpp:1243-1405,
// without the subdivided pairs, which k_fmm_near_adj writes).  The rows of a finest
// cluster share their column list, so a work-group takes a row chunk of one cluster
// (ch_*, at most 16 rows) and a segment of 256 columns: each thread holds one column
// element's geometry in registers and walks the chunk's rows, whose collocation data is
// wave-uniform; a row's entries are stored coalesced.  T is the arithmetic of the
// quadrature (FP32 in the mixed form, after the geometry differences are formed in FP64).
template <typename S, typename T>
__global__ void __launch_bounds__(256) k_fmm_near(Elements E, BMConst bm, const double *adm, const int *ch_cl,
                                                  const int *ch_r0, const int *ch_m, const int *cl_ptr,
                                                  const int *cl_elem, const long long *ptr, const int *col, cx<S> *vals,
                                                  unsigned long long *evals)
{
    const int q = blockIdx.x, c = ch_cl[q], r0 = ch_r0[q], m = ch_m[q];
    const long long cb = ptr[cl_elem[cl_ptr[c]]];
    const int w = (int)(ptr[cl_elem[cl_ptr[c]] + 1] - cb);
    const int u = blockIdx.y * 256 + threadIdx.x;
    unsigned long long nev = 0;
    if (u < w) {
        const int j = col[cb + u];
        NearCol<T> C;
        near_col_load(E, adm, j, C);
        for (int a = 0; a < m; ++a) {
            const int i = cl_elem[r0 + a];
            cx<T> e;
            if (near_entry(E, bm, C, i, j, e, nev)) vals[ptr[i] + u] = cx<S>{S(e.re), S(e.im)};
        }
    }
    if (evals) wave_count(evals, nev);
}

// The same entries with one wave per collocation row: the work-group stages the
// geometry of its 256 columns in LDS, then each wave takes every 4th row of the chunk
// and its lanes walk the columns 64 at a time, reading the column from LDS.  Same
// arithmetic per entry as k_fmm_near (near_entry).
template <typename S, typename T>
__global__ void __launch_bounds__(256) k_fmm_near_rows(Elements E, BMConst bm, const double *adm, const int *ch_cl,
                                                       const int *ch_r0, const int *ch_m, const int *cl_ptr,
                                                       const int *cl_elem, const long long *ptr, const int *col,
                                                       cx<S> *vals, unsigned long long *evals)
{
    __shared__ NearCol<T> cols[256];
    const int q = blockIdx.x, c = ch_cl[q], r0 = ch_r0[q], m = ch_m[q];
    const long long cb = ptr[cl_elem[cl_ptr[c]]];
    const int w = (int)(ptr[cl_elem[cl_ptr[c]] + 1] - cb);
    const int u0 = blockIdx.y * 256, nu = min(256, w - u0);
    if (threadIdx.x < nu) near_col_load(E, adm, col[cb + u0 + threadIdx.x], cols[threadIdx.x]);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned long long nev = 0;
    for (int a = wv; a < m; a += 4) {
        const int i = cl_elem[r0 + a];
        for (int v0 = 0; v0 < nu; v0 += 64) {
            const int v = v0 + lane;
            if (v < nu) {
                const NearCol<T> C = cols[v];
                cx<T> e;
                if (near_entry(E, bm, C, i, col[cb + u0 + v], e, nev)) vals[ptr[i] + u0 + v] = cx<S>{S(e.re), S(e.im)};
            }
        }
    }
    if (evals) wave_count(evals, nev);
}

// near-field CSR of the finest level: row i holds the elements of the near clusters of
// its cluster, cluster by cluster (near lists ascending, elements in cluster order).
// One wave per row; ptr comes from the host (row lengths are per-cluster sums).
__global__ void __launch_bounds__(256) k_near_csr(int n, const int *elem_cl, const int *cl_ptr, const int *cl_elem,
                                                  const int *nn_ptr, const int *nn_idx, const long long *ptr,
                                                  int *rows, int *cols, long long *diag)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int c = elem_cl[i];
    long long q = ptr[i];
    for (int t = nn_ptr[c]; t < nn_ptr[c + 1]; ++t) {
        const int j = nn_idx[t];
        const int b = cl_ptr[j], sz = cl_ptr[j + 1] - b;
        for (int u = lane; u < sz; u += 64) {
            const int e = cl_elem[b + u];
            if (rows) rows[q + u] = i;
            cols[q + u] = e;
            if (e == i) diag[i] = q + u;
        }
        q += sz;
    }
}

// complex multiply-accumulate into separate real and imaginary sums
__device__ __forceinline__ void cmac(double &re, double &im, double ar, double ai, cx<double> v)
{
    re += ar * v.re - ai * v.im;
    im += ar * v.im + ai * v.re;
}

// ---- block-Jacobi preconditioner over the finest clusters ----------------------------
// Blocks are runs of at most kBlk consecutive elements of a finest cluster's element
// list.  A block's rows hold its columns as a contiguous piece of the near-field CSR
// row: the cluster's own segment starts self_off[c] after the row start, the block
// blk_b0[b] - cl_ptr[c] further.  Inverses are stored row-major m x m at boff[b].
constexpr int kBlk = 64;

template <typename S, typename TB>
__global__ void __launch_bounds__(256) k_bj_gather(int n, const int *elem_blk, const int *elem_pos, const int *blk_cl,
                                                   const int *blk_b0, const int *blk_m, const int *cl_ptr,
                                                   const long long *near_ptr, const int *self_off, const cx<S> *nval,
                                                   const long long *boff, cx<TB> *blk)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int b = elem_blk[i], c = blk_cl[b], m = blk_m[b], a = elem_pos[i] - blk_b0[b];
    const cx<S> *src = nval + near_ptr[i] + self_off[c] + (blk_b0[b] - cl_ptr[c]);
    cx<TB> *dst = blk + boff[b] + (size_t)a * m;
    for (int j = lane; j < m; j += 64) dst[j] = cx<TB>{TB(src[j].re), TB(src[j].im)};
}

// the same blocks for a dense operator: block b holds the unknowns [blk_b0[b], + blk_m[b])
// and row i of it is copied from A's row i (one wave per row)
template <typename S, typename TB>
__global__ void __launch_bounds__(256) k_bj_gather_dense(int n, const int *elem_blk, const int *blk_b0, const int *blk_m,
                                                         const cx<S> *A, const long long *boff, cx<TB> *blk)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int b = elem_blk[i], b0 = blk_b0[b], m = blk_m[b], a = i - b0;
    const cx<S> *src = A + (size_t)i * n + b0;
    cx<TB> *dst = blk + boff[b] + (size_t)a * m;
    for (int j = lane; j < m; j += 64) dst[j] = cx<TB>{TB(src[j].re), TB(src[j].im)};
}

// In-place Gauss-Jordan inversion with partial (row) pivoting of one block (m <= 64) by
// one work-group of 4 waves, the block in LDS, lane j owning column j: per pivot step the
// first wave finds the pivot, swaps and scales the pivot row, and after a barrier the four
// waves eliminate their rows (i = wave mod 4) with the same arithmetic per entry as one
// wave would.  The row interchanges are undone as column interchanges at the end, lane =
// row, and the inverse is written back column-major.  TB is the arithmetic and storage
// type (FP32 in the mixed form).
template <typename TB>
__global__ void __launch_bounds__(256) k_bj_invert(const int *blk_m, const long long *boff, cx<TB> *blk, int *err)
{
    __shared__ cx<TB> A[kBlk * kBlk];
    __shared__ int perm[kBlk];
    const int b = blockIdx.x, m = blk_m[b], lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cx<TB> *G = blk + boff[b];
    // row-major m x m from the gather
    if (lane < m)
        for (int i = wv; i < m; i += 4) A[i * kBlk + lane] = G[(size_t)i * m + lane];
    __syncthreads();
    bool singular = false;
    for (int k = 0; k < m; ++k) {
        if (wv == 0) {
            // pivot: the largest |A[i][k]|, i >= k (first on ties)
            TB v = TB(-1);
            int idx = lane;
            if (lane >= k && lane < m) {
                const cx<TB> a = A[lane * kBlk + k];
                v = a.re * a.re + a.im * a.im;
            }
            for (int o = 32; o > 0; o >>= 1) {
                const TB v2 = __shfl_xor(v, o);
                const int i2 = __shfl_xor(idx, o);
                if (v2 > v || (v2 == v && i2 < idx)) {
                    v = v2;
                    idx = i2;
                }
            }
            const int pr = __builtin_amdgcn_readfirstlane(idx);
            if (!(v > TB(0))) singular = true;
            if (lane == 0) perm[k] = pr;
            cx<TB> rk = mk(TB(0), TB(0));
            if (lane < m) {
                rk = A[pr * kBlk + lane];
                if (pr != k) A[pr * kBlk + lane] = A[k * kBlk + lane];
            }
            const cx<TB> pv = {__shfl(rk.re, k), __shfl(rk.im, k)};   // the pivot, from lane k's column
            const cx<TB> inv = cdiv(mk(TB(1), TB(0)), pv);
            rk = lane == k ? inv : rk * inv;
            if (lane < m) A[k * kBlk + lane] = rk;
        }
        __syncthreads();
        // every other row i: A[i][j] -= A[i][k] * rk_j (column k: 0 - A[i][k] * inv).  Wave w
        // takes the rows i = w mod 4, sixteen per batch, their column-k entries read by
        // every lane before any is written.
        const cx<TB> rk = A[k * kBlk + min(lane, m - 1)];
        constexpr int RB = 16;
        for (int i0 = wv; i0 < m; i0 += 4 * RB) {
            cx<TB> f[RB], base[RB];
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int i = min(i0 + 4 * q, m - 1);
                f[q] = A[i * kBlk + k];
                base[q] = A[i * kBlk + min(lane, m - 1)];
            }
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const int i = i0 + 4 * q;
                if (i < m && i != k && lane < m)
                    A[i * kBlk + lane] = (lane == k ? mk(TB(0), TB(0)) : base[q]) - f[q] * rk;
            }
        }
        __syncthreads();
    }
    if (wv == 0) {
        if (singular && lane == 0) atomicOr(err, 16);
        // undo the row interchanges as column interchanges, lane = row
        for (int k = m - 1; k >= 0; --k) {
            const int pk = perm[k];
            if (pk != k && lane < m) {
                const cx<TB> t0 = A[lane * kBlk + k];
                A[lane * kBlk + k] = A[lane * kBlk + pk];
                A[lane * kBlk + pk] = t0;
            }
        }
    }
    __syncthreads();
    // column-major: G[j * m + i] = inverse[i][j], lane = i
    if (lane < m)
        for (int j = wv; j < m; j += 4) G[(size_t)j * m + lane] = A[lane * kBlk + j];
}

// y = M^-1 x: one work-group per block; lane a is row a, the 4 waves split the columns
// (four per step) of the column-major inverse and meet in LDS
template <typename TB, int NR>
__device__ __forceinline__ void bj_apply_item(int b, int n, int nr, const int *blk_b0, const int *blk_m,
                                              const int *cl_elem, const long long *boff, const cx<TB> *minv,
                                              const cx<double> *x, cx<double> *y)
{
    // nr <= NR load cases are present; the others are zero and not stored
    __shared__ cx<double> xs[NR][kBlk];
    __shared__ double red[3][kBlk][2 * NR];
    const int b0 = blk_b0[b], m = blk_m[b];
    const int a = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int t = threadIdx.x; t < NR * m; t += 256) {
        const int r = t / m, j = t - r * m;
        xs[r][j] = r < nr ? x[(size_t)r * n + cl_elem[b0 + j]] : mk(0.0, 0.0);
    }
    __syncthreads();
    const cx<TB> *M = minv + boff[b];
    double re[NR], im[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) re[r] = im[r] = 0.0;
    if (a < m) {
        int j = wv;
        for (; j + 12 < m; j += 16) {
            cx<TB> v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = M[(size_t)(j + 4 * q) * m + a];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < NR; ++r) cmac(re[r], im[r], double(v[q].re), double(v[q].im), xs[r][j + 4 * q]);
        }
        for (; j < m; j += 4) {
            const cx<TB> v = M[(size_t)j * m + a];
#pragma unroll
            for (int r = 0; r < NR; ++r) cmac(re[r], im[r], double(v.re), double(v.im), xs[r][j]);
        }
    }
    if (wv > 0)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            red[wv - 1][a][2 * r] = re[r];
            red[wv - 1][a][2 * r + 1] = im[r];
        }
    __syncthreads();
    if (wv == 0 && a < m) {
        const int e = cl_elem[b0 + a];
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r < nr)
                y[(size_t)r * n + e] = mk(re[r] + red[0][a][2 * r] + red[1][a][2 * r] + red[2][a][2 * r],
                                          im[r] + red[0][a][2 * r + 1] + red[1][a][2 * r + 1] + red[2][a][2 * r + 1]);
    }
}

template <typename TB, int NR>
__global__ void __launch_bounds__(256) k_bj_apply(int n, int nr, const int *blk_b0, const int *blk_m,
                                                  const int *cl_elem, const long long *boff, const cx<TB> *minv,
                                                  const cx<double> *x, cx<double> *y)
{
    bj_apply_item<TB, NR>(blockIdx.x, n, nr, blk_b0, blk_m, cl_elem, boff, minv, x, y);
}

// v_r = w_r / sqrt(re <w_r, w_r>) with the norm read from a device dot result [r][3]
__global__ void k_scale_norm(int n, int nrhs, const cx<double> *dts, const cx<double> *w, cx<double> *v)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nrhs) return;
    const int r = (int)(t / n);
    const double s = sqrt(dts[3 * r].re);
    v[t] = s > 1e-300 ? w[t] * (1.0 / s) : mk(0.0, 0.0);
}

// ---- FMM operator apply ---------------------------------------------------------------
// Each kernel handles NR (1..4) load cases per pass so the operator is read once for
// all of them; vectors are [r][...] with the given strides.

// The T and D applies sum over a cluster's elements or interaction list per sphere
// point; they are latency-bound at these sizes (a few tens of MB per launch), so their
// work-groups have 8 waves that split the sum and meet in LDS.
constexpr int kSumWaves = 8;

// sum over the NW waves of a work-group of each lane's re/im, by halves in LDS (red holds
// [NW/2][64][2 NR] doubles); the result is left in wave 0
template <int NR, int NW>
__device__ __forceinline__ void waves_sum(double (&re)[NR], double (&im)[NR], double *red, int wv, int lane)
{
#pragma unroll
    for (int h = NW / 2; h >= 1; h >>= 1) {
        if (wv >= h && wv < 2 * h)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                red[((wv - h) * 64 + lane) * 2 * NR + 2 * r] = re[r];
                red[((wv - h) * 64 + lane) * 2 * NR + 2 * r + 1] = im[r];
            }
        __syncthreads();
        if (wv < h)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                re[r] += red[(wv * 64 + lane) * 2 * NR + 2 * r];
                im[r] += red[(wv * 64 + lane) * 2 * NR + 2 * r + 1];
            }
        __syncthreads();
    }
}

// Row chunks: the near-field apply below gives each work-group a chunk of at most
// ncgk::kNearRows rows that share one cluster (ncgk::NearChunks: the rows, their value
// offsets and the shared column list).  Each of the 4 waves owns rows a = wave + 4 j and
// streams their values with 2 kNearRowsPerWave independent loads per lane per step; the
// vector the rows multiply is staged in LDS once per work-group.
static_assert(ncgk::kNearRows == 4 * ncgk::kNearRowsPerWave, "4 waves per chunk");

// sum over the wave of re/im for RW rows and NR load cases, then row a gets lane 0's value
template <int RW, int NR>
__device__ __forceinline__ void wave_sum_rows(double (&re)[RW][NR], double (&im)[RW][NR])
{
#pragma unroll
    for (int j = 0; j < RW; ++j)
#pragma unroll
        for (int r = 0; r < NR; ++r)
            for (int o = 32; o > 0; o >>= 1) {
                re[j][r] += __shfl_xor(re[j][r], o);
                im[j][r] += __shfl_xor(im[j][r], o);
            }
}

// values of RW rows for one batch of U steps of 64 columns from b + lane (indices past sl
// clamped to sl - 1)
template <typename S, int RW, int U>
__device__ __forceinline__ void rows_load(const NCG_GLOBAL cx<S> *const (&vr)[RW], int b, int sl, int lane,
                                          cx<S> (&v)[U][RW])
{
#pragma unroll
    for (int q = 0; q < U; ++q) {
        const int u = min(b + 64 * q + lane, sl - 1);
#pragma unroll
        for (int j = 0; j < RW; ++j) v[q][j] = vr[j][u];
    }
}

// re/im[j][r] += this lane's share (u = lane + 64 t < sl) of sum_u v_j[u] * xs[r][u] for
// RW rows v_j and NR vectors in LDS (row stride seg).  v holds the first batch on entry
// (rows_load at b = 0); each batch's successor is loaded before the batch is used, so
// two batches of loads are in flight.  Indices past sl are clamped and weighted by zero.
template <typename S, int NR, int RW, int U>
__device__ __forceinline__ void rows_dot(const NCG_GLOBAL cx<S> *const (&vr)[RW], int sl, const cx<double> *xs, int seg,
                                         int lane,
                                         cx<S> (&v)[U][RW], double (&re)[RW][NR], double (&im)[RW][NR])
{
    for (int b = 0; b < sl; b += 64 * U) {   // wave-uniform trip count
        cx<S> nv[U][RW];
        const bool more = b + 64 * U < sl;
        if (more) rows_load<S, RW, U>(vr, b + 64 * U, sl, lane, nv);
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int u = b + 64 * q + lane;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                cx<double> xv = xs[r * seg + min(u, sl - 1)];
                if (u >= sl) xv = mk(0.0, 0.0);
#pragma unroll
                for (int j = 0; j < RW; ++j) cmac(re[j][r], im[j][r], double(v[q][j].re), double(v[q][j].im), xv);
            }
        }
        if (more)
#pragma unroll
            for (int q = 0; q < U; ++q)
#pragma unroll
                for (int j = 0; j < RW; ++j) v[q][j] = nv[q][j];
    }
}

// near field: all rows of a finest cluster hold the same columns (the elements of its
// near clusters), so a chunk gathers x over that column list into LDS (in segments of
// seg columns) and the values are the only per-nonzero traffic.  The chunk's rows, value
// offsets and column list come from one table, so the first loads of a work-group do not
// depend on each other.
template <typename S, int NR>
__device__ __forceinline__ void near_spmv_item(int q, int n, int seg, const NCG_GLOBAL long long *ch_cb,
                                               const NCG_GLOBAL int *ch_w, const NCG_GLOBAL int *ch_row,
                                               const NCG_GLOBAL long long *ch_off, const NCG_GLOBAL int *col,
                                               const NCG_GLOBAL cx<S> *val, const NCG_GLOBAL cx<double> *x,
                                               NCG_GLOBAL cx<double> *y)
{
    extern __shared__ double xs_sh[];   // [NR][seg] complex
    cx<double> *xs = (cx<double> *)xs_sh;
    // the wave index is uniform: readfirstlane lets per-wave indices and loads go scalar
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long cb = ch_cb[q];
    const int w = ch_w[q];
    constexpr int RW = ncgk::kNearRowsPerWave;
    int row[RW];
    const NCG_GLOBAL cx<S> *vr[RW];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int a = ncgk::kNearRows * q + wv + 4 * j;
        row[j] = ch_row[a];
        vr[j] = val + ch_off[a];
    }
    double re[RW][NR], im[RW][NR];
#pragma unroll
    for (int j = 0; j < RW; ++j)
#pragma unroll
        for (int r = 0; r < NR; ++r) re[j][r] = im[j][r] = 0.0;
    for (int s0 = 0; s0 < w; s0 += seg) {
        const int sl = min(seg, w - s0);
        // unused row slots read the first row of the cluster and are not stored; the first
        // batch of values is requested before the gather so the two latencies overlap
        const NCG_GLOBAL cx<S> *vs[RW];
#pragma unroll
        for (int j = 0; j < RW; ++j) vs[j] = vr[j] + s0;
        cx<S> v[2][RW];
        rows_load<S, RW, 2>(vs, 0, sl, lane, v);
        __syncthreads();
        for (int u = threadIdx.x; u < sl; u += 256) {
            const int e = col[cb + s0 + u];
#pragma unroll
            for (int r = 0; r < NR; ++r) xs[r * seg + u] = x[(size_t)r * n + e];
        }
        __syncthreads();
        rows_dot<S, NR, RW, 2>(vs, sl, xs, seg, lane, v, re, im);
    }
    wave_sum_rows<RW, NR>(re, im);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < RW; ++j)
            if (row[j] >= 0)
#pragma unroll
                for (int r = 0; r < NR; ++r) y[(size_t)r * n + row[j]] = mk(re[j][r], im[j][r]);
}

template <typename S, int NR>
__global__ void __launch_bounds__(256) k_near_spmv(int n, int seg, const long long *ch_cb, const int *ch_w,
                                                   const int *ch_row, const long long *ch_off, const int *col,
                                                   const cx<S> *val, const cx<double> *x, cx<double> *y)
{
    near_spmv_item<S, NR>(blockIdx.x, n, seg, to_global(ch_cb), to_global(ch_w), to_global(ch_row), to_global(ch_off),
                          to_global(col), to_global(val), to_global(x), to_global(y));
}

// t[c, p] = sum over the cluster's elements of T * x.  One work-group per (cluster, tile
// of 64 sphere points); its waves take every kSumWaves-th element of the cluster, four
// per step, and their partial sums meet in LDS.  W is the storage type of t.
template <typename S, typename W, int NR, int NW = kSumWaves, int PL = 1>
__device__ __forceinline__ void apply_T_item(int blk, int n, int ncl, int P, const NCG_GLOBAL int *cl_ptr,
                                             const NCG_GLOBAL int *cl_elem, const NCG_GLOBAL cx<S> *Tv,
                                             const NCG_GLOBAL cx<double> *x, NCG_GLOBAL cx<W> *tw)
{
    // PL points per lane (p, p + 1 with PL = 2: P is even), each summed as with one
    __shared__ double red[(NW > 1 ? NW / 2 : 1) * 64 * 2 * NR * PL];
    const int tiles = (P + 64 * PL - 1) / (64 * PL);
    const int c = blk / tiles, tile = blk % tiles;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = (tile * 64 + lane) * PL;
    const bool on = p < P;
    double re[NR * PL], im[NR * PL];   // [r * PL + h]
#pragma unroll
    for (int r = 0; r < NR * PL; ++r) re[r] = im[r] = 0.0;
    const int k1 = cl_ptr[c + 1];
    int k = cl_ptr[c] + wv;
    if (on) {
        for (; k + 3 * NW < k1; k += 4 * NW) {
            cx<S> a[4][PL];
            int e[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                load_pts<S, PL>(Tv + (size_t)(k + NW * j) * P + p, a[j]);
                e[j] = cl_elem[k + NW * j];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const cx<double> xv = x[(size_t)r * n + e[j]];
#pragma unroll
                    for (int h = 0; h < PL; ++h)
                        cmac(re[r * PL + h], im[r * PL + h], double(a[j][h].re), double(a[j][h].im), xv);
                }
        }
        for (; k < k1; k += NW) {
            cx<S> a[PL];
            load_pts<S, PL>(Tv + (size_t)k * P + p, a);
            const int e = cl_elem[k];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const cx<double> xv = x[(size_t)r * n + e];
#pragma unroll
                for (int h = 0; h < PL; ++h) cmac(re[r * PL + h], im[r * PL + h], double(a[h].re), double(a[h].im), xv);
            }
        }
    }
    waves_sum<NR * PL, NW>(re, im, red, wv, lane);
    if (wv == 0 && on)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            cx<W> o[PL];
#pragma unroll
            for (int h = 0; h < PL; ++h) o[h] = cx<W>{W(re[r * PL + h]), W(im[r * PL + h])};
            store_pts<W, PL>(tw + (size_t)r * ncl * P + (size_t)c * P + p, o);
        }
}

template <typename S, typename W, int NR>
__global__ void __launch_bounds__(64 * kSumWaves) k_fmm_apply_T(int n, int ncl, int P, const int *cl_ptr,
                                                                const int *cl_elem, const cx<S> *Tv, const cx<double> *x,
                                                                cx<W> *tw)
{
    apply_T_item<S, W, NR>(blockIdx.x, n, ncl, P, to_global(cl_ptr), to_global(cl_elem), to_global(Tv), to_global(x),
                           to_global(tw));
}

// s[i, p] = sum over the interaction list of D * t (D is diagonal in p); one work-group
// per (cluster, tile of 64 points), the waves splitting the interaction list four
// entries per step.  WI / WO are the storage types of t and s.
template <typename S, typename WI, typename WO, int NR, int NW = kSumWaves, int PL = 1>
__device__ __forceinline__ void apply_D_item(int blk, int ncl, int P, const NCG_GLOBAL int *fan_ptr,
                                             const NCG_GLOBAL int *fan_idx, const NCG_GLOBAL cx<S> *Dv,
                                             const NCG_GLOBAL cx<WI> *tw, NCG_GLOBAL cx<WO> *sw)
{
    // PL points per lane (p, p + 1 with PL = 2: P is even), each summed as with one
    __shared__ double red[(NW > 1 ? NW / 2 : 1) * 64 * 2 * NR * PL];
    const int tiles = (P + 64 * PL - 1) / (64 * PL);
    const int i = blk / tiles, tile = blk % tiles;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int p = (tile * 64 + lane) * PL;
    const bool on = p < P;
    const size_t ld = (size_t)ncl * P;
    double re[NR * PL], im[NR * PL];   // [r * PL + h]
#pragma unroll
    for (int r = 0; r < NR * PL; ++r) re[r] = im[r] = 0.0;
    const int q1 = fan_ptr[i + 1];
    int q = fan_ptr[i] + wv;
    if (on) {
        for (; q + 3 * NW < q1; q += 4 * NW) {
            cx<S> a[4][PL];
            int jj[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                load_pts<S, PL>(Dv + (size_t)(q + NW * j) * P + p, a[j]);
                jj[j] = fan_idx[q + NW * j];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    cx<WI> t[PL];
                    load_pts<WI, PL>(tw + r * ld + (size_t)jj[j] * P + p, t);
#pragma unroll
                    for (int h = 0; h < PL; ++h)
                        cmac(re[r * PL + h], im[r * PL + h], double(a[j][h].re), double(a[j][h].im),
                             mk(double(t[h].re), double(t[h].im)));
                }
        }
        for (; q < q1; q += NW) {
            cx<S> a[PL];
            load_pts<S, PL>(Dv + (size_t)q * P + p, a);
            const int j = fan_idx[q];
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                cx<WI> t[PL];
                load_pts<WI, PL>(tw + r * ld + (size_t)j * P + p, t);
#pragma unroll
                for (int h = 0; h < PL; ++h)
                    cmac(re[r * PL + h], im[r * PL + h], double(a[h].re), double(a[h].im),
                         mk(double(t[h].re), double(t[h].im)));
            }
        }
    }
    waves_sum<NR * PL, NW>(re, im, red, wv, lane);
    if (wv == 0 && on)
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            cx<WO> o[PL];
#pragma unroll
            for (int h = 0; h < PL; ++h) o[h] = cx<WO>{WO(re[r * PL + h]), WO(im[r * PL + h])};
            store_pts<WO, PL>(sw + r * ld + (size_t)i * P + p, o);
        }
}

template <typename S, typename WI, typename WO, int NR>
__global__ void __launch_bounds__(64 * kSumWaves) k_fmm_apply_D(int ncl, int P, const int *fan_ptr, const int *fan_idx,
                                                                const cx<S> *Dv, const cx<WI> *tw, cx<WO> *sw)
{
    apply_D_item<S, WI, WO, NR>(blockIdx.x, ncl, P, to_global(fan_ptr), to_global(fan_idx), to_global(Dv),
                                to_global(tw), to_global(sw));
}

// y[e] += sum_p S[e, p] s[cluster(e), p].  One work-group of 8 waves per tile of at most
// 64 elements of one cluster (tl_*: cluster, first local element, elements): lane a is
// element a0 + a of the cluster's column-major S block, so every load is a contiguous run
// of the block; the waves split the points, four per step, and the partial sums meet in an
// LDS tree.  s[c, p] is wave-uniform and read from LDS, where the work-group stages it
// kSStage points at a time (read through the scalar cache instead, a launch took 1.4x as
// long).
constexpr int kSStage = 256;
template <typename S, typename W, int NR, int NW = kSumWaves>
__device__ __forceinline__ void apply_S_item(int q, int n, int ncl, int P, const NCG_GLOBAL int *tl_cl,
                                             const NCG_GLOBAL int *tl_a0, const NCG_GLOBAL int *tl_m,
                                             const NCG_GLOBAL int *cl_ptr, const NCG_GLOBAL int *cl_elem,
                                             const NCG_GLOBAL cx<S> *Sv, const NCG_GLOBAL cx<W> *sw,
                                             NCG_GLOBAL cx<double> *y)
{
    static_assert(kSStage % NW == 0, "a stage keeps each point on its wave");
    __shared__ double red[(NW > 1 ? NW / 2 : 1) * 64 * 2 * NR];
    __shared__ cx<double> ss[NR][kSStage];
    const int c = tl_cl[q], a0 = tl_a0[q], m = tl_m[q];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = cl_ptr[c], mc = cl_ptr[c + 1] - b;
    const bool on = lane < m;
    const NCG_GLOBAL cx<S> *blk = Sv + (size_t)b * P + a0 + (on ? lane : 0);
    const size_t ld = (size_t)ncl * P;
    const NCG_GLOBAL cx<W> *sc = sw + (size_t)c * P;
    double re[NR], im[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) re[r] = im[r] = 0.0;
    for (int p0 = 0; p0 < P; p0 += kSStage) {
        const int pn = min(P, p0 + kSStage);
        if (p0 > 0) __syncthreads();   // the previous stage has been read
        for (int u = threadIdx.x; u < pn - p0; u += 64 * NW)
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const cx<W> t = sc[r * ld + p0 + u];
                ss[r][u] = mk(double(t.re), double(t.im));
            }
        __syncthreads();
        if (on) {
            int p = p0 + wv;
            for (; p + 3 * NW < pn; p += 4 * NW) {
                cx<S> a[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = blk[(size_t)(p + NW * j) * mc];
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int r = 0; r < NR; ++r)
                        cmac(re[r], im[r], double(a[j].re), double(a[j].im), ss[r][p + NW * j - p0]);
            }
            for (; p < pn; p += NW) {
                const cx<S> a = blk[(size_t)p * mc];
#pragma unroll
                for (int r = 0; r < NR; ++r) cmac(re[r], im[r], double(a.re), double(a.im), ss[r][p - p0]);
            }
        }
    }
    waves_sum<NR, NW>(re, im, red, wv, lane);
    if (wv == 0 && on) {
        const int e = cl_elem[b + a0 + lane];
#pragma unroll
        for (int r = 0; r < NR; ++r) y[(size_t)r * n + e] += mk(re[r], im[r]);
    }
}

template <typename S, typename W, int NR>
__global__ void __launch_bounds__(64 * kSumWaves) k_fmm_apply_S(int n, int ncl, int P, const int *tl_cl,
                                                                const int *tl_a0, const int *tl_m, const int *cl_ptr,
                                                                const int *cl_elem, const cx<S> *Sv, const cx<W> *sw,
                                                                cx<double> *y)
{
    apply_S_item<S, W, NR>(blockIdx.x, n, ncl, P, to_global(tl_cl), to_global(tl_a0), to_global(tl_m), to_global(cl_ptr),
                           to_global(cl_elem), to_global(Sv), to_global(sw), to_global(y));
}

// ---- FMM evaluation grid (NC_ContributionFMM, src/NC_PostProcessing.cpp:978-1214) ------
// far-field signature of the solution at level 0 (NC_ContributionTvecFMM,
// src/NC_PostProcessing.cpp:302-364): thread per (cluster, sphere point)
// T is the arithmetic of the phase and weight (FP32 in the mixed form, after the
// geometry differences are formed in FP64); sums are FP64.
template <typename T, typename TV>
__global__ void __launch_bounds__(256) k_fmm_Tvec_sol(Elements E, BMConst bm, int P, int ncl, const double *ccen,
                                                      const int *cl_ptr, const int *cl_elem, const double *u,
                                                      const cx<double> *pot, const cx<double> *vel, int nrhs,
                                                      cx<TV> *tv)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncl * P) return;
    const int c = (int)(t / P), p = (int)(t % P);
    const double C[3] = {ccen[3 * c], ccen[3 * c + 1], ccen[3 * c + 2]};
    const T su[3] = {T(u[3 * p]), T(u[3 * p + 1]), T(u[3 * p + 2])};
    const double h = bm.hk / bm.k;
    const T hk = T(bm.hk), kk = T(bm.k);
    for (int r = 0; r < nrhs; ++r) {
        cx<double> acc = mk(0.0, 0.0);
        for (int k = cl_ptr[c]; k < cl_ptr[c + 1]; ++k) {
            const int e = cl_elem[k];
            LocalElem L;
            load_local(E, e, C, L);
            const cx<double> pe = pot[(size_t)r * E.n + e], ve = vel[(size_t)r * E.n + e];
            for (int g = 0; g < 4; ++g) {
                double s, tt, w;
                fmm_gauss(L.nv, g, s, tt, w);
                double q[3], ny[3], J;
                point_on(L, s, tt, q, ny, J);
                const double wga = w * J;
                const T dw = -(T(q[0]) * su[0] + T(q[1]) * su[1] + T(q[2]) * su[2]) * hk;
                T sn, cs;
                sincos_t(dw, &sn, &cs);
                const cx<double> zg = mk(double(cs), double(sn));
                const double dn = double((T(ny[0]) * su[0] + T(ny[1]) * su[1] + T(ny[2]) * su[2]) * kk);
                const cx<double> zH = zg * (dn * wga);
                const cx<double> zG = mk(-zg.im * wga * h, zg.re * wga * h);
                acc += zH * pe - zG * ve;
            }
        }
        acc = acc * (bm.k / kPi4);
        if (bm.tau != 1.0) acc = acc * bm.tau;
        tv[(size_t)r * ncl * P + t] = cx<TV>{TV(acc.re), TV(acc.im)};
    }
}

// translation of the level-0 signatures to the evaluation clusters:
// Z[ipc, p] = sum over far BE clusters of mu(C_ipc - C_ic, s_p) T[ic, p] w_p.
// First pass: one thread per far pair computes the direction and the spherical Hankel
// functions h_n(k |C_ipc - C_ic|), n < L, into [pair][4 + 2 L] doubles.
__global__ void __launch_bounds__(128) k_fmm_pair_hankel(int npair, int L, double k, double h, const double *ecen,
                                                         const double *ccen, const int *pair_ev, const int *pair_cl,
                                                         double *tab)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npair) return;
    const int a = pair_ev[q], c = pair_cl[q];
    double d[3] = {ecen[3 * a] - ccen[3 * c], ecen[3 * a + 1] - ccen[3 * c + 1], ecen[3 * a + 2] - ccen[3 * c + 2]};
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double *o = tab + (size_t)q * (4 + 2 * L);
    o[0] = d[0] / r;
    o[1] = d[1] / r;
    o[2] = d[2] / r;
    o[3] = 0.0;
    hankel_series(L, r * k, h, o + 4);
}

// Second pass: one block per evaluation cluster, threads over sphere points, summing
// its far pairs in registers (pairs of one evaluation cluster are contiguous and are
// staged through LDS in batches).  T is the arithmetic of the series (FP32 in the mixed
// form where the evaluation pairs allow it), TV the storage of the signature tv (FP32
// there too: the kernel is bound by the tv reads, NR per pair and point); the sums over
// pairs are FP64.
template <int NR, typename T, typename TV>
__global__ void __launch_bounds__(256) k_fmm_eval_z(int nev, int L, int P, double h, const int *ev_ptr,
                                                    const int *pair_cl, const double *tab, const double *u,
                                                    const double *wsph, const cx<TV> *tv, int ncl,
                                                    cx<double> *z)
{
    constexpr int QB = 16;
    __shared__ T sh[QB][4 + 2 * kMaxTerms];
    __shared__ T ca[kMaxTerms], cb[kMaxTerms];
    __shared__ int scl[QB];
    const int a = blockIdx.x, w = 4 + 2 * L;
    const size_t ldt = (size_t)ncl * P, ldz = (size_t)nev * P;
    // Legendre recurrence P_n = ((2n-1) c P_{n-1} - (n-1) P_{n-2}) / n as P_n = ca c P_{n-1} - cb P_{n-2}
    for (int t = threadIdx.x; t < L; t += blockDim.x) {
        ca[t] = t >= 2 ? T((double)(2 * t - 1) / (double)t) : T(0);
        cb[t] = t >= 2 ? T((double)(t - 1) / (double)t) : T(0);
    }
    for (int p0 = 0; p0 < P; p0 += blockDim.x) {
        const int p = p0 + threadIdx.x;
        const bool on = p < P;
        T su[3] = {T(0), T(0), T(0)};
        T wp = T(0);
        if (on) { su[0] = T(u[3 * p]); su[1] = T(u[3 * p + 1]); su[2] = T(u[3 * p + 2]); wp = T(wsph[p]); }
        cx<double> acc[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] = mk(0.0, 0.0);
        const int qend = ev_ptr[a + 1];
        for (int q0 = ev_ptr[a]; q0 < qend; q0 += QB) {
            const int nb = min(QB, qend - q0);
            __syncthreads();
            for (int t = threadIdx.x; t < nb * w; t += blockDim.x) sh[t / w][t % w] = T(tab[(size_t)q0 * w + t]);
            for (int t = threadIdx.x; t < nb; t += blockDim.x) scl[t] = pair_cl[q0 + t];
            __syncthreads();
            if (!on) continue;
            for (int qq = 0; qq < nb; ++qq) {
                const T *s = sh[qq];
                // mu = sum_n a_n P_n(cc), a_n = (i h)^n (2n + 1) h_n from the table
                const T cc = s[0] * su[0] + s[1] * su[1] + s[2] * su[2];
                cx<T> mu = mk(s[4], s[5]);   // n = 0
                T pm1 = T(1), pm0 = cc;
                if (L > 1) {
                    mu.re += s[6] * cc;
                    mu.im += s[7] * cc;
                }
                for (int nn = 2; nn < L; ++nn) {
                    const T pn = ca[nn] * cc * pm0 - cb[nn] * pm1;
                    pm1 = pm0;
                    pm0 = pn;
                    mu.re += s[4 + 2 * nn] * pn;
                    mu.im += s[5 + 2 * nn] * pn;
                }
                const cx<double> md = mk(double(mu.re * wp), double(mu.im * wp));
                const int c = scl[qq];
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    const cx<TV> t = tv[r * ldt + (size_t)c * P + p];
                    acc[r] += md * mk(double(t.re), double(t.im));
                }
            }
        }
        if (on)
#pragma unroll
            for (int r = 0; r < NR; ++r) z[r * ldz + (size_t)a * P + p] = acc[r];
    }
}

// far part at each evaluation point: phi = sum_p e^{i h k s_p.(x - C)} Z[p], grad likewise
__global__ void __launch_bounds__(256) k_fmm_eval_far(int m, int P, int nev, double hk, const double *pts,
                                                      const int *pt_ev, const double *ecen, const double *u,
                                                      int nrhs, const cx<double> *z, cx<double> *fphi,
                                                      cx<double> *fgrad)
{
    const int lane = threadIdx.x & 63;
    const int x = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (x >= m) return;
    const int a = pt_ev[x];
    const double d[3] = {pts[3 * x] - ecen[3 * a], pts[3 * x + 1] - ecen[3 * a + 1], pts[3 * x + 2] - ecen[3 * a + 2]};
    for (int r = 0; r < nrhs; ++r) {
        double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int p = lane; p < P; p += 64) {
            const double sp = (d[0] * u[3 * p] + d[1] * u[3 * p + 1] + d[2] * u[3 * p + 2]) * hk;
            double sn, cs;
            sincos(sp, &sn, &cs);
            const cx<double> zi = mk(cs, sn) * z[(size_t)r * nev * P + (size_t)a * P + p];
            v[0] += zi.re;
            v[1] += zi.im;
            for (int kk = 0; kk < 3; ++kk) {
                const double dk = hk * u[3 * p + kk];
                v[2 + 2 * kk] += -zi.im * dk;
                v[3 + 2 * kk] += zi.re * dk;
            }
        }
        for (int c = 0; c < 8; ++c)
            for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o);
        if (lane == 0) {
            fphi[(size_t)r * m + x] = mk(v[0], v[1]);
            for (int kk = 0; kk < 3; ++kk) fgrad[((size_t)r * m + x) * 3 + kk] = mk(v[2 + 2 * kk], v[3 + 2 * kk]);
        }
    }
}

// ======================================================================================
// Grouped launches (ncgk::GNear, GLevel, GBj, GVec, GGm): one launch covers the same
// operation on every problem of a group solved in lockstep.  Work-group b belongs to the
// item d with start[d] <= b < start[d + 1] (scaled by `scale` where an item's count is a
// multiple of a launch-wide factor); it runs the single-problem kernel's work-group
// b - start[d] on that item's arrays, or nothing while the item's skip flag is set.
// ======================================================================================
__device__ __forceinline__ int g_item(const int *start, int nitems, int b, int scale = 1)
{
    int d = 0;
    while (d + 1 < nitems && b >= scale * start[d + 1]) ++d;   // wave-uniform
    return d;
}

template <typename S, int NR>
__global__ void __launch_bounds__(256) kg_near_spmv(const ncgk::GNear *items, const int *start, int nitems, int seg)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GNear it = items[d];
    if (it.skip && *it.skip) return;
    near_spmv_item<S, NR>(blockIdx.x - start[d], it.n, seg, it.cb, it.w, it.row,
                          it.off, it.col, to_global((const cx<S> *)it.val),
                          to_global((const cx<double> *)it.x), to_global((cx<double> *)it.y));
}

// Near field with one work-group per finest cluster: when the cluster's column list fits one
// LDS segment, x over it is gathered once and the cluster's row chunks are streamed one after
// the other, the next chunk's first values requested before the current chunk's rows are
// reduced and stored; each row sums as in near_spmv_item (same lanes, same order).  Wider
// clusters run near_spmv_item chunk by chunk.
template <typename S, int NR>
__device__ __forceinline__ void near_cluster_item(int c, int n, int seg, const NCG_GLOBAL int *cptr,
                                                  const NCG_GLOBAL long long *ch_cb, const NCG_GLOBAL int *ch_w,
                                                  const NCG_GLOBAL int *ch_row, const NCG_GLOBAL long long *ch_off,
                                                  const NCG_GLOBAL int *col, const NCG_GLOBAL cx<S> *val,
                                                  const NCG_GLOBAL cx<double> *x, NCG_GLOBAL cx<double> *y)
{
    extern __shared__ double xs_sh[];   // [NR][seg] complex
    cx<double> *xs = (cx<double> *)xs_sh;
    const int q0 = cptr[c], q1 = cptr[c + 1];
    const int w = ch_w[q0];
    if (w > seg) {
        for (int q = q0; q < q1; ++q) near_spmv_item<S, NR>(q, n, seg, ch_cb, ch_w, ch_row, ch_off, col, val, x, y);
        return;
    }
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long cb = ch_cb[q0];
    constexpr int RW = ncgk::kNearRowsPerWave;
    const NCG_GLOBAL cx<S> *vr[RW];
    int row[RW];
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int a = ncgk::kNearRows * q0 + wv + 4 * j;
        row[j] = ch_row[a];
        vr[j] = val + ch_off[a];
    }
    cx<S> v[2][RW];
    rows_load<S, RW, 2>(vr, 0, w, lane, v);
    for (int u = threadIdx.x; u < w; u += 256) {
        const int e = col[cb + u];
#pragma unroll
        for (int r = 0; r < NR; ++r) xs[r * seg + u] = x[(size_t)r * n + e];
    }
    __syncthreads();
    for (int q = q0; q < q1; ++q) {
        double re[RW][NR], im[RW][NR];
#pragma unroll
        for (int j = 0; j < RW; ++j)
#pragma unroll
            for (int r = 0; r < NR; ++r) re[j][r] = im[j][r] = 0.0;
        rows_dot<S, NR, RW, 2>(vr, w, xs, seg, lane, v, re, im);
        int rq[RW];
#pragma unroll
        for (int j = 0; j < RW; ++j) rq[j] = row[j];
        if (q + 1 < q1) {
#pragma unroll
            for (int j = 0; j < RW; ++j) {
                const int a = ncgk::kNearRows * (q + 1) + wv + 4 * j;
                row[j] = ch_row[a];
                vr[j] = val + ch_off[a];
            }
            rows_load<S, RW, 2>(vr, 0, w, lane, v);
        }
        wave_sum_rows<RW, NR>(re, im);
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < RW; ++j)
                if (rq[j] >= 0)
#pragma unroll
                    for (int r = 0; r < NR; ++r) y[(size_t)r * n + rq[j]] = mk(re[j][r], im[j][r]);
    }
}

template <typename S, int NR>
__global__ void __launch_bounds__(256) kg_near_cluster(const ncgk::GNear *items, const int *start, int nitems,
                                                       int seg)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GNear it = items[d];
    if (it.skip && *it.skip) return;
    near_cluster_item<S, NR>(blockIdx.x - start[d], it.n, seg, it.cptr, it.cb, it.w,
                             it.row, it.off, it.col, to_global((const cx<S> *)it.val),
                             to_global((const cx<double> *)it.x), to_global((cx<double> *)it.y));
}

// Near field as a persistent launch: a grid sized to what the device holds at once walks
// the row chunks of every problem of the group, so no work-group waits for a launch slot
// and the next chunk's first values are in flight while the current chunk's rows are
// reduced and stored.  Two schedules: DYN, work queues, one per XCD (work-groups are
// dispatched to the 8 XCDs in turn, blockIdx.x % 8, and XCD x takes chunks from the x-th
// eighth of the list by an atomic counter queue[32 x]: device-scope atomics are served
// beside memory, so one counter per XCD spreads them over 8 lines); or a static stride
// (work-group b takes chunks b, b + grid, ...).  Every row sums as in near_spmv_item: the
// same lanes take the same columns in the same order, and the wave sum is the same.  With
// DYN the last work-group to leave returns the counters to zero, so the next launch on the
// stream starts from empty queues.
constexpr int kQueueXcds = 8, kQueueStride = 32, kQueueInts = kQueueStride * (kQueueXcds + 1);

template <typename S, int RW>
struct NearUnit {
    int q, w, n;
    long long cb;
    int row[RW];
    const NCG_GLOBAL cx<S> *vr[RW];
    const NCG_GLOBAL int *col;
    const NCG_GLOBAL cx<double> *x;
    NCG_GLOBAL cx<double> *y;
};

template <typename S, int RW>
__device__ __forceinline__ void near_unit(const ncgk::GNear *items, const int *start, int nitems, int b, int wv,
                                          NearUnit<S, RW> &u)
{
    const int d = g_item(start, nitems, b);
    const ncgk::GNear it = items[d];
    u.q = b - start[d];
    u.n = it.n;
    u.cb = it.cb[u.q];
    // a converged problem's chunks are empty units (no loads, no stores)
    u.w = (it.skip && *it.skip) ? 0 : it.w[u.q];
    u.col = it.col;
    u.x = to_global((const cx<double> *)it.x);
    u.y = to_global((cx<double> *)it.y);
#pragma unroll
    for (int j = 0; j < RW; ++j) {
        const int a = ncgk::kNearRows * u.q + wv + 4 * j;
        u.row[j] = it.row[a];
        u.vr[j] = to_global((const cx<S> *)it.val) + it.off[a];
    }
}

template <typename S, int NR, bool DYN>
__global__ void __launch_bounds__(256) kq_near_spmv(const ncgk::GNear *items, const int *start, int nitems, int seg,
                                                    int *queue)
{
    extern __shared__ double xs_sh[];   // [NR][seg] complex
    cx<double> *xs = (cx<double> *)xs_sh;
    __shared__ int s_next[2];
    constexpr int RW = ncgk::kNearRowsPerWave, U = 2;
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int total = start[nitems];
    const int xcd = blockIdx.x % kQueueXcds;
    const int lo = (int)((long long)total * xcd / kQueueXcds), hi = (int)((long long)total * (xcd + 1) / kQueueXcds);
    int *ctr = queue + kQueueStride * xcd;
    // the first unit; indices read back from LDS are uniform (readfirstlane keeps the unit's
    // descriptor, rows and value addresses in scalar registers)
    int b;
    if (DYN) {
        if (threadIdx.x == 0) s_next[0] = lo + atomicAdd(ctr, 1);
        __syncthreads();
        b = __builtin_amdgcn_readfirstlane(s_next[0]);
        if (b >= hi) b = total;
    } else {
        b = blockIdx.x;
    }
    int slot = 1;
    NearUnit<S, RW> cu;
    cx<S> v[U][RW];
    if (b < total) {
        near_unit<S, RW>(items, start, nitems, b, wv, cu);
        if (cu.w > 0) rows_load<S, RW, U>(cu.vr, 0, min(seg, cu.w), lane, v);
    }
    while (b < total) {   // uniform over the work-group
        // DYN: the next unit's index, published by the barrier after the gather
        if (DYN && threadIdx.x == 0) s_next[slot] = lo + atomicAdd(ctr, 1);
        double re[RW][NR], im[RW][NR];
#pragma unroll
        for (int j = 0; j < RW; ++j)
#pragma unroll
            for (int r = 0; r < NR; ++r) re[j][r] = im[j][r] = 0.0;
        for (int s0 = 0; s0 < cu.w; s0 += seg) {
            const int sl = min(seg, cu.w - s0);
            const NCG_GLOBAL cx<S> *vs[RW];
#pragma unroll
            for (int j = 0; j < RW; ++j) vs[j] = cu.vr[j] + s0;
            if (s0 > 0) {   // later segments (rows longer than seg): as near_spmv_item
                __syncthreads();
                rows_load<S, RW, U>(vs, 0, sl, lane, v);
            }
            for (int t = threadIdx.x; t < sl; t += 256) {
                const int e = cu.col[cu.cb + s0 + t];
#pragma unroll
                for (int r = 0; r < NR; ++r) xs[r * seg + t] = cu.x[(size_t)r * cu.n + e];
            }
            __syncthreads();
            rows_dot<S, NR, RW, U>(vs, sl, xs, seg, lane, v, re, im);
        }
        if (cu.w == 0) __syncthreads();   // publish s_next as the gather's barrier would
        int bn;
        if (DYN) {
            bn = __builtin_amdgcn_readfirstlane(s_next[slot]);
            if (bn >= hi) bn = total;
        } else {
            bn = b + gridDim.x;
        }
        // the next unit's first values are requested before this unit's rows are summed
        NearUnit<S, RW> nu;
        if (bn < total) {
            near_unit<S, RW>(items, start, nitems, bn, wv, nu);
            if (nu.w > 0) rows_load<S, RW, U>(nu.vr, 0, min(seg, nu.w), lane, v);
        }
        if (cu.w > 0) {
            wave_sum_rows<RW, NR>(re, im);
            if (lane == 0)
#pragma unroll
                for (int j = 0; j < RW; ++j)
                    if (cu.row[j] >= 0)
#pragma unroll
                        for (int r = 0; r < NR; ++r) cu.y[(size_t)r * cu.n + cu.row[j]] = mk(re[j][r], im[j][r]);
        }
        __syncthreads();   // xs and the other s_next slot are free again
        cu = nu;
        b = bn;
        slot ^= 1;
    }
    if (DYN && threadIdx.x == 0) {
        __threadfence();
        int *done = queue + kQueueStride * kQueueXcds;
        if (atomicAdd(done, 1) == (int)gridDim.x - 1) {
            for (int x = 0; x < kQueueXcds; ++x) atomicExch(queue + kQueueStride * x, 0);
            atomicExch(done, 0);
        }
    }
}

// grouped T and D: two sphere points per lane where the values are FP32 (one 16-byte load
// for both), one for FP64 values; work-groups per item = ncl * ceil(P / (64 * points))
template <typename S>
constexpr int kGroupedPoints = sizeof(S) == 4 ? 2 : 1;

template <typename S, typename W, int NR, int NW>
__global__ void __launch_bounds__(64 * NW) kg_apply_T(const ncgk::GLevel *items, const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GLevel it = items[d];
    if (it.skip && *it.skip) return;
    apply_T_item<S, W, NR, NW, kGroupedPoints<S>>(blockIdx.x - start[d], it.n, it.ncl, it.P, it.cl_ptr, it.cl_elem,
                               to_global((const cx<S> *)it.Tv), to_global((const cx<double> *)it.x),
                               to_global((cx<W> *)it.tw));
}

template <typename S, typename W, int NR, int NW>
__global__ void __launch_bounds__(64 * NW) kg_apply_D(const ncgk::GLevel *items, const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GLevel it = items[d];
    if (it.skip && *it.skip) return;
    apply_D_item<S, W, W, NR, NW, kGroupedPoints<S>>(blockIdx.x - start[d], it.ncl, it.P, it.fan_ptr, it.fan_idx,
                                  to_global((const cx<S> *)it.Dv), to_global((const cx<W> *)it.tw),
                                  to_global((cx<W> *)it.sw));
}

template <typename S, typename W, int NR, int NW>
__global__ void __launch_bounds__(64 * NW) kg_apply_S(const ncgk::GLevel *items, const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GLevel it = items[d];
    if (it.skip && *it.skip) return;
    apply_S_item<S, W, NR, NW>(blockIdx.x - start[d], it.n, it.ncl, it.P, it.tl_cl, it.tl_a0,
                               it.tl_m, it.cl_ptr, it.cl_elem,
                               to_global((const cx<S> *)it.Sv), to_global((const cx<W> *)it.sw),
                               to_global((cx<double> *)it.y));
}

template <typename TB, int NR>
__global__ void __launch_bounds__(256) kg_bj_apply(const ncgk::GBj *items, const int *start, int nitems,
                                                   const int *kp)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GBj it = items[d];
    if (it.skip && *it.skip) return;
    const double *x = (const double *)it.x + (kp ? (size_t)*kp * it.xk : 0);
    bj_apply_item<TB, NR>(blockIdx.x - start[d], it.n, it.nr, it.b0, it.m, it.elem, it.boff, (const cx<TB> *)it.minv,
                          (const cx<double> *)x, (cx<double> *)it.y);
}

__global__ void __launch_bounds__(256) kg_vec(int op, const ncgk::GVec *items, const int *start, int nitems,
                                              const int *kp)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GVec it = items[d];
    if (it.skip && *it.skip) return;
    const long long t = (long long)(blockIdx.x - start[d]) * 256 + threadIdx.x;
    if (t >= it.len) return;
    const cx<double> *a = (const cx<double> *)((const double *)it.a + (kp ? (size_t)*kp * it.ak : 0));
    const cx<double> *b = (const cx<double> *)it.b;
    cx<double> *y = (cx<double> *)it.y;
    switch (op) {
    case ncgk::kVecCopy: y[t] = a[t]; break;
    case ncgk::kVecSub: y[t] = a[t] - b[t]; break;
    case ncgk::kVecAdd: y[t] += a[t]; break;
    default: y[t] = mk(0.0, 0.0); break;
    }
}

// per item kMdLanes * nrhs * nsl work-groups (start holds the prefix of nrhs * nsl): work-group
// lane jj forms the partials of the basis vectors j = jj, jj + kMdLanes, ... below kcount =
// *kp + 1, each as one work-group of the per-vector form would
__global__ void __launch_bounds__(256) kg_multidot(const int *kp, const ncgk::GGm *items, const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x, ncgk::kMdLanes);
    const ncgk::GGm it = items[d];
    if (*it.skip) return;
    const int kcount = *kp + 1;
    const int b = blockIdx.x - ncgk::kMdLanes * start[d];
    const int sl = b % it.nsl, t = b / it.nsl, r = t % it.nrhs, jj = t / it.nrhs;
    for (int k = jj; k < kcount; k += ncgk::kMdLanes) {
        multidot_item(k, r, sl, it.nsl, it.n, it.nrhs, it.slice, (const cx<double> *)it.V, (const cx<double> *)it.W,
                      (cx<double> *)it.part);
        __syncthreads();   // the reduction's LDS before the next vector's partial sums
    }
}

// per item gsb * nrhs work-groups
__global__ void __launch_bounds__(256) kg_gs_update(int kcount, const int *kp, int pass, int m,
                                                    const ncgk::GGm *items, const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GGm it = items[d];
    if (*it.skip) return;
    const int kc = kp ? *kp + 1 : kcount;
    const int b = blockIdx.x - start[d];
    cx<double> *hout = (cx<double> *)it.h + (size_t)pass * it.nrhs * (m + 1);
    gs_update_item(b % it.gsb, b / it.gsb, it.gsb, it.n, it.nrhs, kc, it.nsl, (const cx<double> *)it.part,
                   (const cx<double> *)it.V, (cx<double> *)it.W, hout, pass ? it.npart : nullptr);
}

__global__ void __launch_bounds__(256) kg_gs_normalize(int kh, const int *kp, int m, const ncgk::GGm *items,
                                                       const int *start, int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GGm it = items[d];
    if (*it.skip) return;
    const int k = kp ? *kp : kh;
    const int b = blockIdx.x - start[d];
    cx<double> *v = (cx<double> *)it.V + (size_t)(k + 1) * it.nrhs * it.n;
    cx<double> *nout = (cx<double> *)it.h + (size_t)2 * it.nrhs * (m + 1);
    gs_normalize_item(b % it.gsb, b / it.gsb, it.gsb, it.n, it.gsb, it.npart, (const cx<double> *)it.W, v, nout);
}

// one wave per item, lane r = load case; the problem is done once every load case is
__device__ __forceinline__ void g_set_skip(const ncgk::GGm &it, bool done_r)
{
    const unsigned long long open = __ballot(threadIdx.x < (unsigned)it.nrhs && !done_r);
    if (threadIdx.x == 0) *it.skip = open == 0ull ? 1 : 0;
}

__global__ void __launch_bounds__(64) kg_gm_start(int m, int first, const ncgk::GGm *items, int *kp)
{
    if (kp && blockIdx.x == 0 && threadIdx.x == 0) *kp = 0;
    const ncgk::GGm it = items[blockIdx.x];
    if (!first && *it.skip) return;   // converged in an earlier cycle: no restart
    const int r = threadIdx.x;
    bool dn = true;
    if (r < it.nrhs) {
        const cx<double> *hn = (const cx<double> *)it.h + (size_t)2 * it.nrhs * (m + 1);
        double *sc = it.sc + 4 * r;
        int *ic = it.ic + 4 * r;
        const double beta = sqrt(hn[3 * r].re);
        if (first) {
            sc[1] = beta;      // the first residual is b (x = 0)
            ic[2] = 0;
        }
        sc[0] = beta;
        sc[2] = sc[1] > 0.0 ? beta / sc[1] : 0.0;
        dn = sc[2] < it.tol;   // a NaN residual is not converged
        ic[0] = 0;
        ic[1] = dn ? 1 : 0;
        cx<double> *g = (cx<double> *)it.g + (size_t)r * (m + 1);
        g[0] = mk(beta, 0.0);
    }
    g_set_skip(it, dn);
}

__global__ void __launch_bounds__(64) kg_gm_rotate(int kh, const int *kp, int m, const ncgk::GGm *items)
{
    const ncgk::GGm it = items[blockIdx.x];
    if (*it.skip) return;
    const int k = kp ? *kp : kh;
    const int r = threadIdx.x, kc = k + 1;
    bool dn = true;
    if (r < it.nrhs) {
        int *ic = it.ic + 4 * r;
        double *sc = it.sc + 4 * r;
        dn = ic[1] != 0;
        if (!dn) {
            // separate arrays of the member's GMRES state (restrict: the column's stores do
            // not order the loads of the next rotations)
            const cx<double> *__restrict__ h1 = (const cx<double> *)it.h + (size_t)r * kc;
            const cx<double> *__restrict__ h2 = (const cx<double> *)it.h + (size_t)it.nrhs * (m + 1) + (size_t)r * kc;
            const cx<double> *__restrict__ hn = (const cx<double> *)it.h + (size_t)2 * it.nrhs * (m + 1);
            cx<double> *__restrict__ H = (cx<double> *)it.H + (size_t)r * (m + 1) * m;
            cx<double> *__restrict__ cs = (cx<double> *)it.cs + (size_t)r * m;
            cx<double> *__restrict__ sn = (cx<double> *)it.sn + (size_t)r * m;
            cx<double> *__restrict__ g = (cx<double> *)it.g + (size_t)r * (m + 1);
            // the new column h_i = h1_i + h2_i (i <= k), then the previous rotations; the
            // entry carried down the column stays in registers (no store-then-load chain)
            cx<double> a = h1[0] + h2[0];
            for (int i = 0; i < k; ++i) {   // previous rotations
                const cx<double> bb = h1[i + 1] + h2[i + 1];
                const cx<double> ci = mk(cs[i].re, -cs[i].im), si = mk(sn[i].re, -sn[i].im);
                H[(size_t)i * m + k] = ci * a + si * bb;
                a = mk(0.0, 0.0) - sn[i] * a + cs[i] * bb;
            }
            const cx<double> bb = mk(sqrt(hn[3 * r].re), 0.0);
            const double den = sqrt(a.re * a.re + a.im * a.im + bb.re * bb.re + bb.im * bb.im);
            const cx<double> c = den > 0.0 ? a * (1.0 / den) : mk(1.0, 0.0);
            const cx<double> sv = den > 0.0 ? bb * (1.0 / den) : mk(0.0, 0.0);
            cs[k] = c;
            sn[k] = sv;
            H[(size_t)k * m + k] = mk(den, 0.0);
            H[(size_t)(k + 1) * m + k] = mk(0.0, 0.0);
            const cx<double> gk = g[k];
            g[k + 1] = mk(0.0, 0.0) - sv * gk;
            g[k] = mk(c.re, -c.im) * gk;
            ic[0] = k + 1;
            ic[2] += 1;
            const double gn = sqrt(g[k + 1].re * g[k + 1].re + g[k + 1].im * g[k + 1].im);
            sc[2] = sc[1] > 0.0 ? gn / sc[1] : 0.0;
            dn = sc[2] < it.tol;
            ic[1] = dn ? 1 : 0;
        }
    }
    g_set_skip(it, dn);
}

// y = H^-1 g by back substitution over the k rotated columns, zero past k
__global__ void __launch_bounds__(64) kg_gm_solve(int m, const ncgk::GGm *items)
{
    const ncgk::GGm it = items[blockIdx.x];
    const int r = threadIdx.x;
    if (r >= it.nrhs) return;
    const int kk = it.ic[4 * r];
    const cx<double> *H = (const cx<double> *)it.H + (size_t)r * (m + 1) * m;
    const cx<double> *g = (const cx<double> *)it.g + (size_t)r * (m + 1);
    cx<double> *y = (cx<double> *)it.yv + (size_t)r * m;
    for (int i = m - 1; i >= kk; --i) y[i] = mk(0.0, 0.0);
    for (int i = kk - 1; i >= 0; --i) {
        cx<double> s = g[i];
        for (int c = i + 1; c < kk; ++c) s -= H[(size_t)i * m + c] * y[c];
        y[i] = cdiv(s, H[(size_t)i * m + i]);
    }
}

// Z = V y (to_z) or x += V y over the load case's k basis vectors; per item
// ceil(nrhs n / 256) work-groups (no skip: a problem that converged in this cycle still
// takes its update)
__global__ void __launch_bounds__(256) kg_gm_update(int m, int to_z, const ncgk::GGm *items, const int *start,
                                                    int nitems)
{
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GGm it = items[d];
    const long long t = (long long)(blockIdx.x - start[d]) * 256 + threadIdx.x;
    const long long len = (long long)it.nrhs * it.n;
    if (t >= len) return;
    const int r = (int)(t / it.n);
    const long long i = t % it.n;
    const int kk = it.ic[4 * r];
    const cx<double> *V = (const cx<double> *)it.V, *y = (const cx<double> *)it.yv + (size_t)r * m;
    cx<double> *out = to_z ? (cx<double> *)it.Z : (cx<double> *)it.x;
    cx<double> s = to_z ? mk(0.0, 0.0) : out[t];
    for (int k = 0; k < kk; ++k) s += V[((size_t)k * it.nrhs + r) * it.n + i] * y[k];
    out[t] = s;
}

// skip flags of a group: cleared at the start of a solve, gathered for the host per step
__global__ void kg_gm_clear(const ncgk::GGm *items, int nitems)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nitems) *items[i].skip = 0;
}
// one work-group of 64 threads (nitems <= 64)
__global__ void __launch_bounds__(64) kg_gm_flags(const ncgk::GGm *items, int nitems, int *out, int *kp, int depth)
{
    const int i = threadIdx.x;
    const int slot = kp ? *kp % depth : 0;
    if (i < nitems) out[slot * 64 + i] = *items[i].skip;
    if (kp) {
        __syncthreads();   // every lane has read the counter
        if (i == 0) *kp = *kp + 1;
    }
}
// per item and load case: iterations (counter 2) and relative residual estimate (scalar 2)
__global__ void kg_gm_stats(const ncgk::GGm *items, int *its, double *rel)
{
    const ncgk::GGm it = items[blockIdx.x];
    const int r = threadIdx.x;
    if (r < it.nrhs) {
        its[blockIdx.x * 64 + r] = it.ic[4 * r + 2];
        rel[blockIdx.x * 64 + r] = it.sc[4 * r + 2];
    }
}

// squared norms per load case: partial sums per work-group (gsb * nrhs per item, the
// Gram-Schmidt pass's partition), then a fixed tree per load case
__global__ void __launch_bounds__(256) kg_norm_partial(const ncgk::GNorm *items, const int *start, int nitems)
{
    __shared__ double red[256];
    const int d = g_item(start, nitems, blockIdx.x);
    const ncgk::GNorm it = items[d];
    if (it.skip && *it.skip) return;
    const int b = blockIdx.x - start[d], bx = b % it.gsb, r = b / it.gsb, t = threadIdx.x;
    const cx<double> *v = (const cx<double> *)it.v + (size_t)r * it.n;
    double nn = 0.0;
    for (int i = bx * 256 + t; i < it.n; i += it.gsb * 256) nn += v[i].re * v[i].re + v[i].im * v[i].im;
    red[t] = nn;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2) red[t] += red[t + s2];
        __syncthreads();
    }
    if (t == 0) it.npart[(size_t)r * it.gsb + bx] = red[0];
}
__global__ void __launch_bounds__(256) kg_norm_reduce(const ncgk::GNorm *items, double *out)
{
    __shared__ double red[256];
    const ncgk::GNorm it = items[blockIdx.x];
    if (it.skip && *it.skip) return;
    const int t = threadIdx.x;
    for (int r = 0; r < it.nrhs; ++r) {
        red[t] = t < it.gsb ? it.npart[(size_t)r * it.gsb + t] : 0.0;
        __syncthreads();
        for (int s2 = 128; s2 > 0; s2 >>= 1) {
            if (t < s2) red[t] += red[t + s2];
            __syncthreads();
        }
        if (t == 0) out[blockIdx.x * 64 + r] = red[0];
        __syncthreads();
    }
}

}  // namespace ncg

// ======================================================================================
// host launchers (C++ linkage, declared in ncg_kernels.hpp)
// ======================================================================================
namespace ncgk {
using namespace ncg;

static inline unsigned blocks_for(size_t work, unsigned per) { return (unsigned)((work + per - 1) / per); }

hipError_t set_gauss_tables(const double *x, const double *w)
{
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_gl_x), x, sizeof(double) * 36);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_gl_w), w, sizeof(double) * 36);
}

void element_geometry(hipStream_t s, int n, const double *X, const int8_t *nv, double *cen, double *nrm, double *area,
                      double *jac)
{
    ncg_launch(k_element_geometry, dim3(blocks_for(n, 128)), dim3(128), 0, s, n, X, nv, cen, nrm, area, jac);
}

void near_count(hipStream_t s, const DevElements &D, int *count)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    ncg_launch(k_near_count, dim3(blocks_for(D.n, 4)), dim3(256), 0, s, E, count);
}

void near_fill(hipStream_t s, const DevElements &D, const long long *offset, int *rows, int *cols)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    ncg_launch(k_near_fill, dim3(blocks_for(D.n, 4)), dim3(256), 0, s, E, offset, rows, cols);
}

static BMConst make_bm(const BMParams &b)
{
    BMConst c;
    c.tau = b.tau;
    c.beta_im = b.beta_im;
    c.hk = b.hk;
    c.k2 = b.k * b.k;
    c.k = b.k;
    return c;
}

void assemble_far(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, void *A, int mixed,
                  int rows_per_block, unsigned long long *evals)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    dim3 grid(blocks_for(D.n, 256), blocks_for(D.n, rows_per_block));
    if (mixed)
        ncg_launch(k_assemble_far<float, float>, dim3(grid), dim3(256), 0, s, E, make_bm(b), adm, (cx<float> *)A, rows_per_block, evals);
    else
        ncg_launch(k_assemble_far<double, double>, dim3(grid), dim3(256), 0, s, E, make_bm(b), adm, (cx<double> *)A, rows_per_block, evals);
}

void assemble_near(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const int *rows,
                   const int *cols, long long npairs, void *A, int mixed, int *err, unsigned long long *evals)
{
    if (npairs == 0) return;
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const unsigned g = blocks_for(npairs, 128);
    if (mixed)
        ncg_launch(k_assemble_near<float>, dim3(g), dim3(128), 0, s, E, make_bm(b), adm, rows, cols, npairs, (cx<float> *)A, err, evals);
    else
        ncg_launch(k_assemble_near<double>, dim3(g), dim3(128), 0, s, E, make_bm(b), adm, rows, cols, npairs, (cx<double> *)A, err, evals);
}

void rhs_sources(hipStream_t s, const DevElements &D, const BMParams &b, const int *src, int nsrc, const void *bcval,
                 int nrhs, void *rhs_src, int *err, const int *elem_cluster, const unsigned char *cluster_near,
                 int nclusters)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    for (int r0 = 0; r0 < nrhs; r0 += kRhsChunk)
        ncg_launch(k_rhs_sources, dim3(blocks_for(D.n, 4)), dim3(256), 0, s, E, make_bm(b), src, nsrc, (const cx<double> *)bcval, nrhs, r0,
                                                         (cx<double> *)rhs_src, err, elem_cluster, cluster_near,
                                                         nclusters);
}

void assemble_diag(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const void *bcval,
                   int nrhs, const double *ps, int nps, const double *pw, int npw, const void *rhs_src, int has_src,
                   void *A, int mixed, void *rhs, const long long *diag_index, unsigned long long *evals)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    ncg_launch(k_assemble_diag, dim3(blocks_for(D.n, 64)), dim3(64), 0, s, E, make_bm(b), adm, (const cx<double> *)bcval, nrhs, ps, nps, pw,
                                                       npw, (const cx<double> *)rhs_src, has_src, A, mixed,
                                                       (cx<double> *)rhs, diag_index, evals);
}

void matvec(hipStream_t s, const void *A, int mixed, int acc64, int n, int nrhs, const void *x, const double *d, void *y)
{
    const unsigned g = blocks_for(n, 16);
    for (int r0 = 0; r0 < nrhs; r0 += 4) {
        const int nr = std::min(4, nrhs - r0);
        const cx<double> *xx = (const cx<double> *)x + (size_t)r0 * n;
        cx<double> *yy = (cx<double> *)y + (size_t)r0 * n;
#define NCG_MV(S, ACC)                                                                                 \
    switch (nr) {                                                                                      \
    case 1: ncg_launch(k_matvec<S, ACC, 1>, dim3(g), dim3(256), 0, s, (const cx<S> *)A, n, xx, d, yy); break;             \
    case 2: ncg_launch(k_matvec<S, ACC, 2>, dim3(g), dim3(256), 0, s, (const cx<S> *)A, n, xx, d, yy); break;             \
    case 3: ncg_launch(k_matvec<S, ACC, 3>, dim3(g), dim3(256), 0, s, (const cx<S> *)A, n, xx, d, yy); break;             \
    default: ncg_launch(k_matvec<S, ACC, 4>, dim3(g), dim3(256), 0, s, (const cx<S> *)A, n, xx, d, yy); break;            \
    }
        if (!mixed) { NCG_MV(double, double) }
        else if (acc64) { NCG_MV(float, double) }
        else { NCG_MV(float, float) }
#undef NCG_MV
    }
}

// ---- dense operator applied to many load cases on the matrix cores --------------------
// Y = A X for up to 16 load cases per pass is the real GEMM Y^ = A^ X^ with A^ = A read as
// an n x 2n real matrix (re/im interleaved, as stored) and X^ (2n x 32):
//   X^[2j][c] = Re x_c[j], X^[2j+1][c] = -Im x_c[j]        (c < 16: real parts of Y)
//   X^[2j][16+c] = Im x_c[j], X^[2j+1][16+c] = Re x_c[j]   (imaginary parts of Y)
// One wave owns a 32-row tile and all 32 columns in the 16 accumulators of
// v_mfma_f32_32x32x2f32 (D lane l, register v: row 8 (v / 4) + 4 (l / 32) + v % 4,
// column l % 32).  A work-group of 4 waves covers 128 rows; the columns of A are split
// over gridDim.y work-groups (partial sums [split][32][n], added by k_mfma_reduce).
constexpr int kMfmaRows = 128, kMfmaCols = 32;   // rows per work-group, complex columns per LDS chunk

__global__ void __launch_bounds__(256) k_xhat(int n, int nr, const cx<double> *x, float *xh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 32) return;
    const int j = t / 32, c = t % 32, r = c % 16;
    float v0 = 0.f, v1 = 0.f;
    if (r < nr) {
        const cx<double> a = x[(size_t)r * n + j];
        if (c < 16) { v0 = float(a.re); v1 = float(-a.im); }
        else { v0 = float(a.im); v1 = float(a.re); }
    }
    xh[(size_t)(2 * j) * 32 + c] = v0;
    xh[(size_t)(2 * j + 1) * 32 + c] = v1;
}

typedef float ncg_floatx16 __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(256) k_matvec_mfma(const cx<float> *__restrict__ A, int n, int chunks_per_split,
                                                     const float *__restrict__ xh, float *__restrict__ part)
{
    // row stride 66 floats: the 32 rows a wave reads per step fall on distinct banks
    __shared__ float As[kMfmaRows][2 * kMfmaCols + 2];
    __shared__ float Xs[2 * kMfmaCols][32];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int r0 = blockIdx.x * kMfmaRows;
    const int nchunks = (n + kMfmaCols - 1) / kMfmaCols;
    const int c0 = blockIdx.y * chunks_per_split, c1 = min(nchunks, c0 + chunks_per_split);
    // the next chunk is fetched into registers while the matrix cores work on this one:
    // A: 16 complex per thread (rows t / 32 + 8 i, column t % 32); X^: 8 floats per thread
    constexpr int AQ = kMfmaRows * kMfmaCols / 256, XQ = 2 * kMfmaCols * 32 / 256;
    cx<float> ar[AQ];
    float xr[XQ];
    auto fetch = [&](int ch) {
        const int j0 = ch * kMfmaCols, col = j0 + t % kMfmaCols;
#pragma unroll
        for (int i = 0; i < AQ; ++i) {
            const int row = r0 + t / kMfmaCols + 8 * i;
            ar[i] = row < n && col < n ? A[(size_t)row * n + col] : cx<float>{0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < XQ; ++i) {
            const int k = 2 * j0 + t / 32 + 8 * i;
            xr[i] = k < 2 * n ? xh[(size_t)k * 32 + t % 32] : 0.f;
        }
    };
    ncg_floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    if (c0 < c1) fetch(c0);
    for (int ch = c0; ch < c1; ++ch) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < AQ; ++i) {
            As[t / kMfmaCols + 8 * i][2 * (t % kMfmaCols)] = ar[i].re;
            As[t / kMfmaCols + 8 * i][2 * (t % kMfmaCols) + 1] = ar[i].im;
        }
#pragma unroll
        for (int i = 0; i < XQ; ++i) Xs[t / 32 + 8 * i][t % 32] = xr[i];
        __syncthreads();
        if (ch + 1 < c1) fetch(ch + 1);
#pragma unroll 8
        for (int ks = 0; ks < kMfmaCols; ++ks) {
            const float a = As[wv * 32 + (lane & 31)][2 * ks + (lane >> 5)];
            const float b = Xs[2 * ks + (lane >> 5)][lane & 31];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
    }
    // partial sums of this column split: part[split][column][row]
    float *out = part + (size_t)blockIdx.y * 32 * n;
    const int col = lane & 31;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = r0 + wv * 32 + 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
        if (row < n) out[(size_t)col * n + row] = acc[v];
    }
}

// y[r][i] = sum over the splits of the partial real and imaginary parts, r < nr
__global__ void __launch_bounds__(256) k_mfma_reduce(int n, int nr, int splits, const float *part, cx<double> *y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nr) return;
    const int r = t / n, i = t % n;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < splits; ++s) {
        re += part[((size_t)s * 32 + r) * n + i];
        im += part[((size_t)s * 32 + 16 + r) * n + i];
    }
    y[(size_t)r * n + i] = mk(re, im);
}

// The same product with FP64 accumulation (v_mfma_f64_16x16x4f64; D lane l, register v:
// row 4 v + l / 16, column l % 16): the refinement residual of the mixed form (FP32 A)
// and the FP64 operator.  A wave owns 16 rows and the 32 columns in two accumulators
// (real parts, imaginary parts); a work-group of 4 waves covers 64 rows.
constexpr int kMfma64Rows = 64;

__global__ void __launch_bounds__(256) k_xhat64(int n, int nr, const cx<double> *x, double *xh)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * 32) return;
    const int j = t / 32, c = t % 32, r = c % 16;
    double v0 = 0.0, v1 = 0.0;
    if (r < nr) {
        const cx<double> a = x[(size_t)r * n + j];
        if (c < 16) { v0 = a.re; v1 = -a.im; }
        else { v0 = a.im; v1 = a.re; }
    }
    xh[(size_t)(2 * j) * 32 + c] = v0;
    xh[(size_t)(2 * j + 1) * 32 + c] = v1;
}

typedef double ncg_doublex4 __attribute__((ext_vector_type(4)));

template <typename SA>
__global__ void __launch_bounds__(256) k_matvec_mfma64(const cx<SA> *__restrict__ A, int n, int chunks_per_split,
                                                       const double *__restrict__ xh, double *__restrict__ part)
{
    __shared__ SA As[kMfma64Rows][2 * kMfmaCols + 2];
    __shared__ double Xs[2 * kMfmaCols][32];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int r0 = blockIdx.x * kMfma64Rows;
    const int nchunks = (n + kMfmaCols - 1) / kMfmaCols;
    const int c0 = blockIdx.y * chunks_per_split, c1 = min(nchunks, c0 + chunks_per_split);
    constexpr int AQ = kMfma64Rows * kMfmaCols / 256, XQ = 2 * kMfmaCols * 32 / 256;
    cx<SA> ar[AQ];
    double xr[XQ];
    auto fetch = [&](int ch) {
        const int j0 = ch * kMfmaCols, col = j0 + t % kMfmaCols;
#pragma unroll
        for (int i = 0; i < AQ; ++i) {
            const int row = r0 + t / kMfmaCols + 8 * i;
            ar[i] = row < n && col < n ? A[(size_t)row * n + col] : cx<SA>{SA(0), SA(0)};
        }
#pragma unroll
        for (int i = 0; i < XQ; ++i) {
            const int k = 2 * j0 + t / 32 + 8 * i;
            xr[i] = k < 2 * n ? xh[(size_t)k * 32 + t % 32] : 0.0;
        }
    };
    ncg_doublex4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    if (c0 < c1) fetch(c0);
    for (int ch = c0; ch < c1; ++ch) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < AQ; ++i) {
            As[t / kMfmaCols + 8 * i][2 * (t % kMfmaCols)] = ar[i].re;
            As[t / kMfmaCols + 8 * i][2 * (t % kMfmaCols) + 1] = ar[i].im;
        }
#pragma unroll
        for (int i = 0; i < XQ; ++i) Xs[t / 32 + 8 * i][t % 32] = xr[i];
        __syncthreads();
        if (ch + 1 < c1) fetch(ch + 1);
#pragma unroll 4
        for (int ks = 0; ks < 2 * kMfmaCols / 4; ++ks) {
            const int k = 4 * ks + (lane >> 4);
            const double a = double(As[wv * 16 + (lane & 15)][k]);
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Xs[k][lane & 15], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Xs[k][16 + (lane & 15)], acc1, 0, 0, 0);
        }
    }
    double *out = part + (size_t)blockIdx.y * 32 * n;
    const int col = lane & 15;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = r0 + wv * 16 + 4 * v + (lane >> 4);
        if (row < n) {
            out[(size_t)col * n + row] = acc0[v];
            out[(size_t)(16 + col) * n + row] = acc1[v];
        }
    }
}

__global__ void __launch_bounds__(256) k_mfma64_reduce(int n, int nr, int splits, const double *part, cx<double> *y)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nr) return;
    const int r = t / n, i = t % n;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < splits; ++s) {
        re += part[((size_t)s * 32 + r) * n + i];
        im += part[((size_t)s * 32 + 16 + r) * n + i];
    }
    y[(size_t)r * n + i] = mk(re, im);
}

MfmaPlan mfma_plan(int n, int f64, int splits)
{
    MfmaPlan m;
    m.f64 = f64;
    const int rows = f64 ? kMfma64Rows : kMfmaRows, nchunks = (n + kMfmaCols - 1) / kMfmaCols;
    m.tiles = (n + rows - 1) / rows;
    // enough work-groups to cover the device (1024 for the FP64 tiles of 64 rows, 512 for
    // the FP32 ones of 128), each at least 4 column chunks
    m.splits = std::max(1, std::min(std::min(16, ((f64 ? 1024 : 512) + m.tiles - 1) / m.tiles), (nchunks + 3) / 4));
    if (splits > 0) m.splits = std::min(16, splits);
    m.per = (nchunks + m.splits - 1) / m.splits;
    m.xh_elems = (size_t)2 * n * 32;
    m.part_elems = (size_t)m.splits * 32 * n;
    return m;
}

void mfma_xhat(hipStream_t s, const MfmaPlan &m, int n, int nr, const void *x, void *xh)
{
    if (m.f64) ncg_launch(k_xhat64, dim3(blocks_for((size_t)n * 32, 256)), dim3(256), 0, s, n, nr, (const cx<double> *)x, (double *)xh);
    else ncg_launch(k_xhat, dim3(blocks_for((size_t)n * 32, 256)), dim3(256), 0, s, n, nr, (const cx<double> *)x, (float *)xh);
}

void mfma_product(hipStream_t s, const MfmaPlan &m, const void *A, int a_f32, int n, const void *xh, void *part)
{
    const dim3 grid(m.tiles, m.splits);
    if (!m.f64)
        ncg_launch(k_matvec_mfma, dim3(grid), dim3(256), 0, s, (const cx<float> *)A, n, m.per, (const float *)xh, (float *)part);
    else if (a_f32)
        ncg_launch(k_matvec_mfma64<float>, dim3(grid), dim3(256), 0, s, (const cx<float> *)A, n, m.per, (const double *)xh, (double *)part);
    else
        ncg_launch(k_matvec_mfma64<double>, dim3(grid), dim3(256), 0, s, (const cx<double> *)A, n, m.per, (const double *)xh,
                                                     (double *)part);
}

void mfma_sum(hipStream_t s, const MfmaPlan &m, int n, int nr, const void *part, void *y)
{
    if (m.f64)
        ncg_launch(k_mfma64_reduce, dim3(blocks_for((size_t)n * nr, 256)), dim3(256), 0, s, n, nr, m.splits, (const double *)part,
                                                                        (cx<double> *)y);
    else
        ncg_launch(k_mfma_reduce, dim3(blocks_for((size_t)n * nr, 256)), dim3(256), 0, s, n, nr, m.splits, (const float *)part,
                                                                      (cx<double> *)y);
}

int dot_blocks(int n) { return std::min(kDotBlocks, (int)blocks_for(n, 256)); }

void dots(hipStream_t s, int n, int nrhs, int ndots, const void *const *u, const void *const *v, int conj_mask,
          void *partial, void *out)
{
    const int nb = dot_blocks(n);
    ncg_launch(k_dots, dim3(nb), dim3(256), 0, s, n, nrhs, ndots, (const cx<double> *)u[0], (const cx<double> *)v[0],
                              (const cx<double> *)(ndots > 1 ? u[1] : u[0]), (const cx<double> *)(ndots > 1 ? v[1] : v[0]),
                              (const cx<double> *)(ndots > 2 ? u[2] : u[0]), (const cx<double> *)(ndots > 2 ? v[2] : v[0]),
                              conj_mask, (cx<double> *)partial);
    ncg_launch(k_reduce_partials, dim3(1), dim3(256), 0, s, nb, nrhs, ndots, (const cx<double> *)partial, (cx<double> *)out);
}

void cgs_alpha(hipStream_t s, int nrhs, int first, const void *dts, void *sc, const int *active)
{
    ncg_launch(k_cgs_alpha, dim3(1), dim3(64), 0, s, nrhs, first, (const cx<double> *)dts, (cx<double> *)sc, active);
}
void cgs_beta(hipStream_t s, int nrhs, const void *dts, void *sc, const int *active)
{
    ncg_launch(k_cgs_beta, dim3(1), dim3(64), 0, s, nrhs, (const cx<double> *)dts, (cx<double> *)sc, active);
}
void cgs_update1(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *u, const void *v, void *q,
                 void *uq)
{
    ncg_launch(k_cgs_update1, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, (const cx<double> *)sc, active,
                                                                     (const cx<double> *)u, (const cx<double> *)v,
                                                                     (cx<double> *)q, (cx<double> *)uq);
}
void cgs_update2(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *uq, const void *auq,
                 void *x, void *r)
{
    ncg_launch(k_cgs_update2, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, (const cx<double> *)sc, active,
                                                                     (const cx<double> *)uq, (const cx<double> *)auq,
                                                                     (cx<double> *)x, (cx<double> *)r);
}
void cgs_update3(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *r, const void *q, void *u,
                 void *p)
{
    ncg_launch(k_cgs_update3, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, (const cx<double> *)sc, active,
                                                                     (const cx<double> *)r, (const cx<double> *)q,
                                                                     (cx<double> *)u, (cx<double> *)p);
}
void axpby(hipStream_t s, int n, int nrhs, const void *coef, const void *x, size_t ldx, void *y, size_t ldy)
{
    ncg_launch(k_axpby, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, (const cx<double> *)coef, (const cx<double> *)x,
                                                               ldx, (cx<double> *)y, ldy);
}
void scale_vec(hipStream_t s, int n, int nrhs, const double *d, void *x)
{
    ncg_launch(k_scale_vec, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, d, (cx<double> *)x);
}
void row_scaling(hipStream_t s, const void *A, int mixed, int n, double *d)
{
    if (mixed) ncg_launch(k_row_scaling<float>, dim3(blocks_for(n, 4)), dim3(256), 0, s, (const cx<float> *)A, n, d);
    else ncg_launch(k_row_scaling<double>, dim3(blocks_for(n, 4)), dim3(256), 0, s, (const cx<double> *)A, n, d);
}
void row_scaling_csr(hipStream_t s, int n, const long long *ptr, const void *val, int mixed, double *d)
{
    if (mixed) ncg_launch(k_row_scaling_csr<float>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, ptr, (const cx<float> *)val, d);
    else ncg_launch(k_row_scaling_csr<double>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, ptr, (const cx<double> *)val, d);
}
// slices of k_multidot: about 1024 entries each (one load step of four per thread), as
// many as the partial buffer holds
static int multidot_slices(int n, int nrhs, int kcount, size_t part_cap)
{
    const int nsl = std::max(1, (n + 1023) / 1024);
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)nsl, part_cap / ((size_t)kcount * nrhs)));
}
void multidot(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *w, void *h, void *part,
              size_t part_cap)
{
    const int nsl = multidot_slices(n, nrhs, kcount, part_cap);
    const int slice = (n + nsl - 1) / nsl;
    ncg_launch(k_multidot, dim3(dim3(kcount, nrhs, nsl)), dim3(256), 0, s, n, nrhs, kcount, slice, (const cx<double> *)V,
                                                       (const cx<double> *)w, (cx<double> *)part);
    ncg_launch(k_multidot_sum, dim3(blocks_for((size_t)nrhs * kcount, 256)), dim3(256), 0, s, nrhs, kcount, nsl, (const cx<double> *)part,
                                                                           (cx<double> *)h);
}
static unsigned gs_blocks(int n) { return std::min<unsigned>(256u, blocks_for(n, 256)); }
void gs_pass(hipStream_t s, int n, int nrhs, int kcount, const void *V, void *w, void *h, void *part, size_t part_cap,
             double *npart)
{
    const int nsl = multidot_slices(n, nrhs, kcount, part_cap);
    const int slice = (n + nsl - 1) / nsl;
    ncg_launch(k_multidot, dim3(dim3(kcount, nrhs, nsl)), dim3(256), 0, s, n, nrhs, kcount, slice, (const cx<double> *)V,
                                                       (const cx<double> *)w, (cx<double> *)part);
    ncg_launch(k_gs_update, dim3(dim3(gs_blocks(n), nrhs)), dim3(256), 0, s, n, nrhs, kcount, nsl, (const cx<double> *)part,
                                                         (const cx<double> *)V, (cx<double> *)w, (cx<double> *)h, npart);
}
void gs_normalize(hipStream_t s, int n, int nrhs, const double *npart, const void *w, void *v, void *nout)
{
    ncg_launch(k_gs_normalize, dim3(dim3(gs_blocks(n), nrhs)), dim3(256), 0, s, n, (int)gs_blocks(n), npart, (const cx<double> *)w,
                                                            (cx<double> *)v, (cx<double> *)nout);
}
void multiaxpy(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *h, void *w, int add)
{
    if (add)
        ncg_launch(k_multiaxpy_add, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, kcount, (const cx<double> *)V,
                                                                           (const cx<double> *)h, (cx<double> *)w);
    else
        ncg_launch(k_multiaxpy, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, kcount, (const cx<double> *)V,
                                                                       (const cx<double> *)h, (cx<double> *)w);
}
void ilu_count(hipStream_t s, const void *A, int n, const long long *ptr, const int *ncol, const double *d, double thr,
               int *cnt)
{
    if (ptr)
        ncg_launch(k_ilu_count_csr, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, ptr, ncol, (const cx<double> *)A, d, thr, cnt);
    else
        ncg_launch(k_ilu_count_dense, dim3(blocks_for(n, 4)), dim3(256), 0, s, (const cx<double> *)A, n, d, thr, cnt);
}
void ilu_fill(hipStream_t s, const void *A, int n, const long long *ptr, const int *ncol, const double *d, double thr,
              const int *rp, int *col, void *val)
{
    if (ptr)
        ncg_launch(k_ilu_fill_csr, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, ptr, ncol, (const cx<double> *)A, d, thr, rp, col,
                                                           (cx<double> *)val);
    else
        ncg_launch(k_ilu_fill_dense, dim3(blocks_for(n, 4)), dim3(256), 0, s, (const cx<double> *)A, n, d, thr, rp, col,
                                                          (cx<double> *)val);
}
void ilu_factor(hipStream_t s, int nlev, const int *lev_ptr, const int *lev_rows, const int *rp, const int *col,
                const int *diag, const long long *uoff, const int *upd, void *val, int *err)
{
    ncg_launch(k_ilu_factor, dim3(1), dim3(1024), 0, s, nlev, lev_ptr, lev_rows, rp, col, diag, uoff, upd, (cx<double> *)val, err);
}
void ilu_pack(hipStream_t s, int cnt, const int *map, const void *val, void *out)
{
    if (cnt > 0)
        ncg_launch(k_ilu_pack, dim3(blocks_for(cnt, 256)), dim3(256), 0, s, cnt, map, (const cx<double> *)val, (cx<double> *)out);
}
void ilu_solve(hipStream_t s, int n, int nrhs, const IluSweep &L, const IluSweep &U, const void *x, void *y)
{
    ncg_launch(k_ilu_solve, dim3(nrhs), dim3(ncgk::kIluRows), 0, s, n, L, U, (const cx<double> *)x, (cx<double> *)y);
}
void lu_factor(hipStream_t s, void *A, int n)
{
    for (int i = 0; i < n - 1; ++i) {
        const int rem = n - 1 - i;
        ncg_launch(k_lu_update, dim3(dim3(blocks_for(rem, 256), rem)), dim3(256), 0, s, (cx<double> *)A, n, i);
        ncg_launch(k_lu_scale, dim3(blocks_for(rem, 256)), dim3(256), 0, s, (cx<double> *)A, n, i);
    }
}
void lu_solve(hipStream_t s, const void *A, int n, int nrhs, void *b)
{
    ncg_launch(k_lu_solve, dim3(nrhs), dim3(1024), 0, s, (const cx<double> *)A, n, (cx<double> *)b);
}
void field_prep(hipStream_t s, int n, int nrhs, const int8_t *kind, const double *adm, double tau, const void *bcval,
                const void *x, void *pot, void *vel)
{
    ncg_launch(k_field_prep, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, kind, adm, tau, (const cx<double> *)bcval,
                                                                    (const cx<double> *)x, (cx<double> *)pot,
                                                                    (cx<double> *)vel);
}
void field(hipStream_t s, const DevElements &D, const BMParams &b, const double *pts, int m, int nrhs, const void *pot,
           const void *vel, const double *ps, int nps, const double *pw, int npw, void *phi, void *grad, int fp32,
           int *err, const int *pt_ev, const int *el_ptr, const int *el_idx, const void *fphi, const void *fgrad,
           unsigned long long *evals)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    for (int r0 = 0; r0 < nrhs; r0 += kFieldChunk) {
        if (fp32)
            ncg_launch(k_field<float>, dim3(blocks_for(m, 4)), dim3(256), 0, s, E, make_bm(b), pts, m, nrhs, r0, (const cx<double> *)pot,
                                                            (const cx<double> *)vel, ps, nps, pw, npw, (cx<double> *)phi,
                                                            (cx<double> *)grad, err, pt_ev, el_ptr, el_idx,
                                                            (const cx<double> *)fphi, (const cx<double> *)fgrad, evals);
        else
            ncg_launch(k_field<double>, dim3(blocks_for(m, 4)), dim3(256), 0, s, E, make_bm(b), pts, m, nrhs, r0, (const cx<double> *)pot,
                                                             (const cx<double> *)vel, ps, nps, pw, npw, (cx<double> *)phi,
                                                             (cx<double> *)grad, err, pt_ev, el_ptr, el_idx,
                                                             (const cx<double> *)fphi, (const cx<double> *)fgrad, evals);
    }
}

// ---- FMM launchers -----------------------------------------------------------------
void fmm_D(hipStream_t s, int L, int P, double k, double h, const double *ccen, int ncl, const int *fan_row,
           const int *fan_idx, int nfan, const double *u, void *Dv, int mixed, double *coef)
{
    if (nfan == 0) return;
    ncg_launch(k_fmm_D_coef, dim3(blocks_for(nfan, 128)), dim3(128), 0, s, nfan, L, k, h, ccen, fan_row, fan_idx, coef);
    if (mixed) ncg_launch(k_fmm_D<float>, dim3(nfan), dim3(128), 0, s, L, P, coef, u, (cx<float> *)Dv);
    else ncg_launch(k_fmm_D<double>, dim3(nfan), dim3(128), 0, s, L, P, coef, u, (cx<double> *)Dv);
}
void fmm_T(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, int P, const double *ccen,
           const int *elem_cl, const int *elem_pos, const double *u, void *Tv, int mixed)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const dim3 g(blocks_for(D.n, 4), (P + 63) / 64);
    if (mixed) ncg_launch(k_fmm_T<float>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, P, ccen, elem_cl, elem_pos, u, (cx<float> *)Tv);
    else ncg_launch(k_fmm_T<double>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, P, ccen, elem_cl, elem_pos, u, (cx<double> *)Tv);
}
void fmm_Tvec_bc(hipStream_t s, const DevElements &D, const BMParams &b, int P, int ncl, const double *ccen,
                 const int *cl_ptr, const int *cl_elem, const double *u, const void *bcval, int nrhs, void *tv)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    ncg_launch(k_fmm_Tvec_bc, dim3(blocks_for((size_t)ncl * P, 256)), dim3(256), 0, s, E, make_bm(b), P, ncl, ccen, cl_ptr, cl_elem, u,
                                                                   (const cx<double> *)bcval, nrhs, (cx<double> *)tv);
}
void fmm_S(hipStream_t s, const DevElements &D, const BMParams &b, int P, const double *ccen, const int *elem_cl,
           const int *cl_ptr, const int *cl_elem, const double *u, const double *w, void *Sv, int mixed)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const unsigned g = blocks_for((size_t)D.n * P, 256);
    if (mixed)
        ncg_launch(k_fmm_S<float>, dim3(g), dim3(256), 0, s, E, make_bm(b), P, ccen, elem_cl, cl_ptr, cl_elem, u, w, (cx<float> *)Sv);
    else
        ncg_launch(k_fmm_S<double>, dim3(g), dim3(256), 0, s, E, make_bm(b), P, ccen, elem_cl, cl_ptr, cl_elem, u, w, (cx<double> *)Sv);
}
void fmm_near(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const RowChunks &rc,
              const int *cl_ptr, const int *cl_elem, const long long *ptr, const int *col, int wmax, void *vals,
              int mixed, unsigned long long *evals, int rows_mode)
{
    if (rc.n == 0 || wmax == 0) return;
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const dim3 g(rc.n, (wmax + 255) / 256);
    if (mixed && rows_mode)
        ncg_launch(k_fmm_near_rows<float, float>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, rc.cl, rc.r0, rc.m, cl_ptr, cl_elem, ptr,
                                                        col, (cx<float> *)vals, evals);
    else if (mixed)
        ncg_launch(k_fmm_near<float, float>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, rc.cl, rc.r0, rc.m, cl_ptr, cl_elem, ptr, col,
                                                   (cx<float> *)vals, evals);
    else if (rows_mode)
        ncg_launch(k_fmm_near_rows<double, double>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, rc.cl, rc.r0, rc.m, cl_ptr, cl_elem, ptr,
                                                          col, (cx<double> *)vals, evals);
    else
        ncg_launch(k_fmm_near<double, double>, dim3(g), dim3(256), 0, s, E, make_bm(b), adm, rc.cl, rc.r0, rc.m, cl_ptr, cl_elem, ptr, col,
                                                     (cx<double> *)vals, evals);
}
void adj_positions(hipStream_t s, long long npairs, const int *rows, const int *cols, const int *elem_cl,
                   const int *elem_pos, const int *cl_ptr, const int *nn_ptr, const int *nn_idx, const int *nn_off,
                   const long long *near_ptr, long long *pos)
{
    if (npairs == 0) return;
    ncg_launch(k_adj_positions, dim3(blocks_for(npairs, 256)), dim3(256), 0, s, npairs, rows, cols, elem_cl, elem_pos, cl_ptr, nn_ptr,
                                                            nn_idx, nn_off, near_ptr, pos);
}
void fmm_near_adj(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const int *rows,
                  const int *cols, const long long *pos, long long npairs, void *vals, int mixed, int *err,
                  unsigned long long *evals)
{
    if (npairs == 0) return;
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const unsigned g = blocks_for(npairs, 128);
    if (mixed)
        ncg_launch(k_fmm_near_adj<float>, dim3(g), dim3(128), 0, s, E, make_bm(b), adm, rows, cols, pos, npairs, (cx<float> *)vals, err,
                                                evals);
    else
        ncg_launch(k_fmm_near_adj<double>, dim3(g), dim3(128), 0, s, E, make_bm(b), adm, rows, cols, pos, npairs, (cx<double> *)vals, err,
                                                 evals);
}
void near_csr(hipStream_t s, int n, const int *elem_cl, const int *cl_ptr, const int *cl_elem, const int *nn_ptr,
              const int *nn_idx, const long long *ptr, int *rows, int *cols, long long *diag)
{
    ncg_launch(k_near_csr, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, elem_cl, cl_ptr, cl_elem, nn_ptr, nn_idx, ptr, rows, cols, diag);
}

// load cases in chunks of up to 4 per launch: fn(NR, r0) launches one chunk
template <typename F>
static void rhs_chunks(int nrhs, F &&fn)
{
    for (int r0 = 0; r0 < nrhs; r0 += 4) {
        const int nr = std::min(4, nrhs - r0);
        fn(nr, r0);
    }
}

void bj_setup(hipStream_t s, int n, int nblk, const int *elem_blk, const int *elem_pos, const int *blk_cl,
              const int *blk_b0, const int *blk_m, const int *cl_ptr, const long long *near_ptr, const int *self_off,
              const void *nval, int mixed, const long long *boff, void *blk, int *err)
{
    if (mixed) {
        ncg_launch(k_bj_gather<float, float>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, elem_blk, elem_pos, blk_cl, blk_b0, blk_m, cl_ptr,
                                                                   near_ptr, self_off, (const cx<float> *)nval, boff,
                                                                   (cx<float> *)blk);
        ncg_launch(k_bj_invert<float>, dim3(nblk), dim3(256), 0, s, blk_m, boff, (cx<float> *)blk, err);
    } else {
        ncg_launch(k_bj_gather<double, double>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, elem_blk, elem_pos, blk_cl, blk_b0, blk_m,
                                                                     cl_ptr, near_ptr, self_off,
                                                                     (const cx<double> *)nval, boff, (cx<double> *)blk);
        ncg_launch(k_bj_invert<double>, dim3(nblk), dim3(256), 0, s, blk_m, boff, (cx<double> *)blk, err);
    }
}
void bj_setup_dense(hipStream_t s, int n, int nblk, const int *elem_blk, const int *blk_b0, const int *blk_m,
                    const void *A, int mixed, const long long *boff, void *blk, int *err)
{
    if (mixed) {
        ncg_launch(k_bj_gather_dense<float, float>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, elem_blk, blk_b0, blk_m, (const cx<float> *)A,
                                                                         boff, (cx<float> *)blk);
        ncg_launch(k_bj_invert<float>, dim3(nblk), dim3(256), 0, s, blk_m, boff, (cx<float> *)blk, err);
    } else {
        ncg_launch(k_bj_gather_dense<double, double>, dim3(blocks_for(n, 4)), dim3(256), 0, s, n, elem_blk, blk_b0, blk_m,
                                                                           (const cx<double> *)A, boff,
                                                                           (cx<double> *)blk);
        ncg_launch(k_bj_invert<double>, dim3(nblk), dim3(256), 0, s, blk_m, boff, (cx<double> *)blk, err);
    }
}
template <typename TB>
static void bj_apply_t(hipStream_t s, int n, int nrhs, int nblk, const int *blk_b0, const int *blk_m,
                       const int *cl_elem, const long long *boff, const cx<TB> *M, const void *x, void *y)
{
    // up to 8 load cases per pass over the inverses
    for (int r0 = 0; r0 < nrhs; r0 += 8) {
        const int nr = std::min(8, nrhs - r0);
        const cx<double> *xx = (const cx<double> *)x + (size_t)r0 * n;
        cx<double> *yy = (cx<double> *)y + (size_t)r0 * n;
        switch (nr) {
        case 1: ncg_launch(k_bj_apply<TB, 1>, dim3(nblk), dim3(256), 0, s, n, 1, blk_b0, blk_m, cl_elem, boff, M, xx, yy); break;
        case 2: ncg_launch(k_bj_apply<TB, 2>, dim3(nblk), dim3(256), 0, s, n, 2, blk_b0, blk_m, cl_elem, boff, M, xx, yy); break;
        case 3: ncg_launch(k_bj_apply<TB, 3>, dim3(nblk), dim3(256), 0, s, n, 3, blk_b0, blk_m, cl_elem, boff, M, xx, yy); break;
        case 4: ncg_launch(k_bj_apply<TB, 4>, dim3(nblk), dim3(256), 0, s, n, 4, blk_b0, blk_m, cl_elem, boff, M, xx, yy); break;
        default: ncg_launch(k_bj_apply<TB, 8>, dim3(nblk), dim3(256), 0, s, n, nr, blk_b0, blk_m, cl_elem, boff, M, xx, yy); break;
        }
    }
}
void bj_apply(hipStream_t s, int n, int nrhs, int nblk, const int *blk_b0, const int *blk_m, const int *cl_elem,
              const long long *boff, const void *minv, int mixed, const void *x, void *y)
{
    if (mixed) bj_apply_t(s, n, nrhs, nblk, blk_b0, blk_m, cl_elem, boff, (const cx<float> *)minv, x, y);
    else bj_apply_t(s, n, nrhs, nblk, blk_b0, blk_m, cl_elem, boff, (const cx<double> *)minv, x, y);
}
void scale_norm(hipStream_t s, int n, int nrhs, const void *dts, const void *w, void *v)
{
    ncg_launch(k_scale_norm, dim3(blocks_for((size_t)n * nrhs, 256)), dim3(256), 0, s, n, nrhs, (const cx<double> *)dts,
                                                                    (const cx<double> *)w, (cx<double> *)v);
}

#define NCG_TARGS(...) __VA_ARGS__
#define NCG_NR_DISPATCH_SH(NR, KERNEL, TARGS, GRID, BLOCK, SH, ...)                     \
    switch (NR) {                                                                       \
    case 1: ncg_launch(KERNEL<TARGS, 1>, dim3(GRID), dim3(BLOCK), (SH) * 1, s, __VA_ARGS__); break;         \
    case 2: ncg_launch(KERNEL<TARGS, 2>, dim3(GRID), dim3(BLOCK), (SH) * 2, s, __VA_ARGS__); break;         \
    case 3: ncg_launch(KERNEL<TARGS, 3>, dim3(GRID), dim3(BLOCK), (SH) * 3, s, __VA_ARGS__); break;         \
    default: ncg_launch(KERNEL<TARGS, 4>, dim3(GRID), dim3(BLOCK), (SH) * 4, s, __VA_ARGS__); break;        \
    }

// LDS budget of one near-field chunk: seg columns x NR load cases x 16 bytes
constexpr int kNearLdsBytes = 32768;

template <typename S>
static void near_spmv_t(hipStream_t s, int n, int nrhs, const NearChunks &nc, const int *col, int wmax,
                        const cx<S> *val, const cx<double> *x, cx<double> *y)
{
    rhs_chunks(nrhs, [&](int nr, int r0) {
        // LDS for the longest row (in 64-column steps), at most the budget: more chunks per CU
        const int seg = std::max(64, std::min(kNearLdsBytes / (16 * nr), (wmax + 63) / 64 * 64));
        NCG_NR_DISPATCH_SH(nr, k_near_spmv, S, nc.n, 256, 16 * seg, n, seg, nc.cb, nc.w, nc.row, nc.off, col, val,
                           x + (size_t)r0 * n, y + (size_t)r0 * n);
    });
}
void near_spmv(hipStream_t s, int n, int nrhs, const NearChunks &nc, const int *col, int wmax, const void *val,
               int mixed, const void *x, void *y)
{
    if (mixed)
        near_spmv_t(s, n, nrhs, nc, col, wmax, (const cx<float> *)val, (const cx<double> *)x, (cx<double> *)y);
    else
        near_spmv_t(s, n, nrhs, nc, col, wmax, (const cx<double> *)val, (const cx<double> *)x, (cx<double> *)y);
}
template <typename S, typename W>
static void apply_T_t(hipStream_t s, int n, int ncl, int P, int nrhs, const int *cl_ptr, const int *cl_elem,
                      const cx<S> *Tv, const cx<double> *x, cx<W> *tw)
{
    const unsigned g = (unsigned)ncl * (unsigned)((P + 63) / 64);
    rhs_chunks(nrhs, [&](int nr, int r0) {
        NCG_NR_DISPATCH_SH(nr, k_fmm_apply_T, NCG_TARGS(S, W), g, 64 * kSumWaves, 0, n, ncl, P, cl_ptr, cl_elem, Tv,
                           x + (size_t)r0 * n, tw + (size_t)r0 * ncl * P);
    });
}
void fmm_apply_T(hipStream_t s, int n, int ncl, int P, int nrhs, const int *cl_ptr, const int *cl_elem, const void *Tv,
                 int f32, int tw_f64, const void *x, void *tw)
{
    if (f32 && tw_f64)
        apply_T_t(s, n, ncl, P, nrhs, cl_ptr, cl_elem, (const cx<float> *)Tv, (const cx<double> *)x, (cx<double> *)tw);
    else if (f32)
        apply_T_t(s, n, ncl, P, nrhs, cl_ptr, cl_elem, (const cx<float> *)Tv, (const cx<double> *)x, (cx<float> *)tw);
    else
        apply_T_t(s, n, ncl, P, nrhs, cl_ptr, cl_elem, (const cx<double> *)Tv, (const cx<double> *)x, (cx<double> *)tw);
}
template <typename S, typename WI, typename WO>
static void apply_D_t(hipStream_t s, int ncl, int P, int nrhs, const int *fan_ptr, const int *fan_idx, const cx<S> *Dv,
                      const cx<WI> *tw, cx<WO> *sw)
{
    const unsigned g = (unsigned)ncl * (unsigned)((P + 63) / 64);
    rhs_chunks(nrhs, [&](int nr, int r0) {
        NCG_NR_DISPATCH_SH(nr, k_fmm_apply_D, NCG_TARGS(S, WI, WO), g, 64 * kSumWaves, 0, ncl, P, fan_ptr, fan_idx, Dv,
                           tw + (size_t)r0 * ncl * P, sw + (size_t)r0 * ncl * P);
    });
}
void fmm_apply_D(hipStream_t s, int ncl, int P, int nrhs, const int *fan_ptr, const int *fan_idx, const void *Dv,
                 int f32, int tw_f64, const void *tw, void *sw)
{
    if (f32 && tw_f64)
        apply_D_t(s, ncl, P, nrhs, fan_ptr, fan_idx, (const cx<float> *)Dv, (const cx<double> *)tw, (cx<float> *)sw);
    else if (f32)
        apply_D_t(s, ncl, P, nrhs, fan_ptr, fan_idx, (const cx<float> *)Dv, (const cx<float> *)tw, (cx<float> *)sw);
    else
        apply_D_t(s, ncl, P, nrhs, fan_ptr, fan_idx, (const cx<double> *)Dv, (const cx<double> *)tw, (cx<double> *)sw);
}
template <typename S, typename W>
static void apply_S_t(hipStream_t s, int n, int ncl, int P, int nrhs, const RowChunks &tl, const int *cl_ptr,
                      const int *cl_elem, const cx<S> *Sv, const cx<W> *sw, cx<double> *y)
{
    rhs_chunks(nrhs, [&](int nr, int r0) {
        NCG_NR_DISPATCH_SH(nr, k_fmm_apply_S, NCG_TARGS(S, W), tl.n, 64 * kSumWaves, 0, n, ncl, P, tl.cl, tl.r0, tl.m,
                           cl_ptr, cl_elem, Sv, sw + (size_t)r0 * ncl * P, y + (size_t)r0 * n);
    });
}
void fmm_apply_S(hipStream_t s, int n, int ncl, int P, int nrhs, const RowChunks &tl, const int *cl_ptr,
                 const int *cl_elem, const void *Sv, int f32, const void *sw, void *y)
{
    if (f32)
        apply_S_t(s, n, ncl, P, nrhs, tl, cl_ptr, cl_elem, (const cx<float> *)Sv, (const cx<float> *)sw,
                  (cx<double> *)y);
    else
        apply_S_t(s, n, ncl, P, nrhs, tl, cl_ptr, cl_elem, (const cx<double> *)Sv, (const cx<double> *)sw,
                  (cx<double> *)y);
}
void fmm_Tvec_sol(hipStream_t s, const DevElements &D, const BMParams &b, int P, int ncl, const double *ccen,
                  const int *cl_ptr, const int *cl_elem, const double *u, const void *pot, const void *vel, int nrhs,
                  void *tv, int fp32, int tv32)
{
    Elements E{D.n, D.X, D.nv, D.cen, D.nrm, D.area, D.jac, D.kind};
    const unsigned g = blocks_for((size_t)ncl * P, 256);
    const cx<double> *pp = (const cx<double> *)pot, *vv = (const cx<double> *)vel;
    if (fp32 && tv32)
        ncg_launch(k_fmm_Tvec_sol<float, float>, dim3(g), dim3(256), 0, s, E, make_bm(b), P, ncl, ccen, cl_ptr, cl_elem,
                   u, pp, vv, nrhs, (cx<float> *)tv);
    else if (fp32)
        ncg_launch(k_fmm_Tvec_sol<float, double>, dim3(g), dim3(256), 0, s, E, make_bm(b), P, ncl, ccen, cl_ptr,
                   cl_elem, u, pp, vv, nrhs, (cx<double> *)tv);
    else
        ncg_launch(k_fmm_Tvec_sol<double, double>, dim3(g), dim3(256), 0, s, E, make_bm(b), P, ncl, ccen, cl_ptr,
                   cl_elem, u, pp, vv, nrhs, (cx<double> *)tv);
}
void fmm_eval_hankel(hipStream_t s, int L, double k, double h, const double *ecen, const double *ccen,
                     const int *pair_ev, const int *pair_cl, int npair, double *tab)
{
    if (npair == 0) return;
    ncg_launch(k_fmm_pair_hankel, dim3(blocks_for(npair, 128)), dim3(128), 0, s, npair, L, k, h, ecen, ccen, pair_ev, pair_cl, tab);
}
template <typename T, typename TV>
static void eval_z_t(hipStream_t s, int nev, int L, int P, double h, const int *ev_ptr, const int *pair_cl,
                     const double *tab, const double *u, const double *w, const void *tv, int ncl, int nrhs, void *z)
{
    rhs_chunks(nrhs, [&](int nr, int r0) {
        const cx<TV> *t = (const cx<TV> *)tv + (size_t)r0 * ncl * P;
        cx<double> *zz = (cx<double> *)z + (size_t)r0 * nev * P;
        switch (nr) {
        case 1: ncg_launch(k_fmm_eval_z<1, T, TV>, dim3(nev), dim3(256), 0, s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, t, ncl, zz); break;
        case 2: ncg_launch(k_fmm_eval_z<2, T, TV>, dim3(nev), dim3(256), 0, s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, t, ncl, zz); break;
        case 3: ncg_launch(k_fmm_eval_z<3, T, TV>, dim3(nev), dim3(256), 0, s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, t, ncl, zz); break;
        default: ncg_launch(k_fmm_eval_z<4, T, TV>, dim3(nev), dim3(256), 0, s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, t, ncl, zz); break;
        }
    });
}
void fmm_eval_z(hipStream_t s, int nev, int L, int P, double h, const int *ev_ptr, const int *pair_cl,
                const double *tab, const double *u, const double *w, const void *tv, int ncl, int nrhs, void *z,
                int fp32, int tv32)
{
    if (nev == 0) return;
    if (fp32 && tv32) eval_z_t<float, float>(s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, tv, ncl, nrhs, z);
    else if (fp32) eval_z_t<float, double>(s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, tv, ncl, nrhs, z);
    else eval_z_t<double, double>(s, nev, L, P, h, ev_ptr, pair_cl, tab, u, w, tv, ncl, nrhs, z);
}
void fmm_eval_far(hipStream_t s, int m, int P, int nev, double hk, const double *pts, const int *pt_ev,
                  const double *ecen, const double *u, int nrhs, const void *z, void *fphi, void *fgrad)
{
    ncg_launch(k_fmm_eval_far, dim3(blocks_for(m, 4)), dim3(256), 0, s, m, P, nev, hk, pts, pt_ev, ecen, u, nrhs, (const cx<double> *)z,
                                                    (cx<double> *)fphi, (cx<double> *)fgrad);
}

// ---- grouped launches --------------------------------------------------------------------
#define NCG_G_NR(NR, CALL) \
    switch (NR) {          \
    case 1: CALL(1); break; \
    case 2: CALL(2); break; \
    case 3: CALL(3); break; \
    default: CALL(4); break; \
    }

// work-groups of a persistent launch: what the device holds at once (work-groups per CU of
// the kernel at this LDS size, times the CUs), at most one per unit of work
// (the occupancy query is made once per kernel, LDS size and device)
template <typename K>
static int resident_grid(K kernel, int lds, int units)
{
    static std::mutex mu;
    static std::map<std::tuple<const void *, int, int>, int> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const auto key = std::make_tuple((const void *)kernel, lds, dev);
    int grid = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto f = known.find(key);
        if (f != known.end()) grid = f->second;
    }
    if (!grid) {
        int cus = 0, per = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, lds) != hipSuccess || per < 1) per = 1;
        grid = per * cus;
        std::lock_guard<std::mutex> g(mu);
        known[key] = grid;
    }
    return std::max(1, std::min(units, grid));
}

void g_near_spmv(hipStream_t s, int mixed, int mode, int nr, int seg, const GNear *items, const int *start,
                 int nitems, int blocks, int *queue)
{
    if (blocks <= 0) return;
    const int lds = 16 * seg * std::max(1, std::min(nr, 4));
#define NCG_QCALL(S, NR, DYN)                                                                                      \
    {                                                                                                              \
        const int g = resident_grid(kq_near_spmv<S, NR, DYN>, lds, blocks);                                        \
        ncg_launch(kq_near_spmv<S, NR, DYN>, dim3(g), dim3(256), lds, s, items, start, nitems, seg, queue);                           \
    }
#define NCG_CALL(NR)                                                                                               \
    if (mode == 2 && mixed) NCG_QCALL(float, NR, true)                                                            \
    else if (mode == 2) NCG_QCALL(double, NR, true)                                                               \
    else if (mode == 3 && mixed) NCG_QCALL(float, NR, false)                                                      \
    else if (mode == 3) NCG_QCALL(double, NR, false)                                                              \
    else if (mode && mixed) ncg_launch(kg_near_cluster<float, NR>, dim3(blocks), dim3(256), lds, s, items, start, nitems, seg);        \
    else if (mode) ncg_launch(kg_near_cluster<double, NR>, dim3(blocks), dim3(256), lds, s, items, start, nitems, seg);                \
    else if (mixed) ncg_launch(kg_near_spmv<float, NR>, dim3(blocks), dim3(256), lds, s, items, start, nitems, seg);                   \
    else ncg_launch(kg_near_spmv<double, NR>, dim3(blocks), dim3(256), lds, s, items, start, nitems, seg);
    NCG_G_NR(nr, NCG_CALL)
#undef NCG_CALL
#undef NCG_QCALL
}

// waves per work-group of the grouped far-field applies: 1, 2, 4 or 8
#define NCG_G_NW(NW_, NR_, CALL) \
    switch (NW_) {                \
    case 1: NCG_G_NR(NR_, CALL##1) break; \
    case 2: NCG_G_NR(NR_, CALL##2) break; \
    case 4: NCG_G_NR(NR_, CALL##4) break; \
    default: NCG_G_NR(NR_, CALL##8) break; \
    }
template <typename S, typename W>
static void g_T_t(hipStream_t s, int nw, int nr, const GLevel *items, const int *start, int nitems, int blocks)
{
#define NCG_CALL1(NR) ncg_launch(kg_apply_T<S, W, NR, 1>, dim3(blocks), dim3(64), 0, s, items, start, nitems);
#define NCG_CALL2(NR) ncg_launch(kg_apply_T<S, W, NR, 2>, dim3(blocks), dim3(128), 0, s, items, start, nitems);
#define NCG_CALL4(NR) ncg_launch(kg_apply_T<S, W, NR, 4>, dim3(blocks), dim3(256), 0, s, items, start, nitems);
#define NCG_CALL8(NR) ncg_launch(kg_apply_T<S, W, NR, 8>, dim3(blocks), dim3(512), 0, s, items, start, nitems);
    NCG_G_NW(nw, nr, NCG_CALL)
#undef NCG_CALL1
#undef NCG_CALL2
#undef NCG_CALL4
#undef NCG_CALL8
}
template <typename S, typename W>
static void g_D_t(hipStream_t s, int nw, int nr, const GLevel *items, const int *start, int nitems, int blocks)
{
#define NCG_CALL1(NR) ncg_launch(kg_apply_D<S, W, NR, 1>, dim3(blocks), dim3(64), 0, s, items, start, nitems);
#define NCG_CALL2(NR) ncg_launch(kg_apply_D<S, W, NR, 2>, dim3(blocks), dim3(128), 0, s, items, start, nitems);
#define NCG_CALL4(NR) ncg_launch(kg_apply_D<S, W, NR, 4>, dim3(blocks), dim3(256), 0, s, items, start, nitems);
#define NCG_CALL8(NR) ncg_launch(kg_apply_D<S, W, NR, 8>, dim3(blocks), dim3(512), 0, s, items, start, nitems);
    NCG_G_NW(nw, nr, NCG_CALL)
#undef NCG_CALL1
#undef NCG_CALL2
#undef NCG_CALL4
#undef NCG_CALL8
}
template <typename S, typename W>
static void g_S_t(hipStream_t s, int nw, int nr, const GLevel *items, const int *start, int nitems, int blocks)
{
#define NCG_CALL1(NR) ncg_launch(kg_apply_S<S, W, NR, 1>, dim3(blocks), dim3(64), 0, s, items, start, nitems);
#define NCG_CALL2(NR) ncg_launch(kg_apply_S<S, W, NR, 2>, dim3(blocks), dim3(128), 0, s, items, start, nitems);
#define NCG_CALL4(NR) ncg_launch(kg_apply_S<S, W, NR, 4>, dim3(blocks), dim3(256), 0, s, items, start, nitems);
#define NCG_CALL8(NR) ncg_launch(kg_apply_S<S, W, NR, 8>, dim3(blocks), dim3(512), 0, s, items, start, nitems);
    NCG_G_NW(nw, nr, NCG_CALL)
#undef NCG_CALL1
#undef NCG_CALL2
#undef NCG_CALL4
#undef NCG_CALL8
}
#undef NCG_G_NW
void g_apply_T(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks)
{
    if (blocks <= 0) return;
    if (f32 && !sig64) g_T_t<float, float>(s, nw, nr, items, start, nitems, blocks);
    else if (f32) g_T_t<float, double>(s, nw, nr, items, start, nitems, blocks);
    else g_T_t<double, double>(s, nw, nr, items, start, nitems, blocks);
}
void g_apply_D(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks)
{
    if (blocks <= 0) return;
    if (f32 && !sig64) g_D_t<float, float>(s, nw, nr, items, start, nitems, blocks);
    else if (f32) g_D_t<float, double>(s, nw, nr, items, start, nitems, blocks);
    else g_D_t<double, double>(s, nw, nr, items, start, nitems, blocks);
}
void g_apply_S(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks)
{
    if (blocks <= 0) return;
    if (f32 && !sig64) g_S_t<float, float>(s, nw, nr, items, start, nitems, blocks);
    else if (f32) g_S_t<float, double>(s, nw, nr, items, start, nitems, blocks);
    else g_S_t<double, double>(s, nw, nr, items, start, nitems, blocks);
}
template <typename TB>
static void g_bj_t(hipStream_t s, int nr, const GBj *items, const int *start, int nitems, int blocks, const int *kp)
{
    switch (nr) {
    case 1: ncg_launch(kg_bj_apply<TB, 1>, dim3(blocks), dim3(256), 0, s, items, start, nitems, kp); break;
    case 2: ncg_launch(kg_bj_apply<TB, 2>, dim3(blocks), dim3(256), 0, s, items, start, nitems, kp); break;
    case 3: ncg_launch(kg_bj_apply<TB, 3>, dim3(blocks), dim3(256), 0, s, items, start, nitems, kp); break;
    case 4: ncg_launch(kg_bj_apply<TB, 4>, dim3(blocks), dim3(256), 0, s, items, start, nitems, kp); break;
    default: ncg_launch(kg_bj_apply<TB, 8>, dim3(blocks), dim3(256), 0, s, items, start, nitems, kp); break;
    }
}
void g_bj_apply(hipStream_t s, int f32, int nr, const GBj *items, const int *start, int nitems, int blocks,
                const int *kp)
{
    if (blocks <= 0) return;
    if (f32) g_bj_t<float>(s, nr, items, start, nitems, blocks, kp);
    else g_bj_t<double>(s, nr, items, start, nitems, blocks, kp);
}
#undef NCG_G_NR
void g_vec(hipStream_t s, int op, const GVec *items, const int *start, int nitems, int blocks, const int *kp)
{
    if (blocks > 0) ncg_launch(kg_vec, dim3(blocks), dim3(256), 0, s, op, items, start, nitems, kp);
}
void g_multidot(hipStream_t s, const int *kp, const GGm *items, const int *start, int nitems, int blocks)
{
    if (blocks > 0) ncg_launch(kg_multidot, dim3(blocks), dim3(256), 0, s, kp, items, start, nitems);
}
void g_gs_update(hipStream_t s, int kcount, const int *kp, int pass, int m, const GGm *items, const int *start,
                 int nitems, int blocks)
{
    if (blocks > 0) ncg_launch(kg_gs_update, dim3(blocks), dim3(256), 0, s, kcount, kp, pass, m, items, start, nitems);
}
void g_gs_normalize(hipStream_t s, int k, const int *kp, int m, const GGm *items, const int *start, int nitems,
                    int blocks)
{
    if (blocks > 0) ncg_launch(kg_gs_normalize, dim3(blocks), dim3(256), 0, s, k, kp, m, items, start, nitems);
}
void g_gm_start(hipStream_t s, int m, int first, const GGm *items, int nitems, int *kp)
{
    if (nitems > 0) ncg_launch(kg_gm_start, dim3(nitems), dim3(64), 0, s, m, first, items, kp);
}
void g_gm_rotate(hipStream_t s, int k, const int *kp, int m, const GGm *items, int nitems)
{
    if (nitems > 0) ncg_launch(kg_gm_rotate, dim3(nitems), dim3(64), 0, s, k, kp, m, items);
}
void g_gm_solve(hipStream_t s, int m, const GGm *items, int nitems)
{
    if (nitems > 0) ncg_launch(kg_gm_solve, dim3(nitems), dim3(64), 0, s, m, items);
}
void g_gm_update(hipStream_t s, int m, int to_z, const GGm *items, const int *start, int nitems, int blocks)
{
    if (blocks > 0) ncg_launch(kg_gm_update, dim3(blocks), dim3(256), 0, s, m, to_z, items, start, nitems);
}

void g_gm_clear(hipStream_t s, const GGm *items, int nitems)
{
    if (nitems > 0) ncg_launch(kg_gm_clear, dim3(blocks_for(nitems, 64)), dim3(64), 0, s, items, nitems);
}
void g_gm_flags(hipStream_t s, const GGm *items, int nitems, int *out, int *kp, int depth)
{
    if (nitems > 0 && nitems <= 64) ncg_launch(kg_gm_flags, dim3(1), dim3(64), 0, s, items, nitems, out, kp, depth);
}
void g_gm_stats(hipStream_t s, const GGm *items, int nitems, int *its, double *rel)
{
    if (nitems > 0) ncg_launch(kg_gm_stats, dim3(nitems), dim3(64), 0, s, items, its, rel);
}
void g_norms(hipStream_t s, const GNorm *items, const int *start, int nitems, int blocks, double *out)
{
    if (blocks <= 0) return;
    ncg_launch(kg_norm_partial, dim3(blocks), dim3(256), 0, s, items, start, nitems);
    ncg_launch(kg_norm_reduce, dim3(nitems), dim3(256), 0, s, items, out);
}

}  // namespace ncgk
