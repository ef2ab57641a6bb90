// Host-side launchers of the libncg kernels (implemented in ncg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

namespace ncgk {

// While a recorder is set on the calling thread, the launchers below add their kernels to
// recorder->graph as a chain of kernel nodes (each after the previous one) instead of
// queuing them on a stream; err holds the first failure.  nullptr: launch again.
struct GraphRecorder {
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
    int nodes = 0;
};
void set_graph_recorder(GraphRecorder *r);

// device element arrays (structure of arrays of length n)
struct DevElements {
    int n;
    const double *X;       // [12][n]
    const int8_t *nv;      // [n]
    const double *cen;     // [3][n]
    const double *nrm;     // [3][n]
    const double *area;    // [n]
    const double *jac;     // [n]
    const int8_t *kind;    // [n]
};

struct BMParams {
    double k;        // wave number
    double hk;       // h * k
    double tau;      // +1 exterior, -1 interior
    double beta_im;  // Burton-Miller beta = i * beta_im (h/k exterior, 0 interior)
};

hipError_t set_gauss_tables(const double *x, const double *w);
void element_geometry(hipStream_t s, int n, const double *X, const int8_t *nv, double *cen, double *nrm, double *area,
                      double *jac);
void near_count(hipStream_t s, const DevElements &D, int *count);
void near_fill(hipStream_t s, const DevElements &D, const long long *offset, int *rows, int *cols);

// evals (quadrature launchers): nullptr, or a counter that the launch adds its number of
// Green's-function evaluations to (the measurement pass of a plan)
void assemble_far(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, void *A, int mixed,
                  int rows_per_block, unsigned long long *evals);
void assemble_near(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const int *rows,
                   const int *cols, long long npairs, void *A, int mixed, int *err, unsigned long long *evals = nullptr);
void rhs_sources(hipStream_t s, const DevElements &D, const BMParams &b, const int *src, int nsrc, const void *bcval,
                 int nrhs, void *rhs_src, int *err, const int *elem_cluster = nullptr,
                 const unsigned char *cluster_near = nullptr, int nclusters = 0);
void assemble_diag(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const void *bcval,
                   int nrhs, const double *ps, int nps, const double *pw, int npw, const void *rhs_src, int has_src,
                   void *A, int mixed, void *rhs, const long long *diag_index = nullptr,
                   unsigned long long *evals = nullptr);

// Dense operator applied to up to 16 load cases per pass on the matrix cores: x^ (the load
// cases as a real 2n x 32 operand), the product (column splits of partial sums) and the sum
// of the splits.  f64: FP64 accumulation (A in FP32 or FP64), else FP32 (A in FP32).
struct MfmaPlan {
    int f64 = 0, tiles = 0, splits = 0, per = 0;
    size_t xh_elems = 0, part_elems = 0;     // scratch, in elements of the accumulation type
};
// splits: column splits of the product (0: from the size)
MfmaPlan mfma_plan(int n, int f64, int splits = 0);
void mfma_xhat(hipStream_t s, const MfmaPlan &m, int n, int nr, const void *x, void *xh);
void mfma_product(hipStream_t s, const MfmaPlan &m, const void *A, int a_f32, int n, const void *xh, void *part);
void mfma_sum(hipStream_t s, const MfmaPlan &m, int n, int nr, const void *part, void *y);
void matvec(hipStream_t s, const void *A, int mixed, int acc64, int n, int nrhs, const void *x, const double *d,
            void *y);
int dot_blocks(int n);
void dots(hipStream_t s, int n, int nrhs, int ndots, const void *const *u, const void *const *v, int conj_mask,
          void *partial, void *out);
void cgs_alpha(hipStream_t s, int nrhs, int first, const void *dots, void *sc, const int *active);
void cgs_beta(hipStream_t s, int nrhs, const void *dots, void *sc, const int *active);
void cgs_update1(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *u, const void *v,
                 void *q, void *uq);
void cgs_update2(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *uq, const void *auq,
                 void *x, void *r);
void cgs_update3(hipStream_t s, int n, int nrhs, const void *sc, const int *active, const void *r, const void *q,
                 void *u, void *p);
void axpby(hipStream_t s, int n, int nrhs, const void *coef, const void *x, size_t ldx, void *y, size_t ldy);
void scale_vec(hipStream_t s, int n, int nrhs, const double *d, void *x);
void row_scaling(hipStream_t s, const void *A, int mixed, int n, double *d);
void row_scaling_csr(hipStream_t s, int n, const long long *ptr, const void *val, int mixed, double *d);
// h[r][k] = <V_k, w_r>; part: scratch of part_cap complex values for slice partials
void multidot(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *w, void *h, void *part,
              size_t part_cap);
// one classical Gram-Schmidt pass of GMRES: h[r][k] = <V_k[r], w[r]>, w -= V h (h to the
// device buffer h, [r][kcount]); with npart also the partial sums of |w|^2 per block
void gs_pass(hipStream_t s, int n, int nrhs, int kcount, const void *V, void *w, void *h, void *part, size_t part_cap,
             double *npart);
// v = w / ||w|| from gs_pass's partial sums; the squared norm also goes to nout[r][0]
void gs_normalize(hipStream_t s, int n, int nrhs, const double *npart, const void *w, void *v, void *nout);
void multiaxpy(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *h, void *w, int add);
// incomplete LU of the row-scaled operator (dense A when ptr == nullptr, else the
// near-field CSR ptr/ncol/A); pattern CSR rp/col/val, level schedules from the host
void ilu_count(hipStream_t s, const void *A, int n, const long long *ptr, const int *ncol, const double *d, double thr,
               int *cnt);
void ilu_fill(hipStream_t s, const void *A, int n, const long long *ptr, const int *ncol, const double *d, double thr,
              const int *rp, int *col, void *val);
// upd from uoff[i]: the pivot positions diag[k] of row i's L lower entries (k their
// columns), then for lower entry a and entry e (R entries) the position of u_kj (j = col of
// entry e) in row k's upper part or -1, at L + a R + e
void ilu_factor(hipStream_t s, int nlev, const int *lev_ptr, const int *lev_rows, const int *rp, const int *col,
                const int *diag, const long long *uoff, const int *upd, void *val, int *err);
// one level-scheduled triangular sweep of the incomplete LU: position u of the schedule is
// row row[u] with entries c/v[off[u] .. off[u+1]) (CSR order), the first kIluB of them
// also padded in cd/vd[u * kIluB + q]; d = U's diagonal per position (backward sweep)
constexpr int kIluB = 8;
// rows of a batch of the triangular sweeps = threads of the sweep's work-group, and the
// positions of the LDS ring that keeps the latest solved values of a sweep
constexpr int kIluRows = 256, kIluRing = 8192;
// One triangular sweep's schedule: positions u = 0..n-1 (row row[u]) in batches of
// consecutive levels, bat_ptr[b]..bat_ptr[b+1] (at most kIluRows positions, or one wider
// level), bat_m[b] levels per batch and sub[u] the level of u inside its batch.  The first
// kIluB entries of every position are also in cd/vd[u][kIluB]: cd < 0 is the ring slot
// -1 - cd (position mod kIluRing) of a value the ring still holds, cd >= 0 a column read
// from memory.  bat_flag[b]: bit 0, the batch ends with its stores made visible to the
// work-group; bit 1, some row of the batch reads a padded entry from memory.
struct IluSweep {
    int nbat = 0;
    const int *bat_ptr = nullptr, *bat_m = nullptr, *row = nullptr, *off = nullptr, *c = nullptr, *cd = nullptr;
    const unsigned char *sub = nullptr, *bat_flag = nullptr;
    const void *v = nullptr, *vd = nullptr, *d = nullptr;
};
// out[t] = val[map[t]] (0 where map[t] < 0)
void ilu_pack(hipStream_t s, int cnt, const int *map, const void *val, void *out);
void ilu_solve(hipStream_t s, int n, int nrhs, const IluSweep &L, const IluSweep &U, const void *x, void *y);
void lu_factor(hipStream_t s, void *A, int n);
void lu_solve(hipStream_t s, const void *A, int n, int nrhs, void *b);
void field_prep(hipStream_t s, int n, int nrhs, const int8_t *kind, const double *adm, double tau, const void *bcval,
                const void *x, void *pot, void *vel);
void field(hipStream_t s, const DevElements &D, const BMParams &b, const double *pts, int m, int nrhs, const void *pot,
           const void *vel, const double *ps, int nps, const double *pw, int npw, void *phi, void *grad, int fp32,
           int *err, const int *pt_ev = nullptr, const int *el_ptr = nullptr, const int *el_idx = nullptr,
           const void *fphi = nullptr, const void *fgrad = nullptr, unsigned long long *evals = nullptr);

// fast multipole operators (layouts in ncg_kernels.hip)
// coef: scratch of nfan * (3 + 2 L) doubles (series coefficients per interaction)
void fmm_D(hipStream_t s, int L, int P, double k, double h, const double *ccen, int ncl, const int *fan_row,
           const int *fan_idx, int nfan, const double *u, void *Dv, int mixed, double *coef);
void fmm_T(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, int P, const double *ccen,
           const int *elem_cl, const int *elem_pos, const double *u, void *Tv, int mixed);
void fmm_Tvec_bc(hipStream_t s, const DevElements &D, const BMParams &b, int P, int ncl, const double *ccen,
                 const int *cl_ptr, const int *cl_elem, const double *u, const void *bcval, int nrhs, void *tv);
// S stored per cluster column-major ([point][local element] from cl_ptr[c] * P)
void fmm_S(hipStream_t s, const DevElements &D, const BMParams &b, int P, const double *ccen, const int *elem_cl,
           const int *cl_ptr, const int *cl_elem, const double *u, const double *w, void *Sv, int mixed);
// subdivided pairs of the near field: CSR positions (once per structure) and values
void adj_positions(hipStream_t s, long long npairs, const int *rows, const int *cols, const int *elem_cl,
                   const int *elem_pos, const int *cl_ptr, const int *nn_ptr, const int *nn_idx, const int *nn_off,
                   const long long *near_ptr, long long *pos);
void fmm_near_adj(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const int *rows,
                  const int *cols, const long long *pos, long long npairs, void *vals, int mixed, int *err,
                  unsigned long long *evals = nullptr);
// block-Jacobi preconditioner: blocks of at most 64 consecutive elements of a finest
// cluster, gathered from the near-field CSR and inverted in place (complex FP64)
void bj_setup(hipStream_t s, int n, int nblk, const int *elem_blk, const int *elem_pos, const int *blk_cl,
              const int *blk_b0, const int *blk_m, const int *cl_ptr, const long long *near_ptr, const int *self_off,
              const void *nval, int mixed, const long long *boff, void *blk, int *err);
// the same for a dense operator A (n x n row-major): blocks of consecutive unknowns
void bj_setup_dense(hipStream_t s, int n, int nblk, const int *elem_blk, const int *blk_b0, const int *blk_m,
                    const void *A, int mixed, const long long *boff, void *blk, int *err);
// mixed: the blocks are inverted and stored in FP32 (FP64 otherwise)
void bj_apply(hipStream_t s, int n, int nrhs, int nblk, const int *blk_b0, const int *blk_m, const int *cl_elem,
              const long long *boff, const void *minv, int mixed, const void *x, void *y);
// v = w / ||w|| per load case, the norm read from a dots() result
void scale_norm(hipStream_t s, int n, int nrhs, const void *dts, const void *w, void *v);
void near_csr(hipStream_t s, int n, const int *elem_cl, const int *cl_ptr, const int *cl_elem, const int *nn_ptr,
              const int *nn_idx, const long long *ptr, int *rows, int *cols, long long *diag);
// Row chunks of the FMM kernels: chunk q holds m[q] <= kChunkRows rows of cluster cl[q], starting
// at position r0[q] of the level's element list (cl_elem).
constexpr int kChunkRows = 16;
struct RowChunks {
    int n = 0;
    const int *cl = nullptr, *r0 = nullptr, *m = nullptr;
};
// Near-field row chunks with their addresses resolved on the host: the rows of chunk q share
// the column list col[cb[q] .. cb[q] + w[q]); row slot a < kNearRows holds row
// row[kNearRows q + a] (-1: unused) whose values start at off[kNearRows q + a] (cb[q]
// for an unused slot).  A finest cluster of m elements gives ceil(m / kNearRows) chunks
// of balanced sizes; 4 waves of kNearRowsPerWave rows each work on one chunk.
constexpr int kNearRowsPerWave = 3;
constexpr int kNearRows = 4 * kNearRowsPerWave;
struct NearChunks {
    int n = 0;
    const long long *cb = nullptr, *off = nullptr;
    const int *w = nullptr, *row = nullptr;
};
// near field of the finest level (CSR whose rows in one cluster share their columns)
void near_spmv(hipStream_t s, int n, int nrhs, const NearChunks &nc, const int *col, int wmax, const void *val,
               int mixed, const void *x, void *y);
// far field of one level: f32 selects FP32 T, D, S and FP32 storage of t and s
// tw_f64: t is FP64 whatever f32 says (signatures for the evaluation grid and the
// right-hand side's sources)
void fmm_apply_T(hipStream_t s, int n, int ncl, int P, int nrhs, const int *cl_ptr, const int *cl_elem, const void *Tv,
                 int f32, int tw_f64, const void *x, void *tw);
void fmm_apply_D(hipStream_t s, int ncl, int P, int nrhs, const int *fan_ptr, const int *fan_idx, const void *Dv,
                 int f32, int tw_f64, const void *tw, void *sw);
// tl: tiles of at most 64 elements of one cluster (RowChunks with m <= 64)
void fmm_apply_S(hipStream_t s, int n, int ncl, int P, int nrhs, const RowChunks &tl, const int *cl_ptr,
                 const int *cl_elem, const void *Sv, int f32, const void *sw, void *y);
// regular near-field entries by (row chunk, 256 columns); the subdivided pairs are left to
// fmm_near_adj and the diagonal to assemble_diag.  wmax = longest row.  rows_mode 0: a
// thread per column with its geometry in registers; 1: a wave per row, the columns'
// geometry in LDS (the same values)
void fmm_near(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const RowChunks &rc,
              const int *cl_ptr, const int *cl_elem, const long long *ptr, const int *col, int wmax, void *vals,
              int mixed, unsigned long long *evals = nullptr, int rows_mode = 0);
void fmm_Tvec_sol(hipStream_t s, const DevElements &D, const BMParams &b, int P, int ncl, const double *ccen,
                  const int *cl_ptr, const int *cl_elem, const double *u, const void *pot, const void *vel, int nrhs,
                  void *tv, int fp32, int tv32 = 0);
// evaluation clusters: per far pair the direction and h_n (tab [npair][4 + 2L] doubles),
// then z[r][ev][p] = sum over the pairs of mu * w * tv.  fp32: the signature's phases and
// the series in FP32; tv32 (with fp32): tv stored as FP32 complex, else FP64
void fmm_eval_hankel(hipStream_t s, int L, double k, double h, const double *ecen, const double *ccen,
                     const int *pair_ev, const int *pair_cl, int npair, double *tab);
void fmm_eval_z(hipStream_t s, int nev, int L, int P, double h, const int *ev_ptr, const int *pair_cl,
                const double *tab, const double *u, const double *w, const void *tv, int ncl, int nrhs, void *z,
                int fp32, int tv32 = 0);
void fmm_eval_far(hipStream_t s, int m, int P, int nev, double hk, const double *pts, const int *pt_ev,
                  const double *ecen, const double *u, int nrhs, const void *z, void *fphi, void *fgrad);

// ---------------------------------------------------------------------------------------
// Grouped launches: one launch applies the same operation to several problems (or problem
// levels) of a group solved in lockstep.  `items` is a device array with one entry per
// problem (or level) and `start` (device, nitems + 1 ints) the prefix sums of their
// work-group counts; a work-group finds its item by a scan of `start` (items are few) and
// does nothing while the item's flag *skip (device int, may be null) is nonzero, i.e. once
// that problem's Krylov solve has converged.  Each item's arithmetic is the single-problem
// kernel's, work-group for work-group, so results do not depend on the grouping.
//
// NCG_GLOBAL marks the descriptors' pointers as pointers into global memory in the device
// pass of the kernels' unit (NCG_KERNELS_TU): a pointer read from memory is otherwise
// generic to the compiler, and loads through it are flat loads, which are slower and are
// waited for together with LDS operations.  Every other pass sees plain pointers; the
// layout is the same (64-bit pointers).
#if defined(__HIP_DEVICE_COMPILE__) && defined(NCG_KERNELS_TU)
#define NCG_GLOBAL __attribute__((address_space(1)))
#else
#define NCG_GLOBAL
#endif
struct GNear {      // near field of the finest level (NearChunks, col, values)
    int n, nch, ncl;                        // unknowns, row chunks, finest clusters
    const NCG_GLOBAL int *cptr;             // cluster c owns chunks cptr[c] .. cptr[c + 1]
    const NCG_GLOBAL long long *cb, *off;
    const NCG_GLOBAL int *w, *row, *col;
    const NCG_GLOBAL void *val;
    const NCG_GLOBAL void *x;
    NCG_GLOBAL void *y;
    const NCG_GLOBAL int *skip;
};
struct GLevel {     // one FMM level: T (x -> t), D (t -> s) and S (s -> y += ) of a problem
    int n, ncl, P, ntl;
    const NCG_GLOBAL int *cl_ptr, *cl_elem, *fan_ptr, *fan_idx, *tl_cl, *tl_a0, *tl_m;
    const NCG_GLOBAL void *Tv, *Dv, *Sv;
    const NCG_GLOBAL void *x;
    NCG_GLOBAL void *tw, *sw, *y;
    const NCG_GLOBAL int *skip;
};
struct GBj {        // block-Jacobi blocks (FMM finest clusters or dense diagonal blocks)
    int n, nr, nblk;
    const NCG_GLOBAL int *b0, *m, *elem;
    const NCG_GLOBAL long long *boff;
    const NCG_GLOBAL void *minv;
    const NCG_GLOBAL void *x;               // input; with a step counter kp: x + *kp * xk doubles
    long long xk;
    NCG_GLOBAL void *y;
    const NCG_GLOBAL int *skip;
};
// GMRES(m) state of one problem of a group, on the device: the Hessenberg matrices, Givens
// rotations, right-hand sides g and solutions y of every load case, the Arnoldi work
// arrays, and per load case the scalars [beta, |b|, relative residual, -] and the
// counters [k, done, iterations, -].  skip (= the problem's flag of the grouped launches)
// is set once every load case has converged.
struct GGm {
    int n, nrhs, nsl, slice, gsb;
    const NCG_GLOBAL void *b;               // right-hand sides [nrhs][n] (the Krylov system's)
    NCG_GLOBAL void *x;                     // solution [nrhs][n]
    NCG_GLOBAL void *V;                     // basis [m + 1][nrhs][n]
    NCG_GLOBAL void *W;                     // new Arnoldi vector / residual [nrhs][n]
    NCG_GLOBAL void *Z, *T;                 // preconditioned vector and apply output [nrhs][n]
    NCG_GLOBAL void *part;                  // multidot slice partials
    NCG_GLOBAL void *h;                     // [2][nrhs][m + 1] projections of the two passes, then [nrhs][3] norms
    NCG_GLOBAL double *npart;               // [nrhs][gsb] partial squared norms
    NCG_GLOBAL void *H, *cs, *sn, *g, *yv;  // [nrhs][(m+1) m], [nrhs][m], [nrhs][m], [nrhs][m+1], [nrhs][m]
    NCG_GLOBAL double *sc;                  // [nrhs][4]
    NCG_GLOBAL int *ic;                     // [nrhs][4]
    NCG_GLOBAL int *skip;
    double tol;                             // relative residual target of every load case
};
// element-wise vector operations of a group: per item y = a (op kVecCopy), y = a - b
// (kVecSub), y += a (kVecAdd), y = 0 (kVecZero); len = complex entries
enum { kVecCopy = 0, kVecSub = 1, kVecAdd = 2, kVecZero = 3 };
struct GVec {
    long long len;
    const NCG_GLOBAL void *a, *b;           // with a step counter kp: a + *kp * ak doubles
    long long ak;
    NCG_GLOBAL void *y;
    const NCG_GLOBAL int *skip;
};
// launchers (blocks = start[nitems], the prefix the host uploaded)
// mode 0: a work-group per row chunk; 1: per finest cluster (x gathered once for its chunks);
// 2, 3: persistent work-groups over the row chunks, 2 from per-XCD work queues (`queue`: 288
// device ints, zero between launches; the launch leaves them zero), 3 by a static stride
void g_near_spmv(hipStream_t s, int mixed, int mode, int nr, int seg, const GNear *items, const int *start,
                 int nitems, int blocks, int *queue);
// T and D of every item (each writes its own t or s); S adds into y, so a launch may hold at
// most one level of a problem (the host launches one level index at a time).  sig64: t and
// s in FP64 whatever the values' precision (the refinement residual).  nw: waves per
// work-group (1, 2, 4, 8) splitting each item's sum, met in an LDS tree.  T and D take
// ncl * ceil(P / (64 * pts)) work-groups per item, pts = 2 sphere points per lane for FP32
// values (P is even), 1 for FP64; S one per row tile (GLevel::ntl).
void g_apply_T(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks);
void g_apply_D(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks);
void g_apply_S(hipStream_t s, int f32, int sig64, int nw, int nr, const GLevel *items, const int *start, int nitems,
               int blocks);
// kp (device int, or null): the group's Arnoldi step counter, which offsets each item's
// input by its stride (GBj::xk, GVec::ak) so one launch serves every step
void g_bj_apply(hipStream_t s, int f32, int nr, const GBj *items, const int *start, int nitems, int blocks,
                const int *kp = nullptr);
void g_vec(hipStream_t s, int op, const GVec *items, const int *start, int nitems, int blocks,
           const int *kp = nullptr);
// GMRES(m) of a group in lockstep; at Arnoldi step k the basis holds kcount = k + 1 vectors.
// The step kernels read k from the group's device counter kp (set to 0 by g_gm_start,
// advanced by g_gm_flags at the end of each step), so the launches of an Arnoldi step do
// not depend on k and can be captured once and replayed as a HIP graph; where kp is null
// the k (kcount) argument is used.
//   g_multidot: slice partials of <V_j, W> for j < kcount; per item kMdLanes * nrhs * nsl
//   work-groups (start holds the prefix of nrhs * nsl), lane group jj takes j = jj mod kMdLanes.
//   g_gs_update pass 0/1: projections (multidot partials) summed, w -= V h;
//   pass 1 also leaves the partial norms.  kcount = 0 with pass 1: partial norms of W only.
//   g_gs_normalize: V[k + 1] = W / |W| (k = -1: V[0]), squared norms to h's norm slots.
//   g_gm_start: beta, residual, done flags of a cycle (first: |b| = beta, counters reset);
//   sets *kp = 0.
//   g_gm_rotate: Hessenberg column k, Givens rotations, residual estimate, done flags.
//   g_gm_solve: y = H^-1 g per load case (zero past its k); g_gm_update: Z = V y (to_z) or
//   x += V y.
constexpr int kMdLanes = 16;
void g_multidot(hipStream_t s, const int *kp, const GGm *items, const int *start, int nitems, int blocks);
void g_gs_update(hipStream_t s, int kcount, const int *kp, int pass, int m, const GGm *items, const int *start,
                 int nitems, int blocks);
void g_gs_normalize(hipStream_t s, int k, const int *kp, int m, const GGm *items, const int *start, int nitems,
                    int blocks);
void g_gm_start(hipStream_t s, int m, int first, const GGm *items, int nitems, int *kp);
void g_gm_rotate(hipStream_t s, int k, const int *kp, int m, const GGm *items, int nitems);
void g_gm_solve(hipStream_t s, int m, const GGm *items, int nitems);
void g_gm_update(hipStream_t s, int m, int to_z, const GGm *items, const int *start, int nitems, int blocks);
// skip flags cleared (start of a solve) or gathered into out[nitems] (nitems <= 64): with a
// step counter kp into out + (*kp % depth) * 64, and *kp advanced by one after; per item and
// load case the iterations and residual estimates into its/rel[item][64]
void g_gm_clear(hipStream_t s, const GGm *items, int nitems);
void g_gm_flags(hipStream_t s, const GGm *items, int nitems, int *out, int *kp = nullptr, int depth = 1);
void g_gm_stats(hipStream_t s, const GGm *items, int nitems, int *its, double *rel);
// squared norms |v_r|^2 of every item's load cases into out[item][64]; per item gsb * nrhs
// work-groups (start: their prefix), npart [nrhs][gsb] scratch
struct GNorm {
    int n, nrhs, gsb;
    const NCG_GLOBAL void *v;
    NCG_GLOBAL double *npart;
    const NCG_GLOBAL int *skip;
};
void g_norms(hipStream_t s, const GNorm *items, const int *start, int nitems, int blocks, double *out);

}  // namespace ncgk
