// libncg C ABI (include/ncg.h): problem handles, device buffers, and the host-side
// orchestration of assembly, Krylov solvers and field evaluation.  All arithmetic on
// the solve path runs in the kernels of ncg_kernels.hip; this file only sequences
// launches, moves the few scalars the solvers branch on, and times phases with HIP
// events on the problem's stream.
#include "ncg.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <string>
#include <vector>

#include "ncg_fmm.hpp"
#include "ncg_internal.hpp"
#include "ncg_kernels.hpp"

#include <array>
#include <map>
#include <mutex>

using cd = std::complex<double>;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define NCG_HIP(call)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (call);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return fail(NCG_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));                   \
    } while (0)

// Gauss-Legendre nodes/weights by Newton iteration on P_n (the rules NumCalc tabulates
// in src/NC_IntegrationConstants.h:24-1900), packed rule n at n(n-1)/2, n = 1..8.
void gauss_tables(double *x, double *w)
{
    for (int n = 1; n <= 8; ++n) {
        double *xn = x + n * (n - 1) / 2, *wn = w + n * (n - 1) / 2;
        for (int i = 0; i < (n + 1) / 2; ++i) {
            double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = 0.0;
                for (int k = 1; k <= n; ++k) {
                    double p2 = p1;
                    p1 = p0;
                    p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
                }
                pp = n * (z * p0 - p1) / (z * z - 1.0);
                double dz = p0 / pp;
                z -= dz;
                if (fabs(dz) < 1e-16) break;
            }
            double p0 = 1.0, p1 = 0.0;
            for (int k = 1; k <= n; ++k) {
                double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
            }
            pp = n * (z * p0 - p1) / (z * z - 1.0);
            xn[i] = -z;
            xn[n - 1 - i] = z;
            wn[i] = wn[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
        }
        if (n % 2) xn[n / 2] = 0.0;
    }
}

template <typename T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(NCG_E_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

// device allocations released together
struct Pool {
    std::vector<void *> owned;
    void release()
    {
        for (void *q : owned) hipFree(q);
        owned.clear();
    }
    template <typename T>
    int alloc(T **ptr, size_t count)
    {
        *ptr = nullptr;
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void **)ptr, sizeof(T) * count);
        if (e != hipSuccess) return fail(NCG_E_MEMORY, std::string("hipMalloc (FMM): ") + hipGetErrorString(e));
        owned.push_back((void *)*ptr);
        return 0;
    }
};

// grow-only device buffer (kept across frequencies)
struct Grow {
    void *p = nullptr;
    size_t cap = 0;
    // grows by at least half of the current capacity, so a band of rising frequencies
    // reallocates (and synchronises the device in hipFree) only a few times
    int ensure(size_t bytes)
    {
        if (bytes <= cap && p) return 0;
        const size_t want = std::max<size_t>(std::max<size_t>(bytes, 16), cap + cap / 2);
        hipFree(p);
        p = nullptr;
        cap = 0;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            if (want > bytes && hipMalloc(&p, std::max<size_t>(bytes, 16)) == hipSuccess) {
                cap = std::max<size_t>(bytes, 16);
                return 0;
            }
            return fail(NCG_E_MEMORY, std::string("hipMalloc (FMM): ") + hipGetErrorString(e));
        }
        cap = want;
        return 0;
    }
    void release()
    {
        hipFree(p);
        p = nullptr;
        cap = 0;
    }
    double *d() const { return (double *)p; }
};

// device arrays of one FMM level (layouts in ncg_kernels.hip).  The cluster structure
// is kept while the plan's structure is unchanged; the frequency-dependent arrays (sphere
// points, T, D, S, work vectors) are grow-only buffers.
struct DevLevel {
    int ncl = 0, nfan = 0;
    double *ccen = nullptr;
    int *cl_ptr = nullptr, *cl_elem = nullptr, *elem_cl = nullptr, *elem_pos = nullptr;
    int *fan_ptr = nullptr, *fan_idx = nullptr, *fan_row = nullptr;
    int nch = 0;                      // row chunks: <= 16 elements of one cluster each
    int *ch_cl = nullptr, *ch_r0 = nullptr, *ch_m = nullptr;
    int ntl = 0;                      // tiles: <= 64 elements of one cluster each (local first)
    int *tl_cl = nullptr, *tl_a0 = nullptr, *tl_m = nullptr;
    int P = 0, L = 0;
    int f32 = 0;                      // T, D, S (and t, s) stored in FP32 (mixed form, regular levels only)
    double dmin = 1e300;              // shortest centre distance of an interaction pair
    Grow u, w, T, D, S, tw, sw, tv;   // tw, sw, tv: [nrhs][ncl*P]
    ncgk::RowChunks chunks() const
    {
        ncgk::RowChunks r;
        r.n = nch;
        r.cl = ch_cl;
        r.r0 = ch_r0;
        r.m = ch_m;
        return r;
    }
    ncgk::RowChunks tiles() const
    {
        ncgk::RowChunks r;
        r.n = ntl;
        r.cl = tl_cl;
        r.r0 = tl_a0;
        r.m = tl_m;
        return r;
    }
    void release()
    {
        for (Grow *g : {&u, &w, &T, &D, &S, &tw, &sw, &tv}) g->release();
    }
};

// device arrays of the FMM operator
struct DevFmm {
    unsigned long long sig = 0;   // structure signature of the plan the arrays belong to
    Pool st;
    std::vector<DevLevel> lev;
    long long nnz = 0;
    long long *near_ptr = nullptr, *diag_pos = nullptr;
    int *near_col = nullptr;
    int near_wmax = 0;                 // longest near-field row
    // near-field chunk table (ncgk::NearChunks): per chunk the column base and width, per
    // row slot the row and its value offset
    long long *nd_cb = nullptr, *nd_off = nullptr;
    int *nd_w = nullptr, *nd_row = nullptr;
    int nd_n = 0;
    int *nd_cptr = nullptr;            // finest cluster c owns chunks nd_cptr[c] .. nd_cptr[c + 1]
    ncgk::NearChunks near_chunks() const
    {
        ncgk::NearChunks r;
        r.n = nd_n;
        r.cb = nd_cb;
        r.off = nd_off;
        r.w = nd_w;
        r.row = nd_row;
        return r;
    }
    int *elem_cl_fine = nullptr;
    unsigned char *cl_near = nullptr;
    int ncl_fine = 0;
    long long *adj_pos = nullptr;   // CSR position of each subdivided pair (or -1)
    Grow near_val, xs;            // xs: scaled copy of the input vector
    // block-Jacobi preconditioner: block layout (structure) and inverses (values)
    int *self_off = nullptr, *elem_blk = nullptr, *blk_cl = nullptr, *blk_b0 = nullptr, *blk_m = nullptr;
    long long *boff = nullptr, bj_total = 0;
    int nblk = 0;
    Grow bj_blk;
    int bj_ready = 0, bj_f32 = 0;   // bj_f32: inverses in FP32 (mixed form)
    // evaluation-grid pairs (rebuilt when the pair lists change)
    unsigned long long ev_sig = 0, ev_struct = 0;
    int ev_pure = 0;
    double ev_dmin_far = 0.0;
    Pool evp;
    int *pt_ev = nullptr, *ev_ptr = nullptr, *pair_ev = nullptr, *pair_cl = nullptr, *el_ptr = nullptr,
        *el_idx = nullptr;
    double *ecen = nullptr;
    int npair = 0, nev = 0;
    Grow tab, z, fphi, fgrad;
    Grow dcoef;                   // translation series coefficients while D is built
    void release_structure()
    {
        for (DevLevel &L : lev) L.release();
        lev.clear();
        st.release();
        near_val.release();
        bj_ready = 0;
        sig = 0;
        nnz = 0;
        evp.release();
        ev_sig = 0;
    }
    void release()
    {
        release_structure();
        for (Grow *g : {&xs, &tab, &z, &fphi, &fgrad, &bj_blk, &dcoef}) g->release();
    }
};

// FNV-1a over integer sequences: structure signatures of plans and pair lists
// 64-bit digest for change detection: FNV-1a offset, one multiply-xorshift step per word
struct Fnv {
    unsigned long long h = 1469598103934665603ull;
    void add(long long v)
    {
        h ^= (unsigned long long)v;
        h *= 0x100000001B3ull;
        h ^= h >> 29;
    }
    void add(const std::vector<int> &v)
    {
        add((long long)v.size());
        for (int x : v) add((long long)x);
    }
};

// one timed launch: kernel id, event pair slot, algorithmic bytes, and its algorithmic
// FLOPs: given, or (quad >= 0) the counted evaluations of quadrature kind quad times that
// kind's FLOPs per evaluation
struct KernelRec {
    int id;
    size_t slot;
    double bytes;
    int quad;
    double flops;
};

// Quadrature kinds whose Green's-function evaluations are counted on the device in the
// measurement pass of a plan (the counts depend on the geometry and the plan only): dense
// regular and subdivided pairs, singular diagonal, FMM regular and subdivided near pairs,
// evaluation-grid field.
enum { kQFar = 0, kQDenseNear, kQDiag, kQFmmNear, kQFmmAdj, kQField, kQuadKinds };
// FLOPs per Green's-function evaluation (G, dG/dn_y, dG/dn_x, d2G/dn_x dn_y; FMA = 2,
// sqrt, reciprocal, sin and cos = 1 each), and per plane-wave evaluation of the FMM
// aggregation T (per element Gauss point and sphere point), per disaggregation entry S,
// and per term of the translation series D
constexpr double kGreenFlops = 60.0, kTFlops = 16.0, kSFlops = 20.0, kDTermFlops = 7.0;

// device state of the incomplete LU preconditioner (pattern CSR and level schedules)
struct DevIlu {
    int nnz = 0, nlevL = 0, nlevU = 0;
    int *rp = nullptr, *col = nullptr, *diag = nullptr;
    int *levL_ptr = nullptr, *levL_rows = nullptr, *levU_ptr = nullptr, *levU_rows = nullptr;
    // batches of consecutive levels for the triangular solves
    int nbatL = 0, nbatU = 0;
    int *batL_ptr = nullptr, *batL_m = nullptr, *batU_ptr = nullptr, *batU_m = nullptr;
    unsigned char *subL = nullptr, *subU = nullptr, *flagL = nullptr, *flagU = nullptr;
    double *val = nullptr;
    // level-ordered factors for the triangular solves: offsets, columns, CSR positions
    int *Loff = nullptr, *Lc = nullptr, *Lmap = nullptr, *Uoff = nullptr, *Uc = nullptr, *Umap = nullptr,
        *Udmap = nullptr;
    double *Lv = nullptr, *Uv = nullptr, *Ud = nullptr;
    // the first kIluB entries of every position again, padded: [n][kIluB]
    int *Lcd = nullptr, *Lmapd = nullptr, *Ucd = nullptr, *Umapd = nullptr;
    double *Lvd = nullptr, *Uvd = nullptr;
    int nL = 0, nU = 0;   // strictly lower / upper entries
    std::vector<void *> owned;
    void release()
    {
        for (void *q : owned) hipFree(q);
        owned.clear();
        nnz = nlevL = nlevU = 0;
    }
    template <typename T>
    int alloc(T **ptr, size_t count)
    {
        *ptr = nullptr;
        hipError_t e = hipMalloc((void **)ptr, sizeof(T) * std::max<size_t>(count, 1));
        if (e != hipSuccess) return fail(NCG_E_MEMORY, std::string("hipMalloc (ILU): ") + hipGetErrorString(e));
        owned.push_back((void *)*ptr);
        return 0;
    }
};

}  // namespace

struct ncg_problem {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0, m = 0;
    double h = 1.0;
    int interior = 0;
    // geometry (device)
    double *X = nullptr, *cen = nullptr, *nrm = nullptr, *area = nullptr, *jac = nullptr, *eval = nullptr;
    int8_t *nv = nullptr, *kind = nullptr;
    std::vector<int8_t> kind_h;
    int all_velocity = 0;   // every element has a velocity boundary condition
    long long nnear = 0;
    int *near_rows = nullptr, *near_cols = nullptr;
    // frequency data
    double k = 0.0;
    int nrhs = 0, precision = NCG_FP64;
    int assembled = 0, solved = 0;
    int asm_pending = 0;      // assembly queued on the stream, its error flags not yet read
    void *A = nullptr;
    size_t A_cap = 0;
    double *adm = nullptr;
    double *bcval = nullptr, *rhs = nullptr, *rhs_src = nullptr, *x = nullptr;
    size_t vec_cap = 0;   // complex entries of each [nrhs][n] vector
    double *ps = nullptr, *pw = nullptr;
    int nps = 0, npw = 0, ps_cap = 0, pw_cap = 0;
    int *src = nullptr;
    int nsrc = 0;
    int *err = nullptr;
    // quadrature evaluation counts: device counters of the measurement pass, the counts of
    // each kind and the structure signature they hold for, the kinds being counted
    unsigned long long *evals = nullptr;
    double q_evals[kQuadKinds] = {0};
    uint64_t q_sig[kQuadKinds] = {0};
    int q_pending = 0;
    // Krylov workspace
    std::vector<double *> work;
    size_t work_cap = 0;
    double *scal = nullptr, *partial = nullptr, *dotout = nullptr, *dvec = nullptr;
    size_t partial_cap = 0;   // complex entries of partial
    double *npart = nullptr;  // [64 load cases][256 blocks] partial norms of the Gram-Schmidt passes
    // GMRES: pinned landing slots for two Arnoldi columns and the events that mark them
    cd *hpin = nullptr;
    size_t hpin_slot = 0;
    hipEvent_t hev[2] = {nullptr, nullptr};
    // block-Jacobi of a dense operator: blocks of at most 64 consecutive unknowns
    Pool dbj;
    int dbj_nblk = 0, dbj_ready = 0, dbj_f32 = 0;
    int *dbj_elem_blk = nullptr, *dbj_b0 = nullptr, *dbj_m = nullptr, *dbj_ident = nullptr;
    long long *dbj_boff = nullptr, dbj_total = 0;
    Grow dbj_blk;
    // dense operator on the matrix cores: X^ operand and column-split partial sums
    Grow mm_xh, mm_part;
    int *active = nullptr;
    double *basis = nullptr;
    size_t basis_cap = 0;
    // timings
    EventPair ev_asm, ev_far, ev_solve, ev_field;
    std::vector<EventPair> ev_mv;
    size_t mv_used = 0;
    double t_asm = 0, t_solve = 0, t_field = 0;
    double far_ms_sum = 0, mv_ms_sum = 0;
    long long far_calls = 0, mv_calls = 0;
    double last_evals = 0;
    // fast multipole state
    ncgf::Mesh fm;
    ncgf::Options fo;
    ncgf::Plan plan;
    DevFmm fmm;
    size_t nrhs_fmm = 0;
    DevIlu ilu;
    Grow out_phi, out_grad;   // evaluation-grid results before the copy to the host
    // grouped Krylov solves (this problem as a member or as the group's leader): GMRES state
    // of the member, and for the leader the launch descriptors (device copy and pinned
    // host staging), pinned landing slots for flags and norms, and the events that order
    // the members' streams around the group's
    Grow gm;
    Grow gdesc;
    ncg_tuning tune{8, 8, 8, 0, 0, 0, 0, 0};   // launch shapes (ncg_autotune); a group uses its leader's
    char *gstage = nullptr;
    size_t gstage_cap = 0;
    void *gpin = nullptr;
    int *queue = nullptr;   // work-queue counters of the persistent launches this problem leads (zero between launches)
    hipEvent_t gev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t join = nullptr;
    // Arnoldi steps of the grouped solves this problem leads replayed as captured HIP graphs
    // (off: launched one by one); executable graphs by the digest of every launch argument
    // they hold, most recently used first; no_events: the apply records no timing events
    // (while a step is being captured)
    int graphs = 1, graph_fail = 0, no_events = 0;
    std::vector<std::pair<unsigned long long, hipGraphExec_t>> gcache;
    long long g_captured = 0, g_replayed = 0, g_eager = 0;   // graphs captured, steps replayed / launched one by one
    // per-kernel timing of the operator apply (off unless requested)
    int ktime = 0;
    std::vector<EventPair> kpool;
    std::vector<KernelRec> kpend;
    double k_ms[64] = {0}, k_bytes[64] = {0}, k_flops[64] = {0};
    long long k_calls[64] = {0};

    ncgk::DevElements elements() const { return {n, X, nv, cen, nrm, area, jac, kind}; }
    ncgk::BMParams bm() const
    {
        ncgk::BMParams b;
        b.k = k;
        b.hk = h * k;
        b.tau = interior ? -1.0 : 1.0;
        b.beta_im = interior ? 0.0 : h / k;
        return b;
    }
};

namespace {

int ensure_vectors(ncg_problem *p, int nrhs)
{
    const size_t need = (size_t)nrhs * p->n;
    if (need <= p->vec_cap) return 0;
    hipFree(p->bcval);
    hipFree(p->rhs);
    hipFree(p->rhs_src);
    hipFree(p->x);
    p->vec_cap = 0;
    int rc;
    if ((rc = dalloc(&p->bcval, 2 * need)) || (rc = dalloc(&p->rhs, 2 * need)) || (rc = dalloc(&p->rhs_src, 2 * need)) ||
        (rc = dalloc(&p->x, 2 * need)))
        return rc;
    p->vec_cap = need;
    return 0;
}

int ensure_work(ncg_problem *p, int count)
{
    const size_t need = (size_t)p->nrhs * p->n;
    if ((int)p->work.size() >= count && p->work_cap >= need) return 0;
    for (double *w : p->work) hipFree(w);
    p->work.assign(count, nullptr);
    for (int i = 0; i < count; ++i) {
        int rc = dalloc(&p->work[i], 2 * need);
        if (rc) return rc;
    }
    p->work_cap = need;
    return 0;
}

void record(ncg_problem *p, hipEvent_t e) { hipEventRecord(e, p->stream); }

double elapsed(const EventPair &ev)
{
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev.a, ev.b) != hipSuccess) return 0.0;
    return ms;
}

// matvec with per-call timing (the Krylov loops' dominant kernel)
// Per-kernel timing of the operator apply (enabled by ncg_set_kernel_timing): HIP events
// around each launch on the problem's stream, with the algorithmic bytes of the launch.
// ids 12.. are the grouped launches of the same operations (kernel id + kGroupedKid)
const char *const kKernelNames[] = {"k_near_spmv<f64>",   "k_near_spmv<f32>",   "k_fmm_apply_T<f64>",
                                    "k_fmm_apply_T<f32>", "k_fmm_apply_D<f64>", "k_fmm_apply_D<f32>",
                                    "k_fmm_apply_S<f64>", "k_fmm_apply_S<f32>", "k_matvec<f64>",
                                    "k_matvec<f32>",      "k_matvec_mfma<f32>", "k_matvec_mfma64",
                                    "kg_near_spmv<f64>",  "kg_near_spmv<f32>",  "kg_apply_T<f64>",
                                    "kg_apply_T<f32>",    "kg_apply_D<f64>",    "kg_apply_D<f32>",
                                    "kg_apply_S<f64>",    "kg_apply_S<f32>",
                                    // quadrature and FMM table kernels of the assembly and the field
                                    "k_fmm_near<f64>",    "k_fmm_near<f32>",    "k_fmm_near_adj",
                                    "k_assemble_diag",    "k_fmm_T<f64>",       "k_fmm_T<f32>",
                                    "k_fmm_S<f64>",       "k_fmm_S<f32>",       "k_fmm_D",
                                    "k_assemble_far<f64>", "k_assemble_far<f32>", "k_assemble_near",
                                    "k_field<f64>",       "k_field<f32>"};
constexpr int kNumKernels = 34, kGroupedKid = 12;
enum {
    kKidFmmNear = 20, kKidFmmAdj = 22, kKidDiag, kKidFmmT, kKidFmmS = 26, kKidFmmD = 28, kKidFar = 29,
    kKidDenseNear = 31, kKidField = 32
};

struct KernelScope {
    ncg_problem *p;
    int id;
    double bytes;
    EventPair *ev = nullptr;
    KernelScope(ncg_problem *p_, int id_, double bytes_, int quad = -1, double flops = 0.0);
    ~KernelScope();
};

// Counter of quadrature kind q for the assembly or field pass about to be queued: the
// device counter when kernel timing is on and kind q has no count for structure sig yet
// (the measurement pass; q_collect reads the counts back), else nullptr.
unsigned long long *q_counter(ncg_problem *p, int q, uint64_t sig);
void q_collect(ncg_problem *p);

// Scratch of an allocation-free apply (ncg_apply_noalloc): the FMM signatures t and s of
// every level, or the matrix-core operand and partial sums of a dense operator, carved from
// one caller-independent device array; no events are recorded.
struct MvScratch {
    char *base = nullptr;
    size_t bytes = 0;
};

void mv(ncg_problem *p, int acc64, const double *x, const double *d, double *y, const MvScratch *na = nullptr)
{
    EventPair *e = nullptr;
    if (!na && !p->no_events) {
        if (p->mv_used == p->ev_mv.size()) {
            EventPair ep;
            hipEventCreate(&ep.a);
            hipEventCreate(&ep.b);
            p->ev_mv.push_back(ep);
        }
        e = &p->ev_mv[p->mv_used++];
        record(p, e->a);
    }
    const int ktime_keep = p->ktime;
    if (na) p->ktime = 0;
    const double n = p->n, nr = p->nrhs, chunks = std::ceil(nr / 4.0);
    if (p->plan.method == 0) {
        const int mixed = p->precision == NCG_MIXED;
        // more than 4 load cases: the FP32 operator goes through the matrix cores, up to 16
        // load cases per read of A (the vector kernel reads A once per 4)
        // (FP32 accumulation for the mixed form's Krylov products, FP64 for its refinement
        // residual and for the FP64 operator)
        const bool fp64acc = !mixed || acc64;
        const size_t el = fp64acc ? sizeof(double) : sizeof(float);
        // algorithmic bytes: A once per pass of (4 vector / 16 matrix-core) load cases, x and y
        const double vec = 2.0 * nr * n * 16.0, abytes = n * n * (mixed ? 8.0 : 16.0);
        const ncgk::MfmaPlan mp = ncgk::mfma_plan(p->n, fp64acc, p->tune.mfma_splits);
        void *xh = nullptr, *part = nullptr;
        bool use_mfma = !d && p->nrhs > 4;
        if (use_mfma && na) {
            xh = na->base;
            part = na->base + ((el * mp.xh_elems + 255) & ~(size_t)255);
        } else if (use_mfma) {
            use_mfma = !p->mm_xh.ensure(el * mp.xh_elems) && !p->mm_part.ensure(el * mp.part_elems);
            xh = p->mm_xh.p;
            part = p->mm_part.p;
        }
        if (use_mfma) {
            for (int r0 = 0; r0 < p->nrhs; r0 += 16) {
                const int nq = std::min(16, p->nrhs - r0);
                ncgk::mfma_xhat(p->stream, mp, p->n, nq, x + 2 * (size_t)r0 * p->n, xh);
                {
                    // the product kernel alone: A, this pass's x and y
                    KernelScope ks(p, fp64acc ? 11 : 10, abytes + 2.0 * nq * n * 16.0);
                    ncgk::mfma_product(p->stream, mp, p->A, mixed, p->n, xh, part);
                }
                ncgk::mfma_sum(p->stream, mp, p->n, nq, part, y + 2 * (size_t)r0 * p->n);
            }
        } else {
            KernelScope ks(p, 8 + mixed, chunks * abytes + vec);
            ncgk::matvec(p->stream, p->A, mixed, acc64, p->n, p->nrhs, x, d, y);
        }
    } else {
        // y = Near x + sum over levels S (D (T x))  (src/NC_CommonFunctions.cpp:1247-1316)
        const int mixed = p->precision == NCG_MIXED;
        const double *xin = x;
        if (d) {
            hipMemcpyAsync(p->fmm.xs.p, x, sizeof(cd) * (size_t)p->nrhs * p->n, hipMemcpyDeviceToDevice, p->stream);
            ncgk::scale_vec(p->stream, p->n, p->nrhs, d, p->fmm.xs.p);
            xin = p->fmm.xs.d();
        }
        {
            const double nnz = (double)p->fmm.nnz;
            KernelScope ks(p, mixed, chunks * nnz * (mixed ? 8.0 : 16.0) + 2.0 * nr * n * 16.0);
            ncgk::near_spmv(p->stream, p->n, p->nrhs, p->fmm.near_chunks(), p->fmm.near_col, p->fmm.near_wmax,
                            p->fmm.near_val.p, mixed, xin, y);
        }
        char *sc = na ? na->base : nullptr;
        for (DevLevel &L : p->fmm.lev) {
            // algorithmic bytes: the operator once per pass of at most 4 load cases, the
            // vectors once (t and s in the level's storage precision)
            const double sf = L.f32 ? 8.0 : 16.0, P = L.P, ncl = L.ncl, nfan = L.nfan;
            void *tw = L.tw.p, *sw = L.sw.p;
            if (sc) {
                const size_t sig = ((size_t)p->nrhs * L.ncl * L.P * (L.f32 ? 8 : 16) + 255) & ~(size_t)255;
                tw = sc;
                sw = sc + sig;
                sc += 2 * sig;
            }
            {
                KernelScope ks(p, 2 + L.f32, chunks * n * P * sf + nr * n * 16.0 + nr * ncl * P * sf);
                ncgk::fmm_apply_T(p->stream, p->n, L.ncl, L.P, p->nrhs, L.cl_ptr, L.cl_elem, L.T.p, L.f32, 0, xin, tw);
            }
            {
                KernelScope ks(p, 4 + L.f32, chunks * nfan * P * sf + 2.0 * nr * ncl * P * sf);
                ncgk::fmm_apply_D(p->stream, L.ncl, L.P, p->nrhs, L.fan_ptr, L.fan_idx, L.D.p, L.f32, 0, tw, sw);
            }
            {
                KernelScope ks(p, 6 + L.f32, chunks * n * P * sf + nr * ncl * P * sf + 2.0 * nr * n * 16.0);
                ncgk::fmm_apply_S(p->stream, p->n, L.ncl, L.P, p->nrhs, L.tiles(), L.cl_ptr, L.cl_elem, L.S.p, L.f32,
                                  sw, y);
            }
        }
    }
    p->ktime = ktime_keep;
    if (e) record(p, e->b);
}

// scratch bytes of one allocation-free apply of n_rhs load cases (see MvScratch)
size_t mv_scratch_bytes(const ncg_problem *p, int n_rhs)
{
    if (p->plan.method == 0) {
        if (n_rhs <= 4) return 0;
        const ncgk::MfmaPlan mp = ncgk::mfma_plan(p->n, 1, p->tune.mfma_splits);
        return ((sizeof(double) * mp.xh_elems + 255) & ~(size_t)255) + sizeof(double) * mp.part_elems;
    }
    size_t b = 0;
    for (const DevLevel &L : p->fmm.lev) b += 2 * (((size_t)n_rhs * L.ncl * L.P * (L.f32 ? 8 : 16) + 255) & ~(size_t)255);
    return b;
}

void harvest_mv(ncg_problem *p)
{
    hipStreamSynchronize(p->stream);
    for (size_t i = 0; i < p->mv_used; ++i) {
        p->mv_ms_sum += elapsed(p->ev_mv[i]);
        p->mv_calls++;
    }
    p->mv_used = 0;
    q_collect(p);
    static const double qflops[kQuadKinds] = {kGreenFlops, kGreenFlops, kGreenFlops,
                                              kGreenFlops, kGreenFlops, kGreenFlops};
    for (size_t i = 0; i < p->kpend.size(); ++i) {
        const KernelRec &r = p->kpend[i];
        p->k_ms[r.id] += elapsed(p->kpool[r.slot]);
        p->k_calls[r.id]++;
        p->k_bytes[r.id] += r.bytes;
        p->k_flops[r.id] += r.quad >= 0 ? p->q_evals[r.quad] * qflops[r.quad] : r.flops;
    }
    p->kpend.clear();
}

unsigned long long *q_counter(ncg_problem *p, int q, uint64_t sig)
{
    if (!p->ktime || !p->evals) return nullptr;
    if (p->q_sig[q] == sig && !(p->q_pending & (1 << q))) return nullptr;
    if (!(p->q_pending & (1 << q))) {
        hipMemsetAsync(p->evals + q, 0, sizeof(unsigned long long), p->stream);
        p->q_pending |= 1 << q;
    }
    p->q_sig[q] = sig;
    return p->evals + q;
}

void q_collect(ncg_problem *p)
{
    if (!p->q_pending) return;
    unsigned long long c[kQuadKinds];
    if (hipMemcpyAsync(c, p->evals, sizeof(c), hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
        return;
    for (int q = 0; q < kQuadKinds; ++q)
        if (p->q_pending & (1 << q)) p->q_evals[q] = (double)c[q];
    p->q_pending = 0;
}

KernelScope::KernelScope(ncg_problem *p_, int id_, double bytes_, int quad, double flops)
    : p(p_), id(id_), bytes(bytes_)
{
    if (!p->ktime) return;
    const size_t slot = p->kpend.size();
    if (slot == p->kpool.size()) {
        EventPair e;
        hipEventCreate(&e.a);
        hipEventCreate(&e.b);
        p->kpool.push_back(e);
    }
    p->kpend.push_back(KernelRec{id, slot, bytes, quad, flops});
    ev = &p->kpool[slot];
    record(p, ev->a);
}

KernelScope::~KernelScope()
{
    if (ev) record(p, ev->b);
}

int check_err(ncg_problem *p)
{
    int e = 0;
    NCG_HIP(hipMemcpyAsync(&e, p->err, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    NCG_HIP(hipStreamSynchronize(p->stream));
    if (e) {
        NCG_HIP(hipMemsetAsync(p->err, 0, sizeof(int), p->stream));
        if (e & 8) return fail(NCG_E_ARG, "incomplete LU decomposition: zero pivot");
        if (e & 16) return fail(NCG_E_ARG, "block-Jacobi preconditioner: singular cluster block");
        if (e & 1) return fail(NCG_E_SUBDIV, "number of subels. which are subdivided in a loop must <= 15!");
        if (e & 2) return fail(NCG_E_SUBDIV, "nsbe must be < MSBE(= 220)");
        return fail(NCG_E_SUBDIV, "adaptive subdivision deeper than 24 levels (point on an element?)");
    }
    return 0;
}

// Frequency-independent table of the dense method's near pairs (the pairs NumCalc
// subdivides), built on the first dense assembly.
int ensure_near_table(ncg_problem *p)
{
    if (p->near_rows) return 0;
    hipStream_t s = p->stream;
    const int n = p->n;
    int rc;
    int *cnt = nullptr;
    long long *off = nullptr;
    if ((rc = dalloc(&cnt, n)) || (rc = dalloc(&off, (size_t)n + 1))) {
        hipFree(cnt);
        return rc;
    }
    ncgk::near_count(s, p->elements(), cnt);
    std::vector<int> ch(n);
    hipMemcpyAsync(ch.data(), cnt, sizeof(int) * n, hipMemcpyDeviceToHost, s);
    hipError_t e = hipStreamSynchronize(s);
    std::vector<long long> oh((size_t)n + 1, 0);
    for (int i = 0; i < n; ++i) oh[i + 1] = oh[i] + ch[i];
    if (e == hipSuccess) {
        if ((rc = dalloc(&p->near_rows, (size_t)oh[n])) || (rc = dalloc(&p->near_cols, (size_t)oh[n]))) {
            hipFree(p->near_rows);
            p->near_rows = nullptr;
            hipFree(cnt);
            hipFree(off);
            return rc;
        }
        p->nnear = oh[n];
        hipMemcpyAsync(off, oh.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, s);
        ncgk::near_fill(s, p->elements(), off, p->near_rows, p->near_cols);
        e = hipStreamSynchronize(s);
    }
    hipFree(cnt);
    hipFree(off);
    if (e != hipSuccess) return fail(NCG_E_DEVICE, std::string("near table: ") + hipGetErrorString(e));
    return 0;
}

// ---------------------------------------------------------------------------------- FMM
// host mesh of the planner from the ABI geometry
int build_fmm_mesh(const ncg_geometry *g, ncgf::Mesh &M)
{
    const int n = g->n_elements;
    M = ncgf::Mesh();
    M.n = n;
    M.nv.assign(g->nodes_per_element, g->nodes_per_element + n);
    M.X.assign(g->node_xyz, g->node_xyz + 12 * (size_t)n);
    M.node.assign(4 * (size_t)n, -1);
    if (g->node_index) {
        for (size_t i = 0; i < 4 * (size_t)n; ++i) M.node[i] = g->node_index[i];
    } else {
        // nodes with bitwise equal coordinates are one mesh node
        std::vector<std::pair<std::array<double, 3>, int>> keys;
        std::map<std::array<double, 3>, int> id;
        for (int e = 0; e < n; ++e)
            for (int a = 0; a < M.nv[e]; ++a) {
                std::array<double, 3> k3 = {M.X[12 * (size_t)e + 3 * a], M.X[12 * (size_t)e + 3 * a + 1],
                                            M.X[12 * (size_t)e + 3 * a + 2]};
                auto it = id.find(k3);
                if (it == id.end()) it = id.emplace(k3, (int)id.size()).first;
                M.node[4 * (size_t)e + a] = it->second;
            }
    }
    M.group.assign(n, 0);
    if (g->element_group) M.group.assign(g->element_group, g->element_group + n);
    M.m = g->n_eval > 0 ? g->n_eval : 0;
    if (M.m) M.eval.assign(g->eval_xyz, g->eval_xyz + 3 * (size_t)M.m);
    ncgf::prepare_mesh(M);
    return 0;
}

ncgf::Options fmm_options(const ncg_geometry *g)
{
    ncgf::Options o;
    o.method_bem = g->method.method_bem;
    o.min_terms = g->method.min_expansion_terms == 0 ? 8 : g->method.min_expansion_terms;
    o.clus_edge = g->method.cluster_edge;
    return o;
}

// near-field nonzeros of row DOFs in cluster c: elements of its near clusters
long long near_count(const ncgf::Level &L, int c)
{
    long long s = 0;
    for (int j : L.cl[c].near) s += (long long)L.cl[j].elems.size();
    return s;
}

void fill_info(const ncgf::Plan &P, int n, ncg_fmm_info *info)
{
    memset(info, 0, sizeof(*info));
    info->method = P.method;
    info->levels = (int)P.lev.size();
    for (size_t l = 0; l < P.lev.size() && l < 10; ++l) {
        const ncgf::Level &L = P.lev[l];
        info->clusters[l] = (int)L.cl.size();
        info->radius_max[l] = L.rmax;
        info->radius_min[l] = L.rmin;
        info->expansion_terms[l] = L.L;
        info->points_theta[l] = L.ntheta;
        info->points_phi[l] = L.nphi;
        info->points_sphere[l] = L.P;
        long long fan = 0;
        for (const ncgf::Cluster &c : L.cl) fan += (long long)c.fan.size();
        info->nnz_t[l] = (long long)n * L.P;
        info->nnz_d[l] = fan * L.P;
        info->nnz_s[l] = (long long)n * L.P;
        info->canceled_pairs[l] = L.n_canceled;
        info->far_pairs[l] = L.n_far_geom;
    }
    if (!P.lev.empty()) {
        const ncgf::Level &F = P.lev.back();
        for (int c = 0; c < (int)F.cl.size(); ++c) info->nnz_near += (long long)F.cl[c].elems.size() * near_count(F, c);
    }
    info->eval_clusters = (int)P.ev.size();
    info->eval_radius_max = P.ev_rmax;
    info->eval_radius_min = P.ev.empty() ? 0.0 : P.ev_rmin;
    snprintf(info->warning, sizeof(info->warning), "%s", P.warning.c_str());
}

template <typename T>
int upload(Pool &pool, hipStream_t s, T **dst, const std::vector<T> &src)
{
    int rc = pool.alloc(dst, src.size());
    if (rc) return rc;
    if (!src.empty()) {
        hipError_t e = hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail(NCG_E_DEVICE, std::string("FMM upload: ") + hipGetErrorString(e));
    }
    return 0;
}

template <typename T>
int upload_grow(Grow &g, hipStream_t s, const std::vector<T> &src)
{
    int rc = g.ensure(sizeof(T) * src.size());
    if (rc) return rc;
    if (!src.empty()) {
        hipError_t e = hipMemcpyAsync(g.p, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return fail(NCG_E_DEVICE, std::string("FMM upload: ") + hipGetErrorString(e));
    }
    return 0;
}

// structure signature of a plan: method, clusters, interaction lists, finest near lists
// Within one problem the cluster tree is fixed by (method, levels, cluster edge), and the
// near and interaction lists then follow from the set of cancelled pairs of each level.
unsigned long long plan_signature(const ncgf::Plan &P)
{
    Fnv f;
    f.add((long long)P.method);
    f.add((long long)P.lev.size());
    long long edge;
    memcpy(&edge, &P.clus_edge, sizeof edge);
    f.add(edge);
    for (const ncgf::Level &L : P.lev) {
        f.add((long long)L.cl.size());
        f.add(L.n_far_geom);
        f.add(L.n_canceled);
        f.add((long long)L.cancel_hash);
    }
    return f.h;
}

// cluster structure of every level and the near-field CSR on the device
int fmm_structure(ncg_problem *p)
{
    hipStream_t s = p->stream;
    DevFmm &F = p->fmm;
    const int n = p->n;
    int rc;
    for (const ncgf::Level &L : p->plan.lev) {
        DevLevel D;
        D.ncl = (int)L.cl.size();
        std::vector<double> cc(3 * (size_t)D.ncl);
        std::vector<int> cl_ptr(D.ncl + 1, 0), cl_elem, elem_cl(n, -1), elem_pos(n, -1), fan_ptr(D.ncl + 1, 0), fan_idx,
            fan_row;
        cl_elem.reserve(n);
        for (int c = 0; c < D.ncl; ++c) {
            for (int d = 0; d < 3; ++d) cc[3 * c + d] = L.cl[c].cen[d];
            for (int e : L.cl[c].elems) {
                elem_cl[e] = c;
                elem_pos[e] = (int)cl_elem.size();
                cl_elem.push_back(e);
            }
            cl_ptr[c + 1] = (int)cl_elem.size();
            for (int j : L.cl[c].fan) {
                fan_idx.push_back(j);
                fan_row.push_back(c);
                const double dx = L.cl[c].cen[0] - L.cl[j].cen[0], dy = L.cl[c].cen[1] - L.cl[j].cen[1],
                             dz = L.cl[c].cen[2] - L.cl[j].cen[2];
                D.dmin = std::min(D.dmin, sqrt(dx * dx + dy * dy + dz * dz));
            }
            fan_ptr[c + 1] = (int)fan_idx.size();
        }
        D.nfan = (int)fan_idx.size();
        std::vector<int> ch_cl, ch_r0, ch_m;
        for (int c = 0; c < D.ncl; ++c) {
            const int m = cl_ptr[c + 1] - cl_ptr[c], nch = (m + ncgk::kChunkRows - 1) / ncgk::kChunkRows;
            for (int q = 0, r0 = cl_ptr[c]; q < nch; ++q) {
                const int mq = m / nch + (q < m % nch ? 1 : 0);
                ch_cl.push_back(c);
                ch_r0.push_back(r0);
                ch_m.push_back(mq);
                r0 += mq;
            }
        }
        D.nch = (int)ch_cl.size();
        std::vector<int> tl_cl, tl_a0, tl_m;
        for (int c = 0; c < D.ncl; ++c) {
            const int m = cl_ptr[c + 1] - cl_ptr[c], nt = (m + 63) / 64;
            for (int q = 0, a0 = 0; q < nt; ++q) {
                const int mq = m / nt + (q < m % nt ? 1 : 0);
                tl_cl.push_back(c);
                tl_a0.push_back(a0);
                tl_m.push_back(mq);
                a0 += mq;
            }
        }
        D.ntl = (int)tl_cl.size();
        if ((rc = upload(F.st, s, &D.ccen, cc)) || (rc = upload(F.st, s, &D.ch_cl, ch_cl)) ||
            (rc = upload(F.st, s, &D.tl_cl, tl_cl)) || (rc = upload(F.st, s, &D.tl_a0, tl_a0)) ||
            (rc = upload(F.st, s, &D.tl_m, tl_m)) ||
            (rc = upload(F.st, s, &D.ch_r0, ch_r0)) || (rc = upload(F.st, s, &D.ch_m, ch_m)) || (rc = upload(F.st, s, &D.cl_ptr, cl_ptr)) ||
            (rc = upload(F.st, s, &D.cl_elem, cl_elem)) || (rc = upload(F.st, s, &D.elem_cl, elem_cl)) ||
            (rc = upload(F.st, s, &D.elem_pos, elem_pos)) || (rc = upload(F.st, s, &D.fan_ptr, fan_ptr)) ||
            (rc = upload(F.st, s, &D.fan_idx, fan_idx)) || (rc = upload(F.st, s, &D.fan_row, fan_row)))
            return rc;
        F.lev.push_back(D);
    }
    // near field of the finest level: row i -> elements of the near clusters of its
    // cluster; row lengths are per-cluster sums, the columns are written on the device
    const ncgf::Level &FL = p->plan.lev.back();
    const DevLevel &DF = F.lev.back();
    const int ncf = (int)FL.cl.size();
    std::vector<long long> rowlen(ncf, 0);
    std::vector<int> nn_ptr(ncf + 1, 0), nn_idx, cl_of(n, -1);
    std::vector<unsigned char> cnear((size_t)ncf * ncf, 0);
    for (int c = 0; c < ncf; ++c) {
        for (int j : FL.cl[c].near) {
            rowlen[c] += (long long)FL.cl[j].elems.size();
            nn_idx.push_back(j);
            cnear[(size_t)c * ncf + j] = 1;
        }
        nn_ptr[c + 1] = (int)nn_idx.size();
        for (int e : FL.cl[c].elems) cl_of[e] = c;
    }
    std::vector<long long> ptr(n + 1, 0);
    for (int i = 0; i < n; ++i) ptr[i + 1] = ptr[i] + rowlen[cl_of[i]];
    // chunk table of the near-field apply: runs of at most kNearRows elements of each finest
    // cluster (balanced sizes) with their column base (the cluster's first row), width, rows
    // and value offsets
    const int R = ncgk::kNearRows;
    std::vector<long long> nd_cb, nd_off;
    std::vector<int> nd_w, nd_row, nd_cptr(1, 0);
    for (int c = 0; c < ncf; ++c) {
        const std::vector<int> &el = FL.cl[c].elems;
        const int m = (int)el.size(), nch = (m + R - 1) / R;
        nd_cptr.push_back(nd_cptr.back() + nch);
        for (int q = 0, r0 = 0; q < nch; ++q) {
            const int mq = m / nch + (q < m % nch ? 1 : 0);
            const long long cb = ptr[el[0]];
            nd_cb.push_back(cb);
            nd_w.push_back((int)rowlen[c]);
            for (int a = 0; a < R; ++a) {
                // an unused slot reads the first row and stores nothing
                nd_row.push_back(a < mq ? el[r0 + a] : -1);
                nd_off.push_back(a < mq ? ptr[el[r0 + a]] : cb);
            }
            r0 += mq;
        }
    }
    F.nd_n = (int)nd_cb.size();
    // block-Jacobi layout: offset of the own cluster inside its rows; blocks are runs of
    // at most 64 consecutive elements of a cluster's list, of balanced sizes
    std::vector<int> self_off(ncf, 0), elem_blk(n, -1), blk_cl, blk_b0, blk_m;
    std::vector<long long> boff(1, 0);
    int cstart = 0;
    for (int c = 0; c < ncf; ++c) {
        for (int j : FL.cl[c].near) {
            if (j == c) break;
            self_off[c] += (int)FL.cl[j].elems.size();
        }
        const int m = (int)FL.cl[c].elems.size(), nch = (m + 63) / 64;
        int start = 0;
        for (int q = 0; q < nch; ++q) {
            const int mb = m / nch + (q < m % nch ? 1 : 0);
            const int b = (int)blk_m.size();
            blk_cl.push_back(c);
            blk_b0.push_back(cstart + start);
            blk_m.push_back(mb);
            boff.push_back(boff.back() + (long long)mb * mb);
            for (int t = start; t < start + mb; ++t) elem_blk[FL.cl[c].elems[t]] = b;
            start += mb;
        }
        cstart += m;
    }
    F.nblk = (int)blk_m.size();
    F.bj_total = boff.back();
    F.nnz = ptr[n];
    F.ncl_fine = ncf;
    F.near_wmax = rowlen.empty() ? 0 : (int)*std::max_element(rowlen.begin(), rowlen.end());
    int *d_nn_ptr, *d_nn_idx;
    if ((rc = upload(F.st, s, &F.near_ptr, ptr)) || (rc = F.st.alloc(&F.diag_pos, n)) ||
        (rc = upload(F.st, s, &F.nd_cb, nd_cb)) || (rc = upload(F.st, s, &F.nd_off, nd_off)) ||
        (rc = upload(F.st, s, &F.nd_w, nd_w)) || (rc = upload(F.st, s, &F.nd_row, nd_row)) ||
        (rc = upload(F.st, s, &F.nd_cptr, nd_cptr)) ||
        (rc = F.st.alloc(&F.near_col, F.nnz)) ||
        (rc = upload(F.st, s, &F.elem_cl_fine, cl_of)) || (rc = upload(F.st, s, &F.cl_near, cnear)) ||
        (rc = upload(F.st, s, &d_nn_ptr, nn_ptr)) || (rc = upload(F.st, s, &d_nn_idx, nn_idx)) ||
        (rc = upload(F.st, s, &F.self_off, self_off)) || (rc = upload(F.st, s, &F.boff, boff)) ||
        (rc = upload(F.st, s, &F.elem_blk, elem_blk)) || (rc = upload(F.st, s, &F.blk_cl, blk_cl)) ||
        (rc = upload(F.st, s, &F.blk_b0, blk_b0)) || (rc = upload(F.st, s, &F.blk_m, blk_m)))
        return rc;
    ncgk::near_csr(s, n, F.elem_cl_fine, DF.cl_ptr, DF.cl_elem, d_nn_ptr, d_nn_idx, F.near_ptr, nullptr, F.near_col,
                   F.diag_pos);
    // positions of the subdivided pairs (the dense method's near-pair table) in the CSR
    std::vector<int> nn_off(nn_idx.size(), 0);
    for (int c = 0; c < ncf; ++c) {
        int o = 0;
        for (int t = nn_ptr[c]; t < nn_ptr[c + 1]; ++t) {
            nn_off[t] = o;
            o += (int)FL.cl[nn_idx[t]].elems.size();
        }
    }
    int *d_nn_off;
    if ((rc = ensure_near_table(p)) || (rc = upload(F.st, s, &d_nn_off, nn_off)) ||
        (rc = F.st.alloc(&F.adj_pos, (size_t)p->nnear)))
        return rc;
    ncgk::adj_positions(s, p->nnear, p->near_rows, p->near_cols, F.elem_cl_fine, DF.elem_pos, DF.cl_ptr, d_nn_ptr,
                        d_nn_idx, d_nn_off, F.near_ptr, F.adj_pos);
    F.sig = plan_signature(p->plan);
    return 0;
}

// max over n <= L of (2n+1) max(|y_n(x)|, 1/x): a bound on the largest term of a translation
// series at k d = x (|h_n| ~ |y_n| once n > x, ~ 1/x below), y_n by upward recurrence
double translation_peak(int L, double x)
{
    double y0 = -cos(x) / x, y1 = -cos(x) / (x * x) - sin(x) / x;
    double m = std::max(std::max(std::fabs(y0), 1.0 / x), 3.0 * std::max(std::fabs(y1), 1.0 / x));
    for (int q = 1; q < L; ++q) {
        const double y2 = (2.0 * q + 1.0) / x * y1 - y0;
        y0 = y1;
        y1 = y2;
        m = std::max(m, (2.0 * q + 3.0) * std::max(std::fabs(y1), 1.0 / x));
    }
    return m;
}

// Build and fill the device FMM operator of the current plan (NC_BuildSystemFMBEM,
// src/NC_EquationSystem.cpp:305-365, with the CSR layouts of NC_AllocateNmtx and
// NC_AllocateSDTmtxs*, src/NC_Addresses.cpp:1476-1937).  The structure is rebuilt only
// when the plan's clusters, interaction lists or near lists change.
int assemble_fmm(ncg_problem *p)
{
    hipStream_t s = p->stream;
    DevFmm &F = p->fmm;
    const int n = p->n, nr = p->nrhs;
    const int mixed = p->precision == NCG_MIXED;
    const size_t se = mixed ? 8 : 16;
    int rc;
    const ncgk::BMParams b = p->bm();
    const double h = p->h;
    if (F.sig == 0 || F.sig != plan_signature(p->plan) || F.lev.size() != p->plan.lev.size()) {
        NCG_HIP(hipStreamSynchronize(s));
        F.release_structure();
        if ((rc = fmm_structure(p))) return rc;
    }
    for (size_t l = 0; l < F.lev.size(); ++l) {
        const ncgf::Level &L = p->plan.lev[l];
        DevLevel &D = F.lev[l];
        D.P = L.P;
        D.L = L.L;
        // The translation series sum (2n+1) h_n(k d) terms up to n = L against O(1)
        // aggregation and disaggregation entries; where they are large they cancel, and FP32
        // storage of T, D and S (with FP32 signatures) then breaks the refinement (the inner
        // FP32 operator no longer reduces the FP64 residual).  A level keeps FP64 in the
        // mixed form unless k d_min >= 1 and peak * k d_min <= 2e7, peak the series' largest
        // term bound at d_min.  Measured on head-like meshes: up to 7e7 the refinement
        // converged in one step, 1.2e8 took two (5120 elements at 12.8 kHz), 2.2e8 and more
        // did not converge (5120 elements at 12.4 kHz; 60320 graded elements at 18 kHz, 2.8e9).
        const double kd = p->k * D.dmin;
        D.f32 = mixed && kd >= 1.0 && translation_peak(D.L, kd) * kd <= 2e7;
        const size_t sf = D.f32 ? 8 : 16;
        if ((rc = upload_grow(D.u, s, L.u)) || (rc = upload_grow(D.w, s, L.w)) || (rc = D.T.ensure(sf * (size_t)n * D.P)) ||
            (rc = D.D.ensure(sf * (size_t)std::max(D.nfan, 1) * D.P)) || (rc = D.S.ensure(sf * (size_t)n * D.P)) ||
            (rc = D.tw.ensure(16 * (size_t)nr * D.ncl * D.P)) || (rc = D.sw.ensure(16 * (size_t)nr * D.ncl * D.P)) ||
            (rc = D.tv.ensure(16 * (size_t)nr * D.ncl * D.P)))
            return rc;
    }
    if ((rc = F.near_val.ensure(se * (size_t)F.nnz)) || (rc = F.xs.ensure(16 * (size_t)nr * n))) return rc;
    F.bj_ready = 0;
    // values
    const ncgk::DevElements E = p->elements();
    const DevLevel &DF = F.lev.back();
    {
        KernelScope ks(p, kKidFmmNear + mixed, 0.0, kQFmmNear);
        ncgk::fmm_near(s, E, b, p->adm, DF.chunks(), DF.cl_ptr, DF.cl_elem, F.near_ptr, F.near_col, F.near_wmax,
                       F.near_val.p, mixed, q_counter(p, kQFmmNear, F.sig), p->tune.near_assembly);
    }
    {
        KernelScope ks(p, kKidFmmAdj, 0.0, kQFmmAdj);
        ncgk::fmm_near_adj(s, E, b, p->adm, p->near_rows, p->near_cols, F.adj_pos, p->nnear, F.near_val.p, mixed,
                           p->err, q_counter(p, kQFmmAdj, F.sig));
    }
    if (p->nsrc)
        ncgk::rhs_sources(s, E, b, p->src, p->nsrc, p->bcval, nr, p->rhs_src, p->err, F.elem_cl_fine, F.cl_near,
                          F.ncl_fine);
    {
        KernelScope ks(p, kKidDiag, 0.0, kQDiag);
        ncgk::assemble_diag(s, E, b, p->adm, p->bcval, nr, p->ps, p->nps, p->pw, p->npw, p->rhs_src, p->nsrc > 0,
                            F.near_val.p, mixed, p->rhs, F.diag_pos, q_counter(p, kQDiag, 1));
    }
    for (DevLevel &D : F.lev) {
        if ((rc = F.dcoef.ensure(sizeof(double) * (size_t)std::max(D.nfan, 1) * (3 + 2 * D.L)))) return rc;
        const double nP = (double)n * D.P;
        {
            KernelScope ks(p, kKidFmmD, 0.0, -1, (double)D.nfan * D.P * D.L * kDTermFlops);
            ncgk::fmm_D(s, D.L, D.P, p->k, h, D.ccen, D.ncl, D.fan_row, D.fan_idx, D.nfan, D.u.d(), D.D.p, D.f32,
                        F.dcoef.d());
        }
        {
            KernelScope ks(p, kKidFmmT + D.f32, 0.0, -1, nP * 4.0 * kTFlops);
            ncgk::fmm_T(s, E, b, p->adm, D.P, D.ccen, D.elem_cl, D.elem_pos, D.u.d(), D.T.p, D.f32);
        }
        {
            KernelScope ks(p, kKidFmmS + D.f32, 0.0, -1, nP * kSFlops);
            ncgk::fmm_S(s, E, b, D.P, D.ccen, D.elem_cl, D.cl_ptr, D.cl_elem, D.u.d(), D.w.d(), D.S.p, D.f32);
        }
        if (p->nsrc) {
            // right-hand side far part: b += S D t_v (NC_RHSvecTof*, src/NC_EquationSystem.cpp:1068-1118)
            ncgk::fmm_Tvec_bc(s, E, b, D.P, D.ncl, D.ccen, D.cl_ptr, D.cl_elem, D.u.d(), p->bcval, nr, D.tv.d());
            ncgk::fmm_apply_D(s, D.ncl, D.P, nr, D.fan_ptr, D.fan_idx, D.D.p, D.f32, 1, D.tv.d(), D.sw.p);
            ncgk::fmm_apply_S(s, n, D.ncl, D.P, nr, D.tiles(), D.cl_ptr, D.cl_elem, D.S.p, D.f32, D.sw.p, p->rhs);
        }
    }
    return 0;
}

// evaluation grid through the level-0 expansions (NC_ContributionFMM,
// src/NC_PostProcessing.cpp:978-1214)
int field_fmm(ncg_problem *p, double *pot, double *vel, double *dphi, double *dgrad)
{
    hipStream_t s = p->stream;
    DevFmm &F = p->fmm;
    const ncgf::Level &L0 = p->plan.lev.front();
    DevLevel &D0 = F.lev.front();
    const int nr = p->nrhs, m = p->m, nev = (int)p->plan.ev.size();
    const double fac = 2.0 / sqrt(3.0);
    int rc;
    // The pair lists depend on the level-0 clusters and, through the cancelled-distance
    // test di > min_dist, on k.  Lists built while min_dist was below every geometric-far
    // distance hold for any smaller min_dist as well: they are reused without a rebuild.
    const bool reuse = F.ev_sig != 0 && F.ev_struct == F.sig && F.ev_pure && p->plan.min_dist < F.ev_dmin_far;
    if (!reuse) {
        std::vector<int> pt_ev(m, 0), ev_ptr(nev + 1, 0), pair_ev, pair_cl, el_ptr(nev + 1, 0), el_idx;
        std::vector<double> ecen(3 * (size_t)nev);
        double dmin_far = 1e300;
        bool pure = true;
        for (int a = 0; a < nev; ++a) {
            const ncgf::EvalCluster &E = p->plan.ev[a];
            for (int d = 0; d < 3; ++d) ecen[3 * a + d] = E.cen[d];
            for (int x : E.pts) pt_ev[x] = a;
            for (int c = 0; c < (int)L0.cl.size(); ++c) {
                const ncgf::Cluster &C = L0.cl[c];
                const double dx = C.cen[0] - E.cen[0], dy = C.cen[1] - E.cen[1], dz = C.cen[2] - E.cen[2];
                const double di = sqrt(dx * dx + dy * dy + dz * dz), dk = C.radius + E.radius;
                const bool geo_far = di > fac * dk;
                if (geo_far) dmin_far = std::min(dmin_far, di);
                if (geo_far && di > p->plan.min_dist) {
                    pair_ev.push_back(a);
                    pair_cl.push_back(c);
                } else {
                    if (geo_far) pure = false;
                    el_idx.insert(el_idx.end(), C.elems.begin(), C.elems.end());
                }
            }
            ev_ptr[a + 1] = (int)pair_ev.size();
            el_ptr[a + 1] = (int)el_idx.size();
        }
        Fnv f;
        f.add((long long)F.sig);
        for (const std::vector<int> *v : {&pt_ev, &ev_ptr, &pair_ev, &pair_cl, &el_ptr, &el_idx}) f.add(*v);
        if (f.h != F.ev_sig) {
            NCG_HIP(hipStreamSynchronize(s));
            F.evp.release();
            F.ev_sig = 0;
            if ((rc = upload(F.evp, s, &F.pt_ev, pt_ev)) || (rc = upload(F.evp, s, &F.ev_ptr, ev_ptr)) ||
                (rc = upload(F.evp, s, &F.pair_ev, pair_ev)) || (rc = upload(F.evp, s, &F.pair_cl, pair_cl)) ||
                (rc = upload(F.evp, s, &F.el_ptr, el_ptr)) || (rc = upload(F.evp, s, &F.el_idx, el_idx)) ||
                (rc = upload(F.evp, s, &F.ecen, ecen)))
                return rc;
            F.ev_sig = f.h;
        }
        F.ev_struct = F.sig;
        F.ev_pure = pure;
        F.ev_dmin_far = dmin_far;
        F.npair = (int)pair_ev.size();
        F.nev = nev;
    }
    if ((rc = F.tab.ensure(sizeof(double) * (size_t)std::max(F.npair, 1) * (4 + 2 * D0.L))) ||
        (rc = F.z.ensure(16 * (size_t)nr * nev * L0.P)) || (rc = F.fphi.ensure(16 * (size_t)nr * m)) ||
        (rc = F.fgrad.ensure(48 * (size_t)nr * m)))
        return rc;
    const ncgk::BMParams b = p->bm();
    // Arithmetic of the evaluation pairs' translation series (and of the solution's level-0
    // signature feeding it): FP32 in the mixed form where the rule that keeps FMM levels in
    // FP32 (fmm_operator: k d >= 1 and peak * k d <= 2e7) holds at the nearest far
    // evaluation pair.  Evaluation points usually lie much farther from the level-0
    // clusters than the clusters from each other, so the grid can be FP32 where level 0 is
    // not.
    const double kde = p->k * F.ev_dmin_far;
    const int ev32 = p->precision == NCG_MIXED && F.npair > 0 && kde >= 1.0 && translation_peak(D0.L, kde) * kde <= 2e7;
    // the signature itself stored in FP32 there too (the series' reads of it bound the
    // evaluation); where it is the stored T applied to the potential, only when level 0's
    // T is FP32 as well
    const bool tv_apply = p->all_velocity && p->nsrc == 0;
    const int tv32 = ev32 && (!tv_apply || D0.f32);
    if (tv_apply) {
        // velocity elements with no prescribed velocity: the signature of the solution is
        // the level-0 aggregation T applied to the potential (same sum as below)
        ncgk::fmm_apply_T(s, p->n, D0.ncl, D0.P, nr, D0.cl_ptr, D0.cl_elem, D0.T.p, D0.f32, tv32 ? 0 : 1, pot,
                          D0.tv.d());
    } else {
        ncgk::fmm_Tvec_sol(s, p->elements(), b, D0.P, D0.ncl, D0.ccen, D0.cl_ptr, D0.cl_elem, D0.u.d(), pot, vel, nr,
                           D0.tv.d(), ev32, tv32);
    }
    ncgk::fmm_eval_hankel(s, D0.L, p->k, p->h, F.ecen, D0.ccen, F.pair_ev, F.pair_cl, F.npair, F.tab.d());
    ncgk::fmm_eval_z(s, nev, D0.L, D0.P, p->h, F.ev_ptr, F.pair_cl, F.tab.d(), D0.u.d(), D0.w.d(), D0.tv.d(), D0.ncl,
                     nr, F.z.p, ev32, tv32);
    ncgk::fmm_eval_far(s, m, D0.P, nev, b.hk, p->eval, F.pt_ev, F.ecen, D0.u.d(), nr, F.z.p, F.fphi.p, F.fgrad.p);
    {
        const int f32 = p->precision == NCG_MIXED;
        KernelScope ks(p, kKidField + f32, 0.0, kQField);
        ncgk::field(s, p->elements(), b, p->eval, m, nr, pot, vel, p->ps, p->nps, p->pw, p->npw, dphi, dgrad, f32,
                    p->err, F.pt_ev, F.el_ptr, F.el_idx, F.fphi.p, F.fgrad.p,
                    q_counter(p, kQField, F.sig * 31 + F.ev_sig + 2));
    }
    return 0;
}

// d-vector helpers ------------------------------------------------------------------
int dotv(ncg_problem *p, int ndots, const double *const *u, const double *const *v, int conj_mask, std::vector<cd> &out)
{
    ncgk::dots(p->stream, p->n, p->nrhs, ndots, (const void *const *)u, (const void *const *)v, conj_mask, p->partial,
               p->dotout);
    out.resize((size_t)p->nrhs * 3);
    NCG_HIP(hipMemcpyAsync(out.data(), p->dotout, sizeof(cd) * 3 * p->nrhs, hipMemcpyDeviceToHost, p->stream));
    NCG_HIP(hipStreamSynchronize(p->stream));
    return 0;
}

// host vector to a new device array owned by the ILU (asynchronous on s)
int upload_ilu(DevIlu &I, hipStream_t s, int **dst, const std::vector<int> &v)
{
    int rc = I.alloc(dst, v.size());
    if (rc) return rc;
    if (!v.empty()) NCG_HIP(hipMemcpyAsync(*dst, v.data(), sizeof(int) * v.size(), hipMemcpyHostToDevice, s));
    return 0;
}

// ---------------------------------------------------------------------------------- ILU
// Incomplete LU of NC_IncompleteLUDecomposition (src/NC_CommonFunctions.cpp:1356-1659):
// rows scaled by d_i = sqrt(n_i / sum |a_ij|^2) over the dense row or the near-field row,
// pattern = scaled entries above the method's threshold plus the diagonal, ILU(0) on it.
// The scaling stays implicit: the preconditioned operator is D A (LU)^-1 and the
// right-hand side D b, which is what the reference's in-place scaling of A, S and b gives.
// Pattern counts, entries and factors are computed on the device; the host only turns
// the column lists into level schedules.
int ilu_setup(ncg_problem *p, int degree)
{
    static const double thr_tbem[4] = {1.2, 1.0, 0.8, 0.6}, thr_sl[4] = {0.9, 0.35, 0.07, 0.01},
                        thr_ml[4] = {0.65, 0.15, 0.05, 0.005};
    if (p->precision != NCG_FP64) return fail(NCG_E_ARG, "the ILU preconditioner needs the FP64 operator");
    const int deg = std::min(std::max(degree, 0), 3);
    const double thr = p->plan.method == 0 ? thr_tbem[deg] : p->plan.method == 1 ? thr_sl[deg] : thr_ml[deg];
    hipStream_t s = p->stream;
    const int n = p->n;
    DevIlu &I = p->ilu;
    NCG_HIP(hipStreamSynchronize(s));
    I.release();
    const bool dense = p->plan.method == 0;
    const void *A = dense ? p->A : p->fmm.near_val.p;
    const long long *ptr = dense ? nullptr : p->fmm.near_ptr;
    const int *ncol = dense ? nullptr : p->fmm.near_col;
    int rc;
    int *cnt;
    if ((rc = I.alloc(&cnt, n)) || (rc = I.alloc(&I.rp, n + 1)) || (rc = I.alloc(&I.diag, n))) return rc;
    if (dense) ncgk::row_scaling(s, A, 0, n, p->dvec);
    else ncgk::row_scaling_csr(s, n, ptr, A, 0, p->dvec);
    ncgk::ilu_count(s, A, n, ptr, ncol, p->dvec, thr, cnt);
    std::vector<int> c(n), rp(n + 1, 0);
    NCG_HIP(hipMemcpyAsync(c.data(), cnt, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    NCG_HIP(hipStreamSynchronize(s));
    long long tot = 0;
    for (int i = 0; i < n; ++i) {
        rp[i] = (int)tot;
        tot += c[i];
        if (tot > 0x7fffffffLL) return fail(NCG_E_MEMORY, "incomplete LU pattern exceeds 2^31 entries");
    }
    rp[n] = (int)tot;
    I.nnz = (int)tot;
    if ((rc = I.alloc(&I.col, tot)) || (rc = I.alloc(&I.val, 2 * (size_t)tot))) return rc;
    NCG_HIP(hipMemcpyAsync(I.rp, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    ncgk::ilu_fill(s, A, n, ptr, ncol, p->dvec, thr, I.rp, I.col, I.val);
    std::vector<int> col(tot);
    NCG_HIP(hipMemcpyAsync(col.data(), I.col, sizeof(int) * tot, hipMemcpyDeviceToHost, s));
    NCG_HIP(hipStreamSynchronize(s));
    // diagonal positions and the level of every row in the L and U dependency graphs
    std::vector<int> diag(n, -1), lvL(n, 0), lvU(n, 0);
    int nL = 0, nU = 0;
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            if (col[q] < i) l = std::max(l, lvL[col[q]] + 1);
            else if (col[q] == i) diag[i] = q;
        }
        if (diag[i] < 0) return fail(NCG_E_STATE, "incomplete LU: diagonal missing from the pattern");
        lvL[i] = l;
        nL = std::max(nL, l + 1);
    }
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        for (int q = diag[i] + 1; q < rp[i + 1]; ++q) l = std::max(l, lvU[col[q]] + 1);
        lvU[i] = l;
        nU = std::max(nU, l + 1);
    }
    auto schedule = [&](const std::vector<int> &lv, int nlev, std::vector<int> &lp, std::vector<int> &rows) {
        lp.assign(nlev + 1, 0);
        for (int i = 0; i < n; ++i) lp[lv[i] + 1]++;
        for (int l = 0; l < nlev; ++l) lp[l + 1] += lp[l];
        rows.assign(n, 0);
        std::vector<int> fill(lp.begin(), lp.end() - 1);
        for (int i = 0; i < n; ++i) rows[fill[lv[i]]++] = i;
    };
    std::vector<int> lpL, rowsL, lpU, rowsU;
    schedule(lvL, nL, lpL, rowsL);
    schedule(lvU, nU, lpU, rowsU);
    if ((rc = I.alloc(&I.levL_ptr, nL + 1)) || (rc = I.alloc(&I.levL_rows, n)) || (rc = I.alloc(&I.levU_ptr, nU + 1)) ||
        (rc = I.alloc(&I.levU_rows, n)))
        return rc;
    NCG_HIP(hipMemcpyAsync(I.diag, diag.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.levL_ptr, lpL.data(), sizeof(int) * (nL + 1), hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.levL_rows, rowsL.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.levU_ptr, lpU.data(), sizeof(int) * (nU + 1), hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.levU_rows, rowsU.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    I.nlevL = nL;
    I.nlevU = nU;
    // the factors in the order the solves visit them
    std::vector<int> Loff(n + 1, 0), Uoff(n + 1, 0), Lc, Lmap, Uc, Umap, Udmap(n);
    Lc.reserve(tot);
    Lmap.reserve(tot);
    for (int u = 0; u < n; ++u) {
        const int i = rowsL[u];
        for (int q = rp[i]; q < diag[i]; ++q) {
            Lc.push_back(col[q]);
            Lmap.push_back(q);
        }
        Loff[u + 1] = (int)Lc.size();
    }
    for (int u = 0; u < n; ++u) {
        const int i = rowsU[u];
        for (int q = diag[i] + 1; q < rp[i + 1]; ++q) {
            Uc.push_back(col[q]);
            Umap.push_back(q);
        }
        Uoff[u + 1] = (int)Uc.size();
        Udmap[u] = diag[i];
    }
    I.nL = (int)Lc.size();
    I.nU = (int)Uc.size();
    constexpr int B = ncgk::kIluB;
    std::vector<int> Lcd((size_t)n * B, 0), Lmapd((size_t)n * B, -1), Ucd((size_t)n * B, 0), Umapd((size_t)n * B, -1);
    for (int u = 0; u < n; ++u) {
        for (int q = 0; q < B && Loff[u] + q < Loff[u + 1]; ++q) {
            Lcd[(size_t)u * B + q] = Lc[Loff[u] + q];
            Lmapd[(size_t)u * B + q] = Lmap[Loff[u] + q];
        }
        for (int q = 0; q < B && Uoff[u] + q < Uoff[u + 1]; ++q) {
            Ucd[(size_t)u * B + q] = Uc[Uoff[u] + q];
            Umapd[(size_t)u * B + q] = Umap[Uoff[u] + q];
        }
    }
    // Batches: consecutive levels of at most kIluRows rows in all, as long as every entry of
    // a row whose column lies in the batch is among its first kIluB (those come from the
    // ring); a level wider than kIluRows rows is a batch of its own.  Then where each
    // padded entry is read: from the ring while it still holds the value (position
    // p >= batch end - kIluRing), else from memory (flag bit 1).  A memory read, and every
    // entry past the kIluB-th or of a wide batch, needs the stores of p's batch made visible
    // (flag bit 0) at the end of a batch between p's batch and the batch before the reader.
    auto batches = [&](const std::vector<int> &lp, int nlev, const std::vector<int> &rows, const std::vector<int> &lv,
                       const std::vector<int> &off, const std::vector<int> &cl, std::vector<int> &cdp,
                       std::vector<int> &bptr, std::vector<int> &bm, std::vector<unsigned char> &sub,
                       std::vector<unsigned char> &flag) {
        std::vector<int> pos(n);
        for (int u = 0; u < n; ++u) pos[rows[u]] = u;
        bptr.assign(1, 0);
        bm.clear();
        sub.assign(n, 0);
        auto fits = [&](int l, int l0) {   // rows of level l reach columns of levels >= l0 only early
            for (int u = lp[l]; u < lp[l + 1]; ++u)
                for (int k = off[u] + B; k < off[u + 1]; ++k)
                    if (lv[cl[k]] >= l0) return false;
            return true;
        };
        constexpr int W = ncgk::kIluRows, RING = ncgk::kIluRing;
        for (int l0 = 0; l0 < nlev;) {
            int l1 = l0 + 1, tot = lp[l0 + 1] - lp[l0];
            if (tot <= W)
                while (l1 < nlev && l1 - l0 < 255 && tot + (lp[l1 + 1] - lp[l1]) <= W && fits(l1, l0)) {
                    tot += lp[l1 + 1] - lp[l1];
                    ++l1;
                }
            for (int u = lp[l0]; u < lp[l1]; ++u) sub[u] = (unsigned char)(lv[rows[u]] - l0);
            bptr.push_back(lp[l1]);
            bm.push_back(l1 - l0);
            l0 = l1;
        }
        const int nb = (int)bm.size();
        std::vector<int> bat_of(n);
        for (int b = 0; b < nb; ++b)
            for (int u = bptr[b]; u < bptr[b + 1]; ++u) bat_of[u] = b;
        flag.assign(nb, 0);
        auto visible = [&](int lo, int hi) {
            for (int f = hi; f >= lo && f >= 0; --f)
                if (flag[f] & 1) return true;
            return false;
        };
        for (int b = 0; b < nb; ++b) {
            const int b0 = bptr[b], w = bptr[b + 1] - b0;
            const bool narrow = w <= W;
            int needB = -1;
            for (int u = b0; u < b0 + w; ++u)
                for (int k = off[u]; k < off[u + 1]; ++k) {
                    const int q = k - off[u], j = cl[k], pj = pos[j];
                    if (narrow && q < B && pj >= b0 + w - RING) cdp[(size_t)u * B + q] = -1 - (pj % RING);
                    else {
                        needB = std::max(needB, bat_of[pj]);
                        if (narrow && q < B) flag[b] |= 2;
                    }
                }
            if (needB >= 0 && !visible(needB, b - 1)) flag[b - 1] |= 1;
        }
    };
    std::vector<int> bpL, bmL, bpU, bmU;
    std::vector<unsigned char> sbL, sbU, fgL, fgU;
    batches(lpL, nL, rowsL, lvL, Loff, Lc, Lcd, bpL, bmL, sbL, fgL);
    batches(lpU, nU, rowsU, lvU, Uoff, Uc, Ucd, bpU, bmU, sbU, fgU);
    I.nbatL = (int)bmL.size();
    I.nbatU = (int)bmU.size();
    if ((rc = upload_ilu(I, s, &I.Lcd, Lcd)) || (rc = upload_ilu(I, s, &I.Lmapd, Lmapd)) ||
        (rc = upload_ilu(I, s, &I.Ucd, Ucd)) || (rc = upload_ilu(I, s, &I.Umapd, Umapd)) ||
        (rc = upload_ilu(I, s, &I.batL_ptr, bpL)) || (rc = upload_ilu(I, s, &I.batL_m, bmL)) ||
        (rc = upload_ilu(I, s, &I.batU_ptr, bpU)) || (rc = upload_ilu(I, s, &I.batU_m, bmU)) ||
        (rc = I.alloc(&I.subL, (size_t)n)) || (rc = I.alloc(&I.subU, (size_t)n)) ||
        (rc = I.alloc(&I.flagL, fgL.size())) || (rc = I.alloc(&I.flagU, fgU.size())) ||
        (rc = I.alloc(&I.Lvd, 2 * (size_t)n * B)) || (rc = I.alloc(&I.Uvd, 2 * (size_t)n * B)))
        return rc;
    NCG_HIP(hipMemcpyAsync(I.subL, sbL.data(), (size_t)n, hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.subU, sbU.data(), (size_t)n, hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.flagL, fgL.data(), fgL.size(), hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(I.flagU, fgU.data(), fgU.size(), hipMemcpyHostToDevice, s));
    if ((rc = upload_ilu(I, s, &I.Loff, Loff)) || (rc = upload_ilu(I, s, &I.Lc, Lc)) ||
        (rc = upload_ilu(I, s, &I.Lmap, Lmap)) || (rc = upload_ilu(I, s, &I.Uoff, Uoff)) ||
        (rc = upload_ilu(I, s, &I.Uc, Uc)) || (rc = upload_ilu(I, s, &I.Umap, Umap)) ||
        (rc = upload_ilu(I, s, &I.Udmap, Udmap)) || (rc = I.alloc(&I.Lv, 2 * (size_t)I.nL)) ||
        (rc = I.alloc(&I.Uv, 2 * (size_t)I.nU)) || (rc = I.alloc(&I.Ud, 2 * (size_t)n)))
        return rc;
    // the factorization's update map: per row the pivot positions of its lower entries,
    // then per lower entry (column k) and entry (column j) the position of u_kj in row k's
    // upper part (the first upper column >= j, when it is j), or -1
    std::vector<long long> uoff(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        const long long R = rp[i + 1] - rp[i], Lw = diag[i] - rp[i];
        uoff[i + 1] = uoff[i] + (R <= 64 ? Lw + Lw * R : 0);
    }
    std::vector<int> upd((size_t)std::max<long long>(uoff[n], 1), -1);
    for (int i = 0; i < n; ++i) {
        const int r0 = rp[i], r1 = rp[i + 1], R = r1 - r0, Lw = diag[i] - r0;
        if (R > 64) continue;
        int *um = upd.data() + uoff[i];
        for (int a = 0; a < Lw; ++a) {
            const int k = col[r0 + a];
            um[a] = diag[k];
            int q = diag[k] + 1;
            const int qe = rp[k + 1];
            for (int e = a + 1; e < R; ++e) {
                const int j = col[r0 + e];
                while (q < qe && col[q] < j) ++q;
                um[Lw + (size_t)a * R + e] = q < qe && col[q] == j ? q : -1;
            }
        }
    }
    long long *d_uoff = nullptr;
    int *d_upd = nullptr;
    if ((rc = I.alloc(&d_uoff, (size_t)n + 1)) || (rc = I.alloc(&d_upd, upd.size()))) return rc;
    NCG_HIP(hipMemcpyAsync(d_uoff, uoff.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(d_upd, upd.data(), sizeof(int) * upd.size(), hipMemcpyHostToDevice, s));
    ncgk::ilu_factor(s, nL, I.levL_ptr, I.levL_rows, I.rp, I.col, I.diag, d_uoff, d_upd, I.val, p->err);
    ncgk::ilu_pack(s, I.nL, I.Lmap, I.val, I.Lv);
    ncgk::ilu_pack(s, I.nU, I.Umap, I.val, I.Uv);
    ncgk::ilu_pack(s, n, I.Udmap, I.val, I.Ud);
    ncgk::ilu_pack(s, n * B, I.Lmapd, I.val, I.Lvd);
    ncgk::ilu_pack(s, n * B, I.Umapd, I.val, I.Uvd);
    NCG_HIP(hipGetLastError());
    // the host vectors go out of scope here: the uploads must have completed
    NCG_HIP(hipStreamSynchronize(s));
    return check_err(p);
}

void ilu_apply(ncg_problem *p, const double *x, double *y)
{
    const DevIlu &I = p->ilu;
    ncgk::IluSweep L, U;
    L.nbat = I.nbatL;
    L.bat_ptr = I.batL_ptr;
    L.bat_m = I.batL_m;
    L.sub = I.subL;
    L.bat_flag = I.flagL;
    L.row = I.levL_rows;
    L.off = I.Loff;
    L.c = I.Lc;
    L.v = I.Lv;
    L.cd = I.Lcd;
    L.vd = I.Lvd;
    U.nbat = I.nbatU;
    U.bat_ptr = I.batU_ptr;
    U.bat_m = I.batU_m;
    U.sub = I.subU;
    U.bat_flag = I.flagU;
    U.row = I.levU_rows;
    U.off = I.Uoff;
    U.c = I.Uc;
    U.v = I.Uv;
    U.cd = I.Ucd;
    U.vd = I.Uvd;
    U.d = I.Ud;
    ncgk::ilu_solve(p->stream, p->n, p->nrhs, L, U, x, y);
}

// Waits for a queued assembly and reads its error flags (the subdivision limits of the
// quadrature); the operator may be used only after this.
int assemble_finish(ncg_problem *p)
{
    if (!p->asm_pending) return 0;
    const int dense = p->asm_pending == 2;
    p->asm_pending = 0;
    int rc = check_err(p);
    if (rc) {
        p->assembled = 0;
        return rc;
    }
    q_collect(p);
    // quadrature evaluations of this assembly, when its kinds have been counted
    const int kinds = dense ? (1 << kQFar) | (1 << kQDenseNear) | (1 << kQDiag)
                            : (1 << kQFmmNear) | (1 << kQFmmAdj) | (1 << kQDiag);
    p->last_evals = 0.0;
    bool known = true;
    for (int q = 0; q < kQuadKinds; ++q)
        if (kinds & (1 << q)) {
            if (p->q_sig[q] == 0) known = false;
            p->last_evals += p->q_evals[q];
        }
    if (!known) p->last_evals = 0.0;
    p->t_asm = elapsed(p->ev_asm);
    if (dense) {
        p->far_ms_sum += elapsed(p->ev_far);
        p->far_calls++;
    }
    return 0;
}

// Queues the evaluation-grid field of the solved load cases on the problem's stream into
// the device buffers out_phi / out_grad; field_collect copies them out and reads the error
// flags (ncg_field is the two in a row).
int field_enqueue(ncg_problem *p, bool want_grad)
{
    const int n = p->n, m = p->m, nr = p->nrhs;
    int rc = ensure_work(p, 8);
    if (rc) return rc;
    double *pot = p->work[0], *vel = p->work[1];
    if ((rc = p->out_phi.ensure(16 * (size_t)nr * m)) || (rc = p->out_grad.ensure(48 * (size_t)nr * m))) return rc;
    double *dphi = p->out_phi.d(), *dgrad = p->out_grad.d();
    record(p, p->ev_field.a);
    ncgk::field_prep(p->stream, n, nr, p->kind, p->adm, p->interior ? -1.0 : 1.0, p->bcval, p->x, pot, vel);
    if (p->plan.method != 0) {
        rc = field_fmm(p, pot, vel, dphi, want_grad ? dgrad : nullptr);
        if (rc) return rc;
    } else {
        const int f32 = p->precision == NCG_MIXED;
        KernelScope ks(p, kKidField + f32, 0.0, kQField);
        ncgk::field(p->stream, p->elements(), p->bm(), p->eval, m, nr, pot, vel, p->ps, p->nps, p->pw, p->npw, dphi,
                    want_grad ? dgrad : nullptr, f32, p->err, nullptr, nullptr, nullptr, nullptr, nullptr,
                    q_counter(p, kQField, 1));
    }
    record(p, p->ev_field.b);
    NCG_HIP(hipGetLastError());
    return 0;
}

int field_collect(ncg_problem *p, double *phi, double *grad)
{
    const int m = p->m, nr = p->nrhs;
    hipError_t e = hipMemcpyAsync(phi, p->out_phi.p, sizeof(cd) * nr * m, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess && grad)
        e = hipMemcpyAsync(grad, p->out_grad.p, sizeof(cd) * 3 * nr * m, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return fail(NCG_E_DEVICE, std::string("ncg_field: ") + hipGetErrorString(e));
    p->t_field = elapsed(p->ev_field);
    return check_err(p);
}

// ---------------------------------------------------------------------------------- block-Jacobi
// Inverse of every finest cluster's self block of the near field (FMM plans only),
// computed once per assembly on first use.
int bj_prepare(ncg_problem *p)
{
    DevFmm &F = p->fmm;
    if (p->plan.method == 0) {
        // dense operator: diagonal blocks of at most 64 consecutive unknowns, sizes balanced
        int rc;
        if (p->dbj_ready) return 0;
        const int n = p->n;
        if (!p->dbj_b0) {
            const int nb = (n + 63) / 64;
            std::vector<int> eb(n), b0(nb), m(nb), id(n);
            std::vector<long long> boff(nb + 1, 0);
            for (int b = 0, s0 = 0; b < nb; ++b) {
                m[b] = n / nb + (b < n % nb ? 1 : 0);
                b0[b] = s0;
                for (int i = s0; i < s0 + m[b]; ++i) eb[i] = b;
                boff[b + 1] = boff[b] + (long long)m[b] * m[b];
                s0 += m[b];
            }
            for (int i = 0; i < n; ++i) id[i] = i;
            if ((rc = upload(p->dbj, p->stream, &p->dbj_elem_blk, eb)) || (rc = upload(p->dbj, p->stream, &p->dbj_b0, b0)) ||
                (rc = upload(p->dbj, p->stream, &p->dbj_m, m)) || (rc = upload(p->dbj, p->stream, &p->dbj_ident, id)) ||
                (rc = upload(p->dbj, p->stream, &p->dbj_boff, boff)))
                return rc;
            p->dbj_nblk = nb;
            p->dbj_total = boff[nb];
            NCG_HIP(hipStreamSynchronize(p->stream));
        }
        p->dbj_f32 = p->precision == NCG_MIXED;
        if ((rc = p->dbj_blk.ensure((p->dbj_f32 ? 8 : 16) * (size_t)p->dbj_total))) return rc;
        ncgk::bj_setup_dense(p->stream, n, p->dbj_nblk, p->dbj_elem_blk, p->dbj_b0, p->dbj_m, p->A, p->dbj_f32,
                             p->dbj_boff, p->dbj_blk.p, p->err);
        NCG_HIP(hipGetLastError());
        if ((rc = check_err(p))) return rc;
        p->dbj_ready = 1;
        return 0;
    }
    if (F.bj_ready) return 0;
    const DevLevel &D = F.lev.back();
    int rc;
    // one flag for both: the near values are FP32 exactly in the mixed form, and bj_setup
    // reads them and writes the inverses in that precision
    F.bj_f32 = p->precision == NCG_MIXED;
    if ((rc = F.bj_blk.ensure((F.bj_f32 ? 8 : 16) * (size_t)F.bj_total))) return rc;
    ncgk::bj_setup(p->stream, p->n, F.nblk, F.elem_blk, D.elem_pos, F.blk_cl, F.blk_b0, F.blk_m, D.cl_ptr, F.near_ptr,
                   F.self_off, F.near_val.p, F.bj_f32, F.boff, F.bj_blk.p, p->err);
    NCG_HIP(hipGetLastError());
    if ((rc = check_err(p))) return rc;
    F.bj_ready = 1;
    return 0;
}

void bj_apply_vec(ncg_problem *p, const double *x, double *y)
{
    if (p->plan.method == 0) {
        ncgk::bj_apply(p->stream, p->n, p->nrhs, p->dbj_nblk, p->dbj_b0, p->dbj_m, p->dbj_ident, p->dbj_boff,
                       p->dbj_blk.p, p->dbj_f32, x, y);
        return;
    }
    const DevFmm &F = p->fmm;
    const DevLevel &D = F.lev.back();
    ncgk::bj_apply(p->stream, p->n, p->nrhs, F.nblk, F.blk_b0, F.blk_m, D.cl_elem, F.boff, F.bj_blk.p, F.bj_f32, x,
                   y);
}

// ---------------------------------------------------------------------------------- CGS
// Right-preconditioned CGS of NC_IterativeSolverCGS (src/NC_CommonFunctions.cpp:1066-1208):
// x0 = 0, unconjugated bilinear form, relative error measured against the residual
// after the first iteration, stop below tol or at max_iterations.
int solve_cgs(ncg_problem *p, const ncg_solver_options *o, ncg_solve_stats *st)
{
    const int n = p->n, nr = p->nrhs;
    const bool ilu = o->precond == NCG_PRECOND_ILU, blk = o->precond == NCG_PRECOND_BLOCK;
    int rc = ensure_work(p, ilu || blk ? 10 : 8);
    if (rc) return rc;
    double *R0 = p->work[0], *R = p->work[1], *P = p->work[2], *V = p->work[3], *Q = p->work[4], *U = p->work[5],
           *UQ = p->work[6], *AUQ = p->work[7], *X = p->x;
    const size_t bytes = sizeof(cd) * (size_t)nr * n;
    const double *d = nullptr;
    const double *B = p->rhs;
    double *T = nullptr;
    if (o->precond == NCG_PRECOND_SCALING) {
        if (p->plan.method == 0) ncgk::row_scaling(p->stream, p->A, p->precision == NCG_MIXED, n, p->dvec);
        else ncgk::row_scaling_csr(p->stream, n, p->fmm.near_ptr, p->fmm.near_val.p, p->precision == NCG_MIXED, p->dvec);
        d = p->dvec;
    } else if (ilu) {
        if ((rc = ilu_setup(p, o->ilu_scanning_degree))) return rc;
        double *Bs = p->work[8];
        T = p->work[9];
        NCG_HIP(hipMemcpyAsync(Bs, p->rhs, bytes, hipMemcpyDeviceToDevice, p->stream));
        ncgk::scale_vec(p->stream, n, nr, p->dvec, Bs);
        B = Bs;
        if (st) {
            st->ilu_nonzeros = p->ilu.nnz;
            st->ilu_levels[0] = p->ilu.nlevL;
            st->ilu_levels[1] = p->ilu.nlevU;
        }
    } else if (blk) {
        if ((rc = bj_prepare(p))) return rc;
        T = p->work[9];
    }
    // operator of the Krylov loop: A diag(d) (scaling), D A (LU)^-1 (ILU) or A
    auto op = [&](const double *in, double *out) {
        if (blk) {
            bj_apply_vec(p, in, T);
            mv(p, 1, T, nullptr, out);
            return;
        }
        if (!ilu) {
            mv(p, 1, in, d, out);
            return;
        }
        ilu_apply(p, in, T);
        mv(p, 1, T, nullptr, out);
        ncgk::scale_vec(p->stream, n, nr, p->dvec, out);
    };
    NCG_HIP(hipMemsetAsync(X, 0, bytes, p->stream));
    for (double *w : {R0, R, P, U}) NCG_HIP(hipMemcpyAsync(w, B, bytes, hipMemcpyDeviceToDevice, p->stream));
    std::vector<int> active(nr, 1);
    NCG_HIP(hipMemcpyAsync(p->active, active.data(), sizeof(int) * nr, hipMemcpyHostToDevice, p->stream));
    std::vector<double> err0(nr, 0.0), err(nr, 1.0);
    std::vector<int> iters(nr, 0);
    std::vector<cd> dv;
    int j;
    for (j = 0; j <= o->max_iterations; ++j) {
        op(P, V);
        {
            const double *u[2] = {R, V}, *v[2] = {R0, R0};
            ncgk::dots(p->stream, n, nr, 2, (const void *const *)u, (const void *const *)v, 0, p->partial, p->dotout);
            ncgk::cgs_alpha(p->stream, nr, j == 0, p->dotout, p->scal, p->active);
        }
        ncgk::cgs_update1(p->stream, n, nr, p->scal, p->active, U, V, Q, UQ);
        op(UQ, AUQ);
        ncgk::cgs_update2(p->stream, n, nr, p->scal, p->active, UQ, AUQ, X, R);
        {
            const double *u[2] = {R, R}, *v[2] = {R, R0};
            rc = dotv(p, 2, u, v, 1, dv);
            if (rc) return rc;
            ncgk::cgs_beta(p->stream, nr, p->dotout, p->scal, p->active);
        }
        bool any = false, changed = false;
        for (int r = 0; r < nr; ++r) {
            if (!active[r]) continue;
            const double rr = sqrt(dv[3 * r].real());
            if (j == 0) err0[r] = rr;
            else err[r] = err0[r] > 0.0 ? rr / err0[r] : 0.0;
            if (err[r] < o->tolerance || j == o->max_iterations) {
                active[r] = 0;
                iters[r] = j;
                changed = true;
            } else {
                any = true;
            }
        }
        if (changed)
            NCG_HIP(hipMemcpyAsync(p->active, active.data(), sizeof(int) * nr, hipMemcpyHostToDevice, p->stream));
        if (!any) break;
        ncgk::cgs_update3(p->stream, n, nr, p->scal, p->active, R, Q, U, P);
    }
    if (d) ncgk::scale_vec(p->stream, n, nr, d, X);
    if (ilu) ilu_apply(p, X, X);
    if (blk) {
        bj_apply_vec(p, X, T);
        NCG_HIP(hipMemcpyAsync(X, T, bytes, hipMemcpyDeviceToDevice, p->stream));
    }
    harvest_mv(p);
    int all = 1;
    for (int r = 0; r < nr; ++r) {
        if (err[r] >= o->tolerance) all = 0;
        if (st && r < 64) {
            st->iterations[r] = iters[r];
            st->relative_residual[r] = err[r];
        }
    }
    if (st) st->converged_all = all;
    return all ? 0 : fail(NCG_E_NOCONV, "Warning: Maximum number of iterations is reached!");
}

// ---------------------------------------------------------------------------------- grouped solves
// A group is one or more assembled problems (frequencies of one mesh or of several) solved
// together: each step of the Krylov loop is one launch per operation for all of them
// (ncgk::g_* kernels, one descriptor per problem or problem level), so the operator applies
// of several frequencies share every launch and the host waits once per step for the
// whole group, not once per problem.  GMRES runs on the device: Gram-Schmidt passes,
// Hessenberg columns, Givens rotations and the convergence test (kernels g_gm_*); the host
// only reads each step's per-problem done flags, kFlagDepth steps behind the launches.
// The group runs on its leader's stream; the members' streams are joined before and after.
constexpr int kGroupMax = 64;
constexpr int kFlagDepth = 2;

// descriptor arrays of one launch sequence, built on the host and uploaded at once into
// the leader's device buffer through its pinned staging buffer
struct Stage {
    ncg_problem *lead;
    std::vector<char> host;
    explicit Stage(ncg_problem *l) : lead(l) {}
    template <typename T>
    size_t put(const std::vector<T> &v)
    {
        const size_t off = (host.size() + 255) & ~(size_t)255;
        host.resize(off + sizeof(T) * std::max<size_t>(v.size(), 1), 0);
        if (!v.empty()) memcpy(host.data() + off, v.data(), sizeof(T) * v.size());
        return off;
    }
    template <typename T>
    const T *dev(size_t off) const
    {
        return (const T *)((const char *)lead->gdesc.p + off);
    }
    // the previous upload from the staging buffer has completed when the stream is idle
    int upload(hipStream_t s)
    {
        NCG_HIP(hipStreamSynchronize(s));
        const size_t bytes = std::max<size_t>(host.size(), 256);
        if (bytes > lead->gstage_cap) {
            const size_t cap = std::max(bytes, 2 * lead->gstage_cap);
            if (lead->gstage) hipHostFree(lead->gstage);
            lead->gstage = nullptr;
            lead->gstage_cap = 0;
            NCG_HIP(hipHostMalloc((void **)&lead->gstage, cap, hipHostMallocDefault));
            lead->gstage_cap = cap;
        }
        int rc = lead->gdesc.ensure(bytes);
        if (rc) return rc;
        memcpy(lead->gstage, host.data(), host.size());
        NCG_HIP(hipMemcpyAsync(lead->gdesc.p, lead->gstage, host.size(), hipMemcpyHostToDevice, s));
        return 0;
    }
};

// prefix sums of per-item work-group counts
std::vector<int> prefix(const std::vector<int> &count)
{
    std::vector<int> p(count.size() + 1, 0);
    for (size_t i = 0; i < count.size(); ++i) p[i + 1] = p[i] + count[i];
    return p;
}

// byte offsets of a member's GMRES state in its gm buffer
struct GmLayout {
    size_t H, cs, sn, g, yv, h, sc, ic, skip, total;
};
GmLayout gm_layout(int nrhs, int m)
{
    GmLayout L;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t r = o;
        o += (bytes + 255) & ~(size_t)255;
        return r;
    };
    L.H = take(16 * (size_t)nrhs * (m + 1) * m);
    L.cs = take(16 * (size_t)nrhs * m);
    L.sn = take(16 * (size_t)nrhs * m);
    L.g = take(16 * (size_t)nrhs * (m + 1));
    L.yv = take(16 * (size_t)nrhs * m);
    L.h = take(16 * ((size_t)2 * nrhs * (m + 1) + 3 * (size_t)nrhs));
    L.sc = take(8 * 4 * (size_t)nrhs);
    L.ic = take(4 * 4 * (size_t)nrhs);
    L.skip = take(4);
    L.total = o;
    return L;
}

// leader's pinned landing area: [kFlagDepth][kGroupMax] flags, [kGroupMax][64] norms,
// [kGroupMax][64] iterations and residual estimates; device twin in the leader's gm buffer
// past the member state (gm_out)
constexpr size_t kPinFlags = 0, kPinNorms = 4096, kPinIts = kPinNorms + 8 * 64 * kGroupMax,
                 kPinRel = kPinIts + 4 * 64 * kGroupMax, kPinBytes = kPinRel + 8 * 64 * kGroupMax;

int ensure_pin(ncg_problem *p)
{
    if (p->gpin) return 0;
    NCG_HIP(hipHostMalloc(&p->gpin, kPinBytes, hipHostMallocDefault));
    for (hipEvent_t &e : p->gev)
        if (!e) NCG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return 0;
}

// One apply y = A x of every member, as grouped launches (FMM members with at most 4 load
// cases per pass: near field, then T and D of every level, then S level by level, so each
// y sums as the single-problem apply does) and single-problem applies (dense operators).
// sig64: FMM signatures t and s in FP64 (the refinement residual).
struct ApplyPlan {
    struct Launch {
        int kind, cls, nr, seg, nitems, blocks, kid;
        size_t items, start;
        double bytes;
    };
    std::vector<Launch> seq;
    std::vector<int> single;   // members applied one by one
    int acc64 = 1, sig64 = 0;
};

// near-field launch shape of a group: its leader's tuning (0 chunk per work-group, 1 cluster)
int lead_tune_near(const std::vector<ncg_problem *> &P) { return P[0]->tune.near_mode; }

void plan_apply(Stage &st, const std::vector<ncg_problem *> &P, const std::vector<const double *> &in,
                const std::vector<double *> &out, const std::vector<int *> &skip, int acc64, int sig64, ApplyPlan &A)
{
    A.seq.clear();
    A.single.clear();
    A.acc64 = acc64;
    A.sig64 = sig64;
    int nrhs = 0;
    std::vector<int> fmm;
    for (size_t i = 0; i < P.size(); ++i) {
        if (P[i]->plan.method == 0 || P[i]->nrhs > 4) A.single.push_back((int)i);
        else fmm.push_back((int)i);
        nrhs = std::max(nrhs, P[i]->nrhs);
    }
    if (fmm.empty()) return;
    const int nr = P[fmm[0]]->nrhs;
    // near field: one launch per value precision
    for (int cls = 0; cls < 2; ++cls) {
        std::vector<ncgk::GNear> it;
        std::vector<int> cnt;
        int seg = 64;
        double bytes = 0;
        for (int i : fmm) {
            ncg_problem *p = P[i];
            const DevFmm &F = p->fmm;
            const int mixed = p->precision == NCG_MIXED;
            if (mixed != cls) continue;
            ncgk::GNear g;
            g.n = p->n;
            g.nch = F.nd_n;
            g.ncl = F.ncl_fine;
            g.cptr = F.nd_cptr;
            g.cb = F.nd_cb;
            g.off = F.nd_off;
            g.w = F.nd_w;
            g.row = F.nd_row;
            g.col = F.near_col;
            g.val = F.near_val.p;
            g.x = in[i];
            g.y = out[i];
            g.skip = skip[i];
            it.push_back(g);
            cnt.push_back(lead_tune_near(P) == 1 ? F.ncl_fine : F.nd_n);
            seg = std::max(seg, std::max(64, std::min(32768 / (16 * nr), (F.near_wmax + 63) / 64 * 64)));
            bytes += (double)F.nnz * (mixed ? 8.0 : 16.0) + 2.0 * nr * p->n * 16.0;
        }
        if (it.empty()) continue;
        const std::vector<int> pre = prefix(cnt);
        A.seq.push_back({0, cls, nr, seg, (int)it.size(), pre.back(), cls ? 1 : 0, st.put(it), st.put(pre), bytes});
    }
    // far field: classes 0 FP32 values + FP32 signatures, 1 FP32 values + FP64 signatures,
    // 2 FP64; T and D of every level in one launch per class, S per level and class
    size_t maxlev = 0;
    for (int i : fmm) maxlev = std::max(maxlev, P[i]->fmm.lev.size());
    auto level_item = [&](int i, size_t l, int &cls) {
        ncg_problem *p = P[i];
        DevLevel &L = p->fmm.lev[l];
        cls = L.f32 ? (sig64 ? 1 : 0) : 2;
        ncgk::GLevel g;
        g.n = p->n;
        g.ncl = L.ncl;
        g.P = L.P;
        g.ntl = L.ntl;
        g.cl_ptr = L.cl_ptr;
        g.cl_elem = L.cl_elem;
        g.fan_ptr = L.fan_ptr;
        g.fan_idx = L.fan_idx;
        g.tl_cl = L.tl_cl;
        g.tl_a0 = L.tl_a0;
        g.tl_m = L.tl_m;
        g.Tv = L.T.p;
        g.Dv = L.D.p;
        g.Sv = L.S.p;
        g.x = in[i];
        g.tw = L.tw.p;
        g.sw = L.sw.p;
        g.y = out[i];
        g.skip = skip[i];
        return g;
    };
    for (int kind = 1; kind <= 2; ++kind)
        for (int cls = 0; cls < 3; ++cls) {
            std::vector<ncgk::GLevel> it;
            std::vector<int> cnt;
            double bytes = 0;
            for (int i : fmm)
                for (size_t l = 0; l < P[i]->fmm.lev.size(); ++l) {
                    int c;
                    ncgk::GLevel g = level_item(i, l, c);
                    if (c != cls) continue;
                    it.push_back(g);
                    // grouped T and D: two sphere points per lane for FP32 values (P is even)
                    const int pts = c < 2 ? 2 : 1;
                    cnt.push_back(g.ncl * ((g.P + 64 * pts - 1) / (64 * pts)));
                    const double sf = c == 2 ? 16.0 : 8.0, sw = c == 0 ? 8.0 : 16.0, n = g.n, Pp = g.P, ncl = g.ncl;
                    const double nfan = P[i]->fmm.lev[l].nfan;
                    bytes += kind == 1 ? n * Pp * sf + nr * n * 16.0 + nr * ncl * Pp * sw
                                       : nfan * Pp * sf + 2.0 * nr * ncl * Pp * sw;
                }
            if (it.empty()) continue;
            const std::vector<int> pre = prefix(cnt);
            const int kid = kind == 1 ? (cls == 2 ? 2 : 3) : (cls == 2 ? 4 : 5);
            A.seq.push_back({kind, cls, nr, 0, (int)it.size(), pre.back(), kid, st.put(it), st.put(pre), bytes});
        }
    for (size_t l = 0; l < maxlev; ++l)
        for (int cls = 0; cls < 3; ++cls) {
            std::vector<ncgk::GLevel> it;
            std::vector<int> cnt;
            double bytes = 0;
            for (int i : fmm) {
                if (l >= P[i]->fmm.lev.size()) continue;
                int c;
                ncgk::GLevel g = level_item(i, l, c);
                if (c != cls) continue;
                it.push_back(g);
                cnt.push_back(g.ntl);
                const double sf = c == 2 ? 16.0 : 8.0, sw = c == 0 ? 8.0 : 16.0, n = g.n, Pp = g.P, ncl = g.ncl;
                bytes += n * Pp * sf + nr * ncl * Pp * sw + 2.0 * nr * n * 16.0;
            }
            if (it.empty()) continue;
            const std::vector<int> pre = prefix(cnt);
            A.seq.push_back({3, cls, nr, 0, (int)it.size(), pre.back(), cls == 2 ? 6 : 7, st.put(it), st.put(pre), bytes});
        }
}

void run_apply(const Stage &st, ncg_problem *lead, const std::vector<ncg_problem *> &P,
               const std::vector<const double *> &in, const std::vector<double *> &out, const ApplyPlan &A)
{
    hipStream_t s = lead->stream;
    for (const ApplyPlan::Launch &L : A.seq) {
        KernelScope ks(lead, kGroupedKid + L.kid, L.bytes);
        switch (L.kind) {
        case 0:
            ncgk::g_near_spmv(s, L.cls, lead->tune.near_mode, L.nr, L.seg, st.dev<ncgk::GNear>(L.items),
                              st.dev<int>(L.start), L.nitems, L.blocks, lead->queue);
            break;
        case 1:
            ncgk::g_apply_T(s, L.cls < 2, L.cls == 1, lead->tune.t_waves, L.nr, st.dev<ncgk::GLevel>(L.items),
                            st.dev<int>(L.start), L.nitems, L.blocks);
            break;
        case 2:
            ncgk::g_apply_D(s, L.cls < 2, L.cls == 1, lead->tune.d_waves, L.nr, st.dev<ncgk::GLevel>(L.items),
                            st.dev<int>(L.start), L.nitems, L.blocks);
            break;
        default:
            ncgk::g_apply_S(s, L.cls < 2, L.cls == 1, lead->tune.s_waves, L.nr, st.dev<ncgk::GLevel>(L.items),
                            st.dev<int>(L.start), L.nitems, L.blocks);
            break;
        }
    }
    // single-problem applies on the group's stream (the members' streams are the leader's
    // while the group runs)
    for (int i : A.single) mv(P[i], A.acc64, in[i], nullptr, out[i]);
}

// block-Jacobi z = M^-1 x of every member (passes of up to 8 load cases)
struct BjPlan {
    struct Launch {
        int f32, nr, nitems, blocks;
        size_t items, start;
    };
    std::vector<Launch> seq;
};
// kstride: x is the member's Krylov basis, offset per Arnoldi step by the device step counter
void plan_bj(Stage &st, const std::vector<ncg_problem *> &P, const std::vector<const double *> &x,
             const std::vector<double *> &y, const std::vector<int *> &skip, BjPlan &B, bool kstride = false)
{
    B.seq.clear();
    const int nrhs = P[0]->nrhs;
    for (int r0 = 0; r0 < nrhs; r0 += 8) {
        const int nr = std::min(8, nrhs - r0);
        for (int f32 = 0; f32 < 2; ++f32) {
            std::vector<ncgk::GBj> it;
            std::vector<int> cnt;
            for (size_t i = 0; i < P.size(); ++i) {
                ncg_problem *p = P[i];
                ncgk::GBj g;
                g.n = p->n;
                g.nr = nr;
                if (p->plan.method == 0) {
                    if (p->dbj_f32 != f32) continue;
                    g.nblk = p->dbj_nblk;
                    g.b0 = p->dbj_b0;
                    g.m = p->dbj_m;
                    g.elem = p->dbj_ident;
                    g.boff = p->dbj_boff;
                    g.minv = p->dbj_blk.p;
                } else {
                    const DevFmm &F = p->fmm;
                    if (F.bj_f32 != f32) continue;
                    g.nblk = F.nblk;
                    g.b0 = F.blk_b0;
                    g.m = F.blk_m;
                    g.elem = F.lev.back().cl_elem;
                    g.boff = F.boff;
                    g.minv = F.bj_blk.p;
                }
                g.x = x[i] + 2 * (size_t)r0 * p->n;
                g.xk = kstride ? 2 * (long long)p->nrhs * p->n : 0;
                g.y = y[i] + 2 * (size_t)r0 * p->n;
                g.skip = skip[i];
                it.push_back(g);
                cnt.push_back(g.nblk);
            }
            if (it.empty()) continue;
            const std::vector<int> pre = prefix(cnt);
            B.seq.push_back({f32, nr, (int)it.size(), pre.back(), st.put(it), st.put(pre)});
        }
    }
}
void run_bj(const Stage &st, hipStream_t s, const BjPlan &B, const int *kp = nullptr)
{
    for (const BjPlan::Launch &L : B.seq)
        ncgk::g_bj_apply(s, L.f32, L.nr, st.dev<ncgk::GBj>(L.items), st.dev<int>(L.start), L.nitems, L.blocks, kp);
}

// vector operation of every member: y = a | a - b | y + a | 0 (complex entries nrhs n)
struct VecOp {
    int op = 0, nitems = 0, blocks = 0;
    size_t items = 0, start = 0;
};
VecOp plan_vec(Stage &st, int op, const std::vector<ncg_problem *> &P, const std::vector<const double *> &a,
               const std::vector<const double *> &b, const std::vector<double *> &y, const std::vector<int *> &skip,
               bool kstride = false)
{
    std::vector<ncgk::GVec> it;
    std::vector<int> cnt;
    for (size_t i = 0; i < P.size(); ++i) {
        ncgk::GVec g;
        g.len = (long long)P[i]->nrhs * P[i]->n;
        g.a = a.empty() ? nullptr : a[i];
        g.ak = kstride ? 2 * g.len : 0;
        g.b = b.empty() ? nullptr : b[i];
        g.y = y[i];
        g.skip = skip.empty() ? nullptr : skip[i];
        it.push_back(g);
        cnt.push_back((int)((g.len + 255) / 256));
    }
    const std::vector<int> pre = prefix(cnt);
    VecOp v;
    v.op = op;
    v.nitems = (int)it.size();
    v.blocks = pre.back();
    v.items = st.put(it);
    v.start = st.put(pre);
    return v;
}
void run_vec(const Stage &st, hipStream_t s, const VecOp &v, const int *kp = nullptr)
{
    ncgk::g_vec(s, v.op, st.dev<ncgk::GVec>(v.items), st.dev<int>(v.start), v.nitems, v.blocks, kp);
}

// squared norms of every member's vectors v into the leader's pinned norms [member][64]
struct NormOp {
    int nitems = 0, blocks = 0;
    size_t items = 0, start = 0;
};
NormOp plan_norms(Stage &st, const std::vector<ncg_problem *> &P, const std::vector<const double *> &v)
{
    std::vector<ncgk::GNorm> it;
    std::vector<int> cnt;
    for (size_t i = 0; i < P.size(); ++i) {
        ncgk::GNorm g;
        g.n = P[i]->n;
        g.nrhs = P[i]->nrhs;
        g.gsb = std::min(256, (P[i]->n + 255) / 256);
        g.v = v[i];
        g.npart = P[i]->npart;
        g.skip = nullptr;
        it.push_back(g);
        cnt.push_back(g.gsb * g.nrhs);
    }
    const std::vector<int> pre = prefix(cnt);
    NormOp o;
    o.nitems = (int)it.size();
    o.blocks = pre.back();
    o.items = st.put(it);
    o.start = st.put(pre);
    return o;
}
// The leader's gm buffer holds, past its own GMRES state, the device side of the gathers:
// [kGroupMax][64] norms and residual estimates (doubles), iterations (ints), and
// [kFlagDepth][kGroupMax] flags.
constexpr size_t kTailNorms = 0, kTailRel = 8 * 64 * kGroupMax, kTailIts = 16 * 64 * kGroupMax,
                 kTailFlags = 20 * 64 * kGroupMax, kTailK = kTailFlags + 4 * kFlagDepth * kGroupMax,
                 kLeadTail = kTailK + 256;
char *lead_tail(ncg_problem *lead) { return (char *)lead->gm.p + lead->gm.cap - kLeadTail; }

// runs the norms and waits for them: out[i][r] = |v_i,r|
int run_norms(const Stage &st, ncg_problem *lead, const NormOp &o, std::vector<std::vector<double>> &out,
              const std::vector<ncg_problem *> &P)
{
    hipStream_t s = lead->stream;
    double *dev = (double *)(lead_tail(lead) + kTailNorms);
    ncgk::g_norms(s, st.dev<ncgk::GNorm>(o.items), st.dev<int>(o.start), o.nitems, o.blocks, dev);
    double *pin = (double *)((char *)lead->gpin + kPinNorms);
    NCG_HIP(hipMemcpyAsync(pin, dev, 8 * 64 * (size_t)o.nitems, hipMemcpyDeviceToHost, s));
    NCG_HIP(hipStreamSynchronize(s));
    out.assign(P.size(), std::vector<double>());
    for (size_t i = 0; i < P.size(); ++i) {
        out[i].resize(P[i]->nrhs);
        for (int r = 0; r < P[i]->nrhs; ++r) out[i][r] = sqrt(pin[64 * i + r]);
    }
    return 0;
}


// ---- captured Arnoldi steps
// A step may be captured when none of its launches is timed on its own (kernel timing
// records events around single launches) and every member applied on its own is a dense
// operator (its apply's launch arguments are few and go into the digest).
bool graphs_allowed(const std::vector<ncg_problem *> &P)
{
    if (!P[0]->graphs || P[0]->graph_fail) return false;
    for (ncg_problem *p : P)
        if (p->ktime || (p->plan.method != 0 && p->nrhs > 4)) return false;
    return true;
}

void key_bj(Fnv &key, const BjPlan &B)
{
    key.add((long long)B.seq.size());
    for (const BjPlan::Launch &L : B.seq) {
        key.add(L.f32);
        key.add(L.nr);
        key.add(L.nitems);
        key.add(L.blocks);
        key.add((long long)L.items);
        key.add((long long)L.start);
    }
}

// the digest of an apply's launch arguments; a dense member applied on its own gets its
// matrix-core buffers sized here, before any capture, so its apply allocates nothing
void key_apply(Fnv &key, ncg_problem *lead, const std::vector<ncg_problem *> &P, const std::vector<const double *> &in,
               const std::vector<double *> &out, const ApplyPlan &A)
{
    key.add((long long)(uintptr_t)lead->queue);
    key.add(lead->tune.near_mode);
    key.add(lead->tune.t_waves);
    key.add(lead->tune.d_waves);
    key.add(lead->tune.s_waves);
    key.add(A.acc64);
    key.add(A.sig64);
    key.add((long long)A.seq.size());
    for (const ApplyPlan::Launch &L : A.seq) {
        key.add(L.kind);
        key.add(L.cls);
        key.add(L.nr);
        key.add(L.seg);
        key.add(L.nitems);
        key.add(L.blocks);
        key.add(L.kid);
        key.add((long long)L.items);
        key.add((long long)L.start);
    }
    for (int i : A.single) {
        ncg_problem *p = P[i];
        const int mixed = p->precision == NCG_MIXED;
        const bool fp64acc = !mixed || A.acc64;
        const size_t el = fp64acc ? sizeof(double) : sizeof(float);
        if (p->nrhs > 4) {
            const ncgk::MfmaPlan mp = ncgk::mfma_plan(p->n, fp64acc, p->tune.mfma_splits);
            if (p->mm_xh.ensure(el * mp.xh_elems) || p->mm_part.ensure(el * mp.part_elems)) key.add(-1);
        }
        key.add((long long)(uintptr_t)p);
        key.add((long long)(uintptr_t)p->A);
        key.add(p->n);
        key.add(p->nrhs);
        key.add(mixed);
        key.add(p->tune.mfma_splits);
        key.add((long long)(uintptr_t)p->mm_xh.p);
        key.add((long long)p->mm_xh.cap);
        key.add((long long)(uintptr_t)p->mm_part.p);
        key.add((long long)p->mm_part.cap);
        key.add((long long)(uintptr_t)in[i]);
        key.add((long long)(uintptr_t)out[i]);
    }
}

constexpr size_t kGraphCache = 32;

// The executable graph of `launch` (the launches it makes on the leader's stream) for the
// digest `key`: from the leader's cache, or recorded now — the launches become a chain of
// kernel nodes (ncgk::set_graph_recorder) without stream capture, so the other lanes'
// allocations and waits meanwhile are unaffected.  nullptr where recording fails: the
// caller then launches one by one, and the leader stops trying.
template <typename F>
hipGraphExec_t cached_graph(ncg_problem *lead, const std::vector<ncg_problem *> &P, unsigned long long key,
                            F &&launch)
{
    std::vector<std::pair<unsigned long long, hipGraphExec_t>> &c = lead->gcache;
    for (size_t i = 0; i < c.size(); ++i)
        if (c[i].first == key) {
            std::rotate(c.begin(), c.begin() + i, c.begin() + i + 1);
            return c[0].second;
        }
    ncgk::GraphRecorder rec;
    hipError_t e = hipGraphCreate(&rec.graph, 0);
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) {
        for (ncg_problem *p : P) p->no_events = 1;
        ncgk::set_graph_recorder(&rec);
        launch();
        ncgk::set_graph_recorder(nullptr);
        for (ncg_problem *p : P) p->no_events = 0;
        e = rec.err;
        if (e == hipSuccess && rec.nodes > 0) e = hipGraphInstantiate(&x, rec.graph, nullptr, nullptr, 0);
        hipGraphDestroy(rec.graph);
    }
    if (e != hipSuccess || !x) {
        (void)hipGetLastError();
        lead->graph_fail = 1;
        return nullptr;
    }
    if (c.size() >= kGraphCache) {
        // the evicted graph may still run on the stream
        hipStreamSynchronize(lead->stream);
        hipGraphExecDestroy(c.back().second);
        c.pop_back();
    }
    c.insert(c.begin(), std::make_pair(key, x));
    lead->g_captured++;
    return x;
}

// Restarted GMRES(m) of every member on its system A x = b (b, x: device [nrhs][n]), x
// starting from zero, per member relative tolerance tol[i]; bj: right block-Jacobi
// preconditioning (x = M^-1 u).  acc64 selects the dense matvec accumulation (FP32 operator:
// 0 = FP32 inner solve).  iters / relres per member and load case.
int gmres_group(const std::vector<ncg_problem *> &P, const std::vector<const double *> &b,
                const std::vector<double *> &x, const std::vector<double> &tol, int restart, int maxit, int acc64,
                int bj, std::vector<std::vector<int>> &iters, std::vector<std::vector<double>> &relres)
{
    const int G = (int)P.size();
    ncg_problem *lead = P[0];
    hipStream_t s = lead->stream;
    const int m = std::max(2, std::min(restart, 200));
    int rc;
    if ((rc = ensure_pin(lead))) return rc;
    std::vector<GmLayout> lay(G);
    for (int i = 0; i < G; ++i) {
        ncg_problem *p = P[i];
        const size_t vlen = (size_t)p->nrhs * p->n, need = (size_t)(m + 1) * vlen * 2;
        if (need > p->basis_cap) {
            NCG_HIP(hipStreamSynchronize(s));
            hipFree(p->basis);
            p->basis = nullptr;
            p->basis_cap = 0;
            if ((rc = dalloc(&p->basis, need))) return rc;
            p->basis_cap = need;
        }
        if ((rc = ensure_work(p, 10))) return rc;
        // multidot partials: at least one slice per basis vector and load case
        if ((size_t)(m + 1) * p->nrhs > p->partial_cap) {
            NCG_HIP(hipStreamSynchronize(s));
            hipFree(p->partial);
            p->partial = nullptr;
            p->partial_cap = 0;
            if ((rc = dalloc(&p->partial, 2 * (size_t)(m + 1) * p->nrhs))) return rc;
            p->partial_cap = (size_t)(m + 1) * p->nrhs;
        }
        lay[i] = gm_layout(p->nrhs, m);
        if ((rc = p->gm.ensure(lay[i].total + (i == 0 ? kLeadTail : 0)))) return rc;
    }
    char *tail = lead_tail(lead);
    int *dflags = (int *)(tail + kTailFlags), *dits = (int *)(tail + kTailIts), *kstep = (int *)(tail + kTailK);
    double *drel = (double *)(tail + kTailRel);
    // descriptors
    Stage st(lead);
    std::vector<ncgk::GGm> gi(G);
    std::vector<int> c_md(G), c_gs(G), c_vec(G);
    std::vector<int *> skip(G);
    std::vector<const double *> Zc(G), Wc(G), xc(G), bc(G);
    std::vector<double *> Z(G), W(G), T(G), xv(G);
    for (int i = 0; i < G; ++i) {
        ncg_problem *p = P[i];
        const GmLayout &L = lay[i];
        char *base = (char *)p->gm.p;
        ncgk::GGm &g = gi[i];
        g.n = p->n;
        g.nrhs = p->nrhs;
        g.gsb = std::min(256, (p->n + 255) / 256);
        g.nsl = (int)std::max<size_t>(
            1, std::min<size_t>((size_t)std::max(1, (p->n + 1023) / 1024), p->partial_cap / ((size_t)(m + 1) * p->nrhs)));
        g.slice = (p->n + g.nsl - 1) / g.nsl;
        g.b = b[i];
        g.x = x[i];
        g.V = p->basis;
        g.W = W[i] = p->work[0];
        g.T = T[i] = p->work[1];
        g.Z = Z[i] = p->work[2];
        g.part = p->partial;
        g.h = base + L.h;
        g.npart = p->npart;
        g.H = base + L.H;
        g.cs = base + L.cs;
        g.sn = base + L.sn;
        g.g = base + L.g;
        g.yv = base + L.yv;
        g.sc = (double *)(base + L.sc);
        g.ic = (int *)(base + L.ic);
        g.skip = skip[i] = (int *)(base + L.skip);
        g.tol = tol[i];
        c_md[i] = p->nrhs * g.nsl;
        c_gs[i] = g.gsb * p->nrhs;
        c_vec[i] = (int)(((size_t)p->nrhs * p->n + 255) / 256);
        Zc[i] = Z[i];
        Wc[i] = W[i];
        xc[i] = x[i];
        bc[i] = b[i];
        xv[i] = x[i];
    }
    const size_t o_gm = st.put(gi);
    const std::vector<int> p_md = prefix(c_md), p_gs = prefix(c_gs), p_vec = prefix(c_vec);
    const size_t o_md = st.put(p_md), o_gs = st.put(p_gs), o_vec = st.put(p_vec);
    const std::vector<int *> noskip(G, nullptr);
    const VecOp v_zero = plan_vec(st, ncgk::kVecZero, P, {}, {}, xv, {});
    const VecOp v_w_b = plan_vec(st, ncgk::kVecCopy, P, bc, {}, W, {});
    const VecOp v_w_res = plan_vec(st, ncgk::kVecSub, P, bc, std::vector<const double *>(T.begin(), T.end()), W, skip);
    const VecOp v_x_add = plan_vec(st, ncgk::kVecAdd, P, std::vector<const double *>(T.begin(), T.end()), {}, xv, {});
    // one Arnoldi step, the same launches for every k (the kernels read k from the device
    // counter kstep): Z = M^-1 V_k (or Z = V_k), W = A Z, two Gram-Schmidt passes, V_k+1,
    // the Hessenberg column and its rotations, the done flags into slot k % kFlagDepth
    BjPlan bj_step;
    VecOp cp_step;
    {
        std::vector<const double *> Vb(G);
        for (int i = 0; i < G; ++i) Vb[i] = P[i]->basis;
        if (bj) plan_bj(st, P, Vb, Z, skip, bj_step, true);
        else cp_step = plan_vec(st, ncgk::kVecCopy, P, Vb, {}, Z, skip, true);
    }
    BjPlan bj_end;
    if (bj) plan_bj(st, P, Zc, T, noskip, bj_end);
    ApplyPlan ap_step, ap_res;
    plan_apply(st, P, Zc, W, skip, acc64, 0, ap_step);
    plan_apply(st, P, xc, T, skip, acc64, 0, ap_res);
    if ((rc = st.upload(s))) return rc;
    const ncgk::GGm *dgm = st.dev<ncgk::GGm>(o_gm);
    const int *d_md = st.dev<int>(o_md), *d_gs = st.dev<int>(o_gs), *d_vec = st.dev<int>(o_vec);
    int *pin_flags = (int *)((char *)lead->gpin + kPinFlags);
    auto arnoldi_step = [&]() {
        if (bj) run_bj(st, s, bj_step, kstep);
        else run_vec(st, s, cp_step, kstep);
        run_apply(st, lead, P, Zc, W, ap_step);
        ncgk::g_multidot(s, kstep, dgm, d_md, G, ncgk::kMdLanes * p_md.back());
        ncgk::g_gs_update(s, 0, kstep, 0, m, dgm, d_gs, G, p_gs.back());
        ncgk::g_multidot(s, kstep, dgm, d_md, G, ncgk::kMdLanes * p_md.back());
        ncgk::g_gs_update(s, 0, kstep, 1, m, dgm, d_gs, G, p_gs.back());
        ncgk::g_gs_normalize(s, 0, kstep, m, dgm, d_gs, G, p_gs.back());
        ncgk::g_gm_rotate(s, 0, kstep, m, dgm, G);
        ncgk::g_gm_flags(s, dgm, G, dflags, kstep, kFlagDepth);
    };
    // the step as a captured graph, where the group allows it (none of its launches is
    // timed on its own)
    hipGraphExec_t step_graph = nullptr;
    if (graphs_allowed(P)) {
        Fnv key;
        key.add((long long)m);
        key.add((long long)bj);
        key.add((long long)acc64);
        key.add((long long)G);
        key.add((long long)(uintptr_t)lead->gdesc.p);
        key.add((long long)(uintptr_t)tail);
        key.add((long long)o_gm);
        key.add((long long)o_md);
        key.add((long long)o_gs);
        key.add((long long)p_md.back());
        key.add((long long)p_gs.back());
        key_bj(key, bj_step);
        key.add((long long)cp_step.nitems);
        key.add((long long)cp_step.blocks);
        key.add((long long)cp_step.items);
        key.add((long long)cp_step.start);
        key_apply(key, lead, P, Zc, W, ap_step);
        step_graph = cached_graph(lead, P, key.h, arnoldi_step);
    }

    run_vec(st, s, v_zero);
    ncgk::g_gm_clear(s, dgm, G);
    int total = 0;
    bool first = true, all_done = false;
    std::vector<int> flags(G, 0);
    while (true) {
        if (first) {
            run_vec(st, s, v_w_b);
        } else {
            run_apply(st, lead, P, xc, T, ap_res);
            run_vec(st, s, v_w_res);
        }
        // |W|, V_0 = W / |W|, beta and the done flags of this cycle; the step counter to 0
        ncgk::g_gs_update(s, 0, nullptr, 1, m, dgm, d_gs, G, p_gs.back());
        ncgk::g_gs_normalize(s, -1, nullptr, m, dgm, d_gs, G, p_gs.back());
        ncgk::g_gm_start(s, m, first ? 1 : 0, dgm, G, kstep);
        first = false;
        // Arnoldi steps; the flags of step k are read after step k + kFlagDepth - 1 is queued
        int launched = 0, checked = 0;
        auto check = [&](int step) -> int {
            NCG_HIP(hipEventSynchronize(lead->gev[step % kFlagDepth]));
            const int *f = pin_flags + (step % kFlagDepth) * kGroupMax;
            all_done = true;
            for (int i = 0; i < G; ++i) {
                flags[i] = f[i];
                if (!f[i]) all_done = false;
            }
            return 0;
        };
        // the flags after the cycle start
        ncgk::g_gm_flags(s, dgm, G, dflags + (kFlagDepth - 1) * kGroupMax);
        NCG_HIP(hipMemcpyAsync(pin_flags + (kFlagDepth - 1) * kGroupMax, dflags + (kFlagDepth - 1) * kGroupMax, 4 * G,
                               hipMemcpyDeviceToHost, s));
        NCG_HIP(hipEventRecord(lead->gev[kFlagDepth - 1], s));
        if ((rc = check(kFlagDepth - 1))) return rc;
        for (int k = 0; k < m && total < maxit && !all_done; ++k) {
            if (step_graph) {
                NCG_HIP(hipGraphLaunch(step_graph, s));
                lead->g_replayed++;
            } else {
                arnoldi_step();
                lead->g_eager++;
            }
            NCG_HIP(hipMemcpyAsync(pin_flags + (k % kFlagDepth) * kGroupMax, dflags + (k % kFlagDepth) * kGroupMax,
                                   4 * G, hipMemcpyDeviceToHost, s));
            NCG_HIP(hipEventRecord(lead->gev[k % kFlagDepth], s));
            ++total;
            launched = k + 1;
            if (launched - checked >= kFlagDepth) {
                if ((rc = check(checked))) return rc;
                ++checked;
            }
        }
        while (!all_done && checked < launched) {
            if ((rc = check(checked))) return rc;
            ++checked;
        }
        // x += M^-1 V y (or V y)
        ncgk::g_gm_solve(s, m, dgm, G);
        if (bj) {
            ncgk::g_gm_update(s, m, 1, dgm, d_vec, G, p_vec.back());
            run_bj(st, s, bj_end);
            run_vec(st, s, v_x_add);
        } else {
            ncgk::g_gm_update(s, m, 0, dgm, d_vec, G, p_vec.back());
        }
        if (all_done || total >= maxit) break;
    }
    // iterations and residual estimates
    int *pin_its = (int *)((char *)lead->gpin + kPinIts);
    double *pin_rel = (double *)((char *)lead->gpin + kPinRel);
    ncgk::g_gm_stats(s, dgm, G, dits, drel);
    NCG_HIP(hipMemcpyAsync(pin_its, dits, 4 * 64 * (size_t)G, hipMemcpyDeviceToHost, s));
    NCG_HIP(hipMemcpyAsync(pin_rel, drel, 8 * 64 * (size_t)G, hipMemcpyDeviceToHost, s));
    NCG_HIP(hipStreamSynchronize(s));
    NCG_HIP(hipGetLastError());
    iters.assign(G, std::vector<int>());
    relres.assign(G, std::vector<double>());
    for (int i = 0; i < G; ++i)
        for (int r = 0; r < P[i]->nrhs; ++r) {
            iters[i].push_back(pin_its[64 * i + r]);
            relres[i].push_back(pin_rel[64 * i + r]);
        }
    return 0;
}

// Mixed precision for every member: FP32 operator, inner GMRES with FP32 matvec
// accumulation and FP64 Arnoldi, outer refinement x += d with the residual b - A x applied
// with FP64 accumulation and, on FMM levels, FP64 signatures.  tol of the inner solve per
// member: what its worst load case still needs (a factor 2 to spare), at most 1/2 and not
// below 1e-5.  iters / relres per member and load case (relres: the refinement residual).
int mixed_group(const std::vector<ncg_problem *> &P, const ncg_solver_options *o, int bj,
                std::vector<std::vector<int>> &iters, std::vector<std::vector<double>> &relres, std::vector<int> &steps)
{
    const int G = (int)P.size();
    ncg_problem *lead = P[0];
    hipStream_t s = lead->stream;
    int rc;
    for (ncg_problem *p : P)
        if ((rc = ensure_work(p, 10))) return rc;
    if ((rc = ensure_pin(lead)) || (rc = lead->gm.ensure(gm_layout(lead->nrhs, std::max(2, std::min(o->restart > 0 ? o->restart : 50, 200))).total + kLeadTail)))
        return rc;
    std::vector<const double *> rhs(G), Rc(G), xc(G), Dc(G);
    std::vector<double *> R(G), D(G), X(G), T(G);
    for (int i = 0; i < G; ++i) {
        rhs[i] = P[i]->rhs;
        Rc[i] = R[i] = P[i]->work[8];
        Dc[i] = D[i] = P[i]->work[9];
        xc[i] = X[i] = P[i]->x;
        T[i] = P[i]->work[1];
    }
    const double target = o->refinement_tolerance > 0 ? o->refinement_tolerance : o->tolerance;
    const int nsteps = o->refinement_steps > 0 ? o->refinement_steps : 5;
    std::vector<std::vector<double>> bnorm, rn;
    iters.assign(G, std::vector<int>());
    relres.assign(G, std::vector<double>());
    steps.assign(G, 0);
    for (int i = 0; i < G; ++i) {
        iters[i].assign(P[i]->nrhs, 0);
        relres[i].assign(P[i]->nrhs, 1.0);
    }
    {
        Stage st(lead);
        const VecOp x0 = plan_vec(st, ncgk::kVecZero, P, {}, {}, X, {});
        const VecOp r0 = plan_vec(st, ncgk::kVecCopy, P, rhs, {}, R, {});
        const NormOp nb = plan_norms(st, P, rhs);
        if ((rc = st.upload(s))) return rc;
        run_vec(st, s, x0);
        run_vec(st, s, r0);
        if ((rc = run_norms(st, lead, nb, bnorm, P))) return rc;
    }
    std::vector<int> active;
    for (int step = 0; step < nsteps + 1; ++step) {
        // relative refinement residual of every member still refining
        std::vector<ncg_problem *> A;
        std::vector<int> idx;
        {
            Stage st(lead);
            std::vector<ncg_problem *> all(P);
            const NormOp nr = plan_norms(st, all, Rc);
            if ((rc = st.upload(s))) return rc;
            if ((rc = run_norms(st, lead, nr, rn, all))) return rc;
        }
        std::vector<double> tol;
        for (int i = 0; i < G; ++i) {
            if (step > 0 && std::find(active.begin(), active.end(), i) == active.end()) continue;
            double worst = 0.0;
            bool ok = true;
            for (int r = 0; r < P[i]->nrhs; ++r) {
                relres[i][r] = bnorm[i][r] > 0 ? rn[i][r] / bnorm[i][r] : 0.0;
                if (!(relres[i][r] < target)) ok = false;
                worst = std::max(worst, relres[i][r]);
            }
            if (ok || step == nsteps) continue;
            A.push_back(P[i]);
            idx.push_back(i);
            tol.push_back(std::min(0.5, std::max(1e-5, 0.5 * target / worst)));
        }
        active = idx;
        if (A.empty()) break;
        std::vector<const double *> bA;
        std::vector<double *> DA, TA, XA;
        std::vector<const double *> DAc, XAc;
        for (int i : idx) {
            bA.push_back(Rc[i]);
            DA.push_back(D[i]);
            DAc.push_back(Dc[i]);
            TA.push_back(T[i]);
            XA.push_back(X[i]);
            XAc.push_back(xc[i]);
            steps[i]++;
        }
        std::vector<std::vector<int>> it;
        std::vector<std::vector<double>> rr;
        if ((rc = gmres_group(A, bA, DA, tol, o->restart > 0 ? o->restart : 50, o->max_iterations, 0, bj, it, rr)))
            return rc;
        for (size_t a = 0; a < idx.size(); ++a)
            for (int r = 0; r < P[idx[a]]->nrhs; ++r) iters[idx[a]][r] += it[a][r];
        // x += d ; r = b - A x with FP64 accumulation and signatures
        Stage st(lead);
        std::vector<const double *> rhsA, TAc(TA.begin(), TA.end());
        std::vector<double *> RA;
        for (int i : idx) {
            rhsA.push_back(rhs[i]);
            RA.push_back(R[i]);
        }
        const VecOp addd = plan_vec(st, ncgk::kVecAdd, A, DAc, {}, XA, {});
        const VecOp res = plan_vec(st, ncgk::kVecSub, A, rhsA, TAc, RA, {});
        ApplyPlan ap;
        plan_apply(st, A, XAc, TA, std::vector<int *>(A.size(), nullptr), 1, 1, ap);
        if ((rc = st.upload(s))) return rc;
        run_vec(st, s, addd);
        run_apply(st, lead, A, XAc, TA, ap);
        run_vec(st, s, res);
    }
    return 0;
}

// Joins the members' streams into the leader's for the duration of a group solve: the
// leader waits for what is queued on each member, the members' launches go to the leader's
// stream, and afterwards each member's stream waits for the group.
struct GroupStreams {
    std::vector<ncg_problem *> P;
    std::vector<hipStream_t> saved;
    explicit GroupStreams(const std::vector<ncg_problem *> &p) : P(p), saved(p.size(), nullptr) {}
    int join()
    {
        for (size_t i = 1; i < P.size(); ++i) {
            if (!P[i]->join) NCG_HIP(hipEventCreateWithFlags(&P[i]->join, hipEventDisableTiming));
            NCG_HIP(hipEventRecord(P[i]->join, P[i]->stream));
            NCG_HIP(hipStreamWaitEvent(P[0]->stream, P[i]->join, 0));
            saved[i] = P[i]->stream;
            P[i]->stream = P[0]->stream;
        }
        return 0;
    }
    void leave()
    {
        if (P.size() < 2) return;
        if (!P[0]->join && hipEventCreateWithFlags(&P[0]->join, hipEventDisableTiming) != hipSuccess) {
            hipStreamSynchronize(P[0]->stream);
            for (size_t i = 1; i < P.size(); ++i)
                if (saved[i]) P[i]->stream = saved[i];
            return;
        }
        hipEventRecord(P[0]->join, P[0]->stream);
        for (size_t i = 1; i < P.size(); ++i)
            if (saved[i]) {
                P[i]->stream = saved[i];
                hipStreamWaitEvent(saved[i], P[0]->join, 0);
            }
    }
};

// The grouped solves are GMRES: the FP64 operator's, or the mixed form's FP64 refinement
// around FP32 GMRES; both take the block-Jacobi or no preconditioner.
int group_options(int precision, const ncg_solver_options *o)
{
    if (precision == NCG_MIXED && o->solver != NCG_SOLVER_GMRES)
        return fail(NCG_E_ARG, "the mixed-precision form solves with GMRES (NCG_SOLVER_GMRES)");
    if (o->solver != NCG_SOLVER_GMRES) return fail(NCG_E_ARG, "a group solve runs GMRES (NCG_SOLVER_GMRES)");
    if (o->precond != NCG_PRECOND_BLOCK && o->precond != NCG_PRECOND_NONE)
        return fail(NCG_E_ARG, "GMRES and the mixed-precision form take the block-Jacobi or no preconditioner");
    return 0;
}

// GMRES (FP64 operator) or mixed-precision refinement of every member of a group.
int solve_group(const std::vector<ncg_problem *> &P, const ncg_solver_options *o, ncg_solve_stats *st)
{
    const int G = (int)P.size();
    ncg_problem *lead = P[0];
    const int bj = o->precond == NCG_PRECOND_BLOCK;
    GroupStreams gs(P);
    int rc = gs.join();
    if (rc) return rc;
    record(lead, lead->ev_solve.a);
    std::vector<std::vector<int>> it;
    std::vector<std::vector<double>> rr;
    std::vector<int> rsteps(G, 0);
    for (ncg_problem *p : P)
        if (bj && !rc) rc = bj_prepare(p);
    const bool mixed = lead->precision == NCG_MIXED;
    double tol = o->tolerance;
    if (!rc) {
        if (mixed) {
            rc = mixed_group(P, o, bj, it, rr, rsteps);
            tol = o->refinement_tolerance > 0 ? o->refinement_tolerance : o->tolerance;
        } else {
            std::vector<const double *> b(G);
            std::vector<double *> x(G);
            for (int i = 0; i < G; ++i) {
                b[i] = P[i]->rhs;
                x[i] = P[i]->x;
            }
            rc = gmres_group(P, b, x, std::vector<double>(G, o->tolerance), o->restart > 0 ? o->restart : 50,
                             o->max_iterations, 1, bj, it, rr);
        }
    }
    record(lead, lead->ev_solve.b);
    for (ncg_problem *p : P) harvest_mv(p);
    gs.leave();
    if (rc) return rc;
    const double ms = elapsed(lead->ev_solve);
    int all = 1;
    for (int i = 0; i < G; ++i) {
        ncg_problem *p = P[i];
        int conv = 1;
        for (int r = 0; r < p->nrhs; ++r)
            if (!(rr[i][r] < tol)) conv = 0;
        if (!conv) all = 0;
        p->t_solve = ms;
        p->solved = 1;
        if (st) {
            ncg_solve_stats &s = st[i];
            memset(&s, 0, sizeof(s));
            for (int r = 0; r < p->nrhs && r < 64; ++r) {
                s.iterations[r] = it[i][r];
                s.relative_residual[r] = rr[i][r];
            }
            s.converged_all = conv;
            s.refinement_steps = rsteps[i];
            s.ms_assemble = p->t_asm;
            s.ms_solve = ms;
            s.ms_field = p->t_field;
        }
    }
    if (all) return 0;
    return fail(NCG_E_NOCONV, mixed ? "mixed-precision refinement did not reach its tolerance"
                                    : "GMRES: maximum number of iterations reached");
}

}  // namespace

// ==================================================================================== API

extern "C" {

const char *ncg_last_error(void) { return g_err.c_str(); }

int ncg_device_count(int32_t *count)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return 0;
}

int ncg_problem_create(const ncg_geometry *g, int32_t device, ncg_problem **out)
{
    if (!g || !out || g->n_elements <= 0 || !g->nodes_per_element || !g->node_xyz || !g->bc_kind)
        return fail(NCG_E_ARG, "ncg_problem_create: missing geometry");
    if (g->n_eval > 0 && !g->eval_xyz) return fail(NCG_E_ARG, "ncg_problem_create: missing evaluation points");
    const int n = g->n_elements;
    std::vector<int8_t> nvh(n), kh(n);
    for (int e = 0; e < n; ++e) {
        if (g->nodes_per_element[e] != 3 && g->nodes_per_element[e] != 4)
            return fail(NCG_E_ARG, "number of nodes of an element must be 3 or 4!");
        if (g->bc_kind[e] < 0 || g->bc_kind[e] > 2) return fail(NCG_E_ARG, "bc_kind must be 0, 1 or 2");
        nvh[e] = (int8_t)g->nodes_per_element[e];
        kh[e] = (int8_t)g->bc_kind[e];
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NCG_E_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev) return fail(NCG_E_ARG, "device index out of range");
    NCG_HIP(hipSetDevice(device));
    if (g->method.method_bem != 0 && g->method.method_bem != 1 && g->method.method_bem != 4)
        return fail(NCG_E_ARG, "methodBEM_ must be in {0,1,4}!");
    if (g->method.min_expansion_terms != 0 && (g->method.min_expansion_terms < 5 || g->method.min_expansion_terms > 10))
        return fail(NCG_E_ARG, "minExpansionTermsFMM_  must be in [5 10]!");
    ncg_problem *p = new ncg_problem();
    p->device = device;
    build_fmm_mesh(g, p->fm);
    p->fo = fmm_options(g);
    p->n = n;
    p->m = g->n_eval > 0 ? g->n_eval : 0;
    p->h = g->harmonic_time_factor == -1.0 ? -1.0 : 1.0;
    p->interior = g->interior ? 1 : 0;
    p->kind_h = kh;
    p->all_velocity = std::all_of(kh.begin(), kh.end(), [](int8_t k) { return k == NCG_BC_VELOCITY; });
    auto bail = [&](int rc) {
        ncg_problem_destroy(p);
        return rc;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NCG_E_DEVICE, "hipStreamCreate failed"));
    for (EventPair *e : {&p->ev_asm, &p->ev_far, &p->ev_solve, &p->ev_field}) {
        hipEventCreate(&e->a);
        hipEventCreate(&e->b);
    }
    double glx[36], glw[36];
    gauss_tables(glx, glw);
    if (ncgk::set_gauss_tables(glx, glw) != hipSuccess) return bail(fail(NCG_E_DEVICE, "constant table upload failed"));
    std::vector<double> Xh((size_t)12 * n, 0.0);
    for (int e = 0; e < n; ++e)
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 3; ++d) Xh[(size_t)(3 * a + d) * n + e] = g->node_xyz[(size_t)e * 12 + 3 * a + d];
    int rc;
    if ((rc = dalloc(&p->X, 12 * (size_t)n)) || (rc = dalloc(&p->cen, 3 * (size_t)n)) ||
        (rc = dalloc(&p->nrm, 3 * (size_t)n)) || (rc = dalloc(&p->area, n)) || (rc = dalloc(&p->jac, n)) ||
        (rc = dalloc(&p->nv, n)) || (rc = dalloc(&p->kind, n)) || (rc = dalloc(&p->eval, 3 * (size_t)p->m)) ||
        (rc = dalloc(&p->adm, 2 * (size_t)n)) || (rc = dalloc(&p->err, 1)) || (rc = dalloc(&p->evals, (size_t)kQuadKinds)) ||
        (rc = dalloc(&p->dvec, n)) || (rc = dalloc(&p->active, 64)) || (rc = dalloc(&p->queue, 320)))
        return bail(rc);
    const int nb = ncgk::dot_blocks(n);
    // partial sums of the dot kernels: 3 dots x 64 load cases per block, and at least
    // 256 slices for 8 basis vectors of 2 load cases in GMRES projections
    p->partial_cap = std::max<size_t>((size_t)nb * 3 * 64, 4096);
    if ((rc = dalloc(&p->partial, 2 * p->partial_cap)) || (rc = dalloc(&p->dotout, (size_t)2 * 3 * 64 * 256)) ||
        (rc = dalloc(&p->scal, (size_t)2 * 8 * 64 * 2)) || (rc = dalloc(&p->npart, (size_t)64 * 256)))
        return bail(rc);
    hipStream_t s = p->stream;
    hipMemcpyAsync(p->X, Xh.data(), sizeof(double) * Xh.size(), hipMemcpyHostToDevice, s);
    hipMemcpyAsync(p->nv, nvh.data(), n, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(p->kind, kh.data(), n, hipMemcpyHostToDevice, s);
    if (p->m) hipMemcpyAsync(p->eval, g->eval_xyz, sizeof(double) * 3 * p->m, hipMemcpyHostToDevice, s);
    hipMemsetAsync(p->err, 0, sizeof(int), s);
    hipMemsetAsync(p->queue, 0, 320 * sizeof(int), s);
    ncgk::element_geometry(s, n, p->X, p->nv, p->cen, p->nrm, p->area, p->jac);
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(fail(NCG_E_DEVICE, std::string("geometry kernels: ") + hipGetErrorString(e)));
    *out = p;
    return 0;
}

int ncg_problem_destroy(ncg_problem *p)
{
    if (!p) return 0;
    hipSetDevice(p->device);
    if (p->stream) hipStreamSynchronize(p->stream);
    for (void *q : {(void *)p->X, (void *)p->cen, (void *)p->nrm, (void *)p->area, (void *)p->jac, (void *)p->eval,
                    (void *)p->nv, (void *)p->kind, (void *)p->near_rows, (void *)p->near_cols, p->A, (void *)p->adm,
                    (void *)p->bcval, (void *)p->rhs, (void *)p->rhs_src, (void *)p->x, (void *)p->ps, (void *)p->pw,
                    (void *)p->src, (void *)p->err, (void *)p->evals, (void *)p->scal, (void *)p->partial,
                    (void *)p->dotout, (void *)p->dvec, (void *)p->active, (void *)p->basis, (void *)p->npart,
                    (void *)p->queue})
        if (q) hipFree(q);
    for (double *w : p->work) hipFree(w);
    p->fmm.release();
    p->ilu.release();
    p->dbj.release();
    p->dbj_blk.release();
    p->mm_xh.release();
    p->mm_part.release();
    p->out_phi.release();
    p->out_grad.release();
    for (EventPair *e : {&p->ev_asm, &p->ev_far, &p->ev_solve, &p->ev_field}) {
        if (e->a) hipEventDestroy(e->a);
        if (e->b) hipEventDestroy(e->b);
    }
    for (EventPair &e : p->ev_mv) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    for (EventPair &e : p->kpool) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    for (hipEvent_t e : p->hev)
        if (e) hipEventDestroy(e);
    if (p->hpin) hipHostFree(p->hpin);
    p->gm.release();
    p->gdesc.release();
    if (p->gstage) hipHostFree(p->gstage);
    if (p->gpin) hipHostFree(p->gpin);
    for (hipEvent_t e : p->gev)
        if (e) hipEventDestroy(e);
    if (p->join) hipEventDestroy(p->join);
    if (p->stream) hipStreamSynchronize(p->stream);
    for (auto &g : p->gcache) hipGraphExecDestroy(g.second);
    p->gcache.clear();
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

}  // extern "C"

namespace {

// Queues the assembly of one frequency step on the problem's stream; with wait, also
// waits for it and reads its error flags (ncg_assemble), else the flags are read by
// assemble_finish before the operator is used.
int assemble_impl(ncg_problem *p, const ncg_frequency *f, int32_t precision, bool wait)
{
    if (!p || !f) return fail(NCG_E_ARG, "ncg_assemble: null argument");
    if (!(f->k > 0.0)) return fail(NCG_E_ARG, "ncg_assemble: wave number must be positive");
    if (f->n_rhs < 1 || f->n_rhs > 64) return fail(NCG_E_ARG, "ncg_assemble: n_rhs must be in [1, 64]");
    if (precision != NCG_FP64 && precision != NCG_MIXED) return fail(NCG_E_ARG, "ncg_assemble: unknown precision");
    if ((f->n_point_sources > 0 && !f->point_sources) || (f->n_plane_waves > 0 && !f->plane_waves))
        return fail(NCG_E_ARG, "ncg_assemble: missing source data");
    NCG_HIP(hipSetDevice(p->device));
    const int n = p->n, nr = f->n_rhs;
    hipStream_t s = p->stream;
    p->k = f->k;
    p->nrhs = nr;
    p->precision = precision;
    p->assembled = p->solved = 0;
    int rc;
    if ((rc = ensure_vectors(p, nr))) return rc;
    // method and FMM plan of this frequency (NC_DeclareArrays, src/NC_Addresses.cpp:58-287)
    if (!ncgf::make_plan(p->fm, p->fo, f->k, p->plan)) return fail(NCG_E_ARG, p->plan.error);
    const size_t abytes = p->plan.method ? 0 : (size_t)n * n * (precision == NCG_MIXED ? 8 : 16);
    if (abytes > p->A_cap) {
        hipFree(p->A);
        p->A = nullptr;
        p->A_cap = 0;
        if (hipMalloc(&p->A, abytes) != hipSuccess) return fail(NCG_E_MEMORY, "cannot allocate the dense operator");
        p->A_cap = abytes;
    }
    // admittances only where the element kind says so; prescribed values per load case
    std::vector<double> admh(2 * (size_t)n, 0.0);
    if (f->admittance)
        for (int e = 0; e < n; ++e)
            if (p->kind_h[e] == NCG_BC_ADMITTANCE) { admh[2 * e] = f->admittance[2 * e]; admh[2 * e + 1] = f->admittance[2 * e + 1]; }
    std::vector<double> bch(2 * (size_t)nr * n, 0.0);
    if (f->bc_value) memcpy(bch.data(), f->bc_value, sizeof(double) * bch.size());
    // source columns: elements with a nonzero prescribed value in some load case
    // (the Ifcrh3 test of src/NC_EquationSystem.cpp:268-279)
    std::vector<int> srch;
    for (int e = 0; e < n; ++e)
        for (int r = 0; r < nr; ++r) {
            const double re = bch[2 * ((size_t)r * n + e)], im = bch[2 * ((size_t)r * n + e) + 1];
            if (sqrt(re * re + im * im) > 1.0e-14) { srch.push_back(e); break; }
        }
    if ((int)srch.size() > p->nsrc || !p->src) {
        hipFree(p->src);
        if ((rc = dalloc(&p->src, srch.size() + 1))) return rc;
    }
    p->nsrc = (int)srch.size();
    auto upload_src = [&](double **dst, int *cap, const double *srcp, int cnt) -> int {
        if (cnt > *cap || !*dst) {
            hipFree(*dst);
            int r2 = dalloc(dst, 6 * (size_t)std::max(cnt, 1));
            if (r2) return r2;
            *cap = std::max(cnt, 1);
        }
        if (cnt) hipMemcpyAsync(*dst, srcp, sizeof(double) * 6 * cnt, hipMemcpyHostToDevice, s);
        return 0;
    };
    for (int q = 0; q < f->n_point_sources; ++q)
        if (f->point_sources[6 * q + 5] < 0 || f->point_sources[6 * q + 5] >= nr)
            return fail(NCG_E_ARG, "point source rhs index out of range");
    for (int q = 0; q < f->n_plane_waves; ++q)
        if (f->plane_waves[6 * q + 5] < 0 || f->plane_waves[6 * q + 5] >= nr)
            return fail(NCG_E_ARG, "plane wave rhs index out of range");
    if ((rc = upload_src(&p->ps, &p->ps_cap, f->point_sources, f->n_point_sources))) return rc;
    if ((rc = upload_src(&p->pw, &p->pw_cap, f->plane_waves, f->n_plane_waves))) return rc;
    p->nps = f->n_point_sources;
    p->npw = f->n_plane_waves;
    NCG_HIP(hipMemcpyAsync(p->adm, admh.data(), sizeof(double) * admh.size(), hipMemcpyHostToDevice, s));
    NCG_HIP(hipMemcpyAsync(p->bcval, bch.data(), sizeof(double) * bch.size(), hipMemcpyHostToDevice, s));
    if (p->nsrc) NCG_HIP(hipMemcpyAsync(p->src, srch.data(), sizeof(int) * p->nsrc, hipMemcpyHostToDevice, s));
    const ncgk::BMParams b = p->bm();
    const int mixed = precision == NCG_MIXED;
    if (p->plan.method != 0) {
        record(p, p->ev_asm.a);
        record(p, p->ev_far.a);
        if ((rc = assemble_fmm(p))) return rc;
        record(p, p->ev_far.b);
        record(p, p->ev_asm.b);
        NCG_HIP(hipGetLastError());
        p->last_evals = 0.0;
        p->assembled = 1;
        p->asm_pending = 1;
        return wait ? assemble_finish(p) : 0;
    }
    if ((rc = ensure_near_table(p))) return rc;
    record(p, p->ev_asm.a);
    record(p, p->ev_far.a);
    p->dbj_ready = 0;
    // quadrature evaluations are counted once per problem, in the first assembly with
    // kernel timing on (measurement)
    {
        KernelScope ks(p, kKidFar + mixed, 0.0, kQFar);
        ncgk::assemble_far(s, p->elements(), b, p->adm, p->A, mixed,
                           p->tune.assemble_rows > 0 ? p->tune.assemble_rows : 32, q_counter(p, kQFar, 1));
    }
    record(p, p->ev_far.b);
    {
        KernelScope ks(p, kKidDenseNear, 0.0, kQDenseNear);
        ncgk::assemble_near(s, p->elements(), b, p->adm, p->near_rows, p->near_cols, p->nnear, p->A, mixed, p->err,
                            q_counter(p, kQDenseNear, 1));
    }
    if (p->nsrc) ncgk::rhs_sources(s, p->elements(), b, p->src, p->nsrc, p->bcval, nr, p->rhs_src, p->err);
    {
        KernelScope ks(p, kKidDiag, 0.0, kQDiag);
        ncgk::assemble_diag(s, p->elements(), b, p->adm, p->bcval, nr, p->ps, p->nps, p->pw, p->npw, p->rhs_src,
                            p->nsrc > 0, p->A, mixed, p->rhs, nullptr, q_counter(p, kQDiag, 1));
    }
    record(p, p->ev_asm.b);
    NCG_HIP(hipGetLastError());
    p->last_evals = 0.0;
    p->assembled = 1;
    p->asm_pending = 2;   // dense: also the far-field kernel timing
    return wait ? assemble_finish(p) : 0;
}

}  // namespace

extern "C" {

int ncg_assemble(ncg_problem *p, const ncg_frequency *f, int32_t precision)
{
    return assemble_impl(p, f, precision, true);
}


int ncg_solve(ncg_problem *p, const ncg_solver_options *o, ncg_solve_stats *st)
{
    if (!p || !o) return fail(NCG_E_ARG, "ncg_solve: null argument");
    if (!p->assembled) return fail(NCG_E_STATE, "ncg_solve: assemble first");
    if (o->max_iterations < 0) return fail(NCG_E_ARG, "ncg_solve: negative iteration limit");
    NCG_HIP(hipSetDevice(p->device));
    int rcf = assemble_finish(p);
    if (rcf) return rcf;
    if (o->precond < NCG_PRECOND_ILU || o->precond > NCG_PRECOND_BLOCK)
        return fail(NCG_E_ARG, "ncg_solve: precond must be NCG_PRECOND_ILU, _SCALING, _NONE or _BLOCK");
    NCG_HIP(hipSetDevice(p->device));
    if (st) memset(st, 0, sizeof(*st));
    record(p, p->ev_solve.a);
    int rc;
    if (o->solver == NCG_SOLVER_LU) {
        if (p->precision != NCG_FP64) return fail(NCG_E_ARG, "the direct solver needs the FP64 operator");
        if (p->plan.method != 0) return fail(NCG_E_ARG, "the direct solver needs the dense operator (TBEM)");
        const size_t bytes = sizeof(cd) * (size_t)p->nrhs * p->n;
        NCG_HIP(hipMemcpyAsync(p->x, p->rhs, bytes, hipMemcpyDeviceToDevice, p->stream));
        ncgk::lu_factor(p->stream, p->A, p->n);
        ncgk::lu_solve(p->stream, p->A, p->n, p->nrhs, p->x);
        NCG_HIP(hipGetLastError());
        p->assembled = 0;   // the operator now holds its factors
        rc = 0;
        if (st) st->converged_all = 1;
    } else if (p->precision == NCG_MIXED || o->solver == NCG_SOLVER_GMRES) {
        if ((rc = group_options(p->precision, o))) return rc;
        std::vector<ncg_problem *> one(1, p);
        return solve_group(one, o, st);
    } else {
        rc = solve_cgs(p, o, st);
    }
    record(p, p->ev_solve.b);
    NCG_HIP(hipStreamSynchronize(p->stream));
    p->t_solve = elapsed(p->ev_solve);
    if (st) {
        st->ms_assemble = p->t_asm;
        st->ms_solve = p->t_solve;
        st->ms_field = p->t_field;
    }
    if (rc == 0 || rc == NCG_E_NOCONV) p->solved = 1;
    return rc;
}

static int valid_waves(int w) { return w == 1 || w == 2 || w == 4 || w == 8; }

int ncg_set_tuning(ncg_problem *p, const ncg_tuning *t)
{
    if (!p || !t) return fail(NCG_E_ARG, "ncg_set_tuning: null argument");
    if (!valid_waves(t->t_waves) || !valid_waves(t->d_waves) || !valid_waves(t->s_waves))
        return fail(NCG_E_ARG, "ncg_set_tuning: waves per work-group must be 1, 2, 4 or 8");
    if (t->near_mode < 0 || t->near_mode > 3) return fail(NCG_E_ARG, "ncg_set_tuning: near_mode must be 0, 1, 2 or 3");
    if (t->mfma_splits < 0 || t->mfma_splits > 16) return fail(NCG_E_ARG, "ncg_set_tuning: mfma_splits must be 0..16");
    if (t->assemble_rows != 0 && (t->assemble_rows < 8 || t->assemble_rows > 256))
        return fail(NCG_E_ARG, "ncg_set_tuning: assemble_rows must be 0 or 8..256");
    if (t->near_assembly != 0 && t->near_assembly != 1)
        return fail(NCG_E_ARG, "ncg_set_tuning: near_assembly must be 0 or 1");
    p->tune = *t;
    return 0;
}

int ncg_set_graphs(ncg_problem *p, int32_t on)
{
    if (!p) return fail(NCG_E_ARG, "ncg_set_graphs: null problem");
    p->graphs = on ? 1 : 0;
    p->graph_fail = 0;
    return 0;
}

int ncg_graph_stats(ncg_problem *p, int64_t *out)
{
    if (!p || !out) return fail(NCG_E_ARG, "ncg_graph_stats: null argument");
    out[0] = p->g_captured;
    out[1] = p->g_replayed;
    out[2] = p->g_eager;
    out[3] = (int64_t)p->gcache.size();
    return 0;
}

int ncg_get_tuning(ncg_problem *p, ncg_tuning *t)
{
    if (!p || !t) return fail(NCG_E_ARG, "ncg_get_tuning: null argument");
    *t = p->tune;
    return 0;
}

}  // extern "C"

namespace {

// summed kernel time (ms) of the listed timing ids on p since the last reset, then reset
double take_ms(ncg_problem *p, std::initializer_list<int> ids)
{
    harvest_mv(p);
    double t = 0.0;
    for (int id : ids) t += p->k_ms[id];
    for (int id = 0; id < kNumKernels; ++id) {
        p->k_ms[id] = p->k_bytes[id] = p->k_flops[id] = 0.0;
        p->k_calls[id] = 0;
    }
    return t;
}

// A candidate replaces the choice so far only when it is faster by more than kTuneMargin:
// the first candidate of each kind is the default shape, so timing noise between nearly
// equal shapes leaves the default in place and the choice is the same from run to run.
constexpr double kTuneMargin = 0.03;
inline bool tune_better(double t, double tbest) { return t < (1.0 - kTuneMargin) * tbest; }

int autotune_group(const std::vector<ncg_problem *> &P, int reps, ncg_tuning &best)
{
    ncg_problem *lead = P[0];
    hipStream_t s = lead->stream;
    const int G = (int)P.size();
    int rc;
    best = lead->tune;
    std::vector<int> keep(G);
    for (int i = 0; i < G; ++i) {
        keep[i] = P[i]->ktime;
        P[i]->ktime = 1;
        if ((rc = ensure_work(P[i], 8))) return rc;
        take_ms(P[i], {});
    }
    std::vector<const double *> in(G);
    std::vector<double *> out(G);
    for (int i = 0; i < G; ++i) {
        in[i] = P[i]->x;
        out[i] = P[i]->work[1];
    }
    const std::vector<int *> noskip(G, nullptr);
    const bool fmm = std::any_of(P.begin(), P.end(), [](ncg_problem *p) { return p->plan.method != 0 && p->nrhs <= 4; });
    if (fmm) {
        // near-field mode (changes no value)
        double tbest = 1e300;
        for (int mode = 0; mode < 4; ++mode) {
            lead->tune.near_mode = mode;
            Stage st(lead);
            ApplyPlan ap;
            plan_apply(st, P, in, out, noskip, 1, 0, ap);
            if ((rc = st.upload(s))) return rc;
            run_apply(st, lead, P, in, out, ap);   // warm
            take_ms(lead, {});
            for (int r = 0; r < reps; ++r) run_apply(st, lead, P, in, out, ap);
            const double t = take_ms(lead, {kGroupedKid + 0, kGroupedKid + 1});
            if (tune_better(t, tbest)) {
                tbest = t;
                best.near_mode = mode;
            }
        }
        lead->tune.near_mode = best.near_mode;
        // near-field quadrature mapping (changes no value): the leader's regular near field
        // assembled again in both mappings
        {
            ncg_problem *p = lead;
            DevFmm &F = p->fmm;
            const DevLevel &DF = F.lev.back();
            EventPair ev = p->ev_far;
            double tq = 1e300;
            for (int mode = 0; mode < 2; ++mode) {
                record(p, ev.a);
                for (int r = 0; r < std::max(1, reps / 2); ++r)
                    ncgk::fmm_near(s, p->elements(), p->bm(), p->adm, DF.chunks(), DF.cl_ptr, DF.cl_elem, F.near_ptr,
                                   F.near_col, F.near_wmax, F.near_val.p, p->precision == NCG_MIXED, nullptr, mode);
                record(p, ev.b);
                NCG_HIP(hipStreamSynchronize(s));
                const double t = elapsed(ev);
                if (tune_better(t, tq)) {
                    tq = t;
                    best.near_assembly = mode;
                }
            }
        }
        // waves per work-group of T, D and S, each kind on its own
        double tT = 1e300, tD = 1e300, tS = 1e300;
        Stage st(lead);
        ApplyPlan ap;
        plan_apply(st, P, in, out, noskip, 1, 0, ap);
        if ((rc = st.upload(s))) return rc;
        for (int w : {8, 4, 2, 1}) {
            lead->tune.t_waves = lead->tune.d_waves = lead->tune.s_waves = w;
            run_apply(st, lead, P, in, out, ap);
            take_ms(lead, {});
            for (int r = 0; r < reps; ++r) run_apply(st, lead, P, in, out, ap);
            harvest_mv(lead);
            const double a = lead->k_ms[kGroupedKid + 2] + lead->k_ms[kGroupedKid + 3];
            const double b = lead->k_ms[kGroupedKid + 4] + lead->k_ms[kGroupedKid + 5];
            const double c = lead->k_ms[kGroupedKid + 6] + lead->k_ms[kGroupedKid + 7];
            take_ms(lead, {});
            if (tune_better(a, tT)) { tT = a; best.t_waves = w; }
            if (tune_better(b, tD)) { tD = b; best.d_waves = w; }
            if (tune_better(c, tS)) { tS = c; best.s_waves = w; }
        }
    }
    // dense operators: matrix-core splits (more than 4 load cases) and far-field rows
    std::vector<ncg_problem *> dense;
    for (ncg_problem *p : P)
        if (p->plan.method == 0) dense.push_back(p);
    if (!dense.empty()) {
        if (dense[0]->nrhs > 4) {
            double tbest = 1e300;
            for (int sp : {0, 1, 2, 4, 8, 16}) {
                for (ncg_problem *p : dense) p->tune.mfma_splits = sp;
                double t = 0.0;
                for (ncg_problem *p : dense) {
                    mv(p, 0, p->x, nullptr, p->work[1]);
                    take_ms(p, {});
                    for (int r = 0; r < reps; ++r) mv(p, 0, p->x, nullptr, p->work[1]);
                    t += take_ms(p, {10, 11});
                }
                if (tune_better(t, tbest)) {
                    tbest = t;
                    best.mfma_splits = sp;
                }
            }
        }
        double tbest = 1e300;
        ncg_problem *p = dense[0];
        EventPair ev = p->ev_far;
        for (int rows : {32, 16, 64, 128}) {
            record(p, ev.a);
            for (int r = 0; r < std::max(1, reps / 2); ++r)
                ncgk::assemble_far(s, p->elements(), p->bm(), p->adm, p->A, p->precision == NCG_MIXED, rows, nullptr);
            record(p, ev.b);
            NCG_HIP(hipStreamSynchronize(s));
            const double t = elapsed(ev);
            if (tune_better(t, tbest)) {
                tbest = t;
                best.assemble_rows = rows;
            }
        }
    }
    NCG_HIP(hipGetLastError());
    NCG_HIP(hipStreamSynchronize(s));
    best.tuned = 1;
    for (int i = 0; i < G; ++i) {
        P[i]->tune = best;
        P[i]->ktime = keep[i];
        take_ms(P[i], {});
    }
    return 0;
}

}  // namespace

extern "C" {

int ncg_autotune(ncg_problem *const *probs, int32_t count, int32_t reps, ncg_tuning *out)
{
    if (!probs || count < 1 || count > kGroupMax) return fail(NCG_E_ARG, "ncg_autotune: 1..64 problems");
    std::vector<ncg_problem *> P(probs, probs + count);
    for (ncg_problem *p : P) {
        if (!p) return fail(NCG_E_ARG, "ncg_autotune: null problem");
        if (!p->assembled) return fail(NCG_E_STATE, "ncg_autotune: assemble every problem first");
        if (p->device != P[0]->device || p->nrhs != P[0]->nrhs || p->precision != P[0]->precision)
            return fail(NCG_E_ARG, "ncg_autotune: the problems must share device, n_rhs and precision");
    }
    NCG_HIP(hipSetDevice(P[0]->device));
    int rc;
    for (ncg_problem *p : P)
        if ((rc = assemble_finish(p))) return rc;
    GroupStreams gs(P);
    if ((rc = gs.join())) return rc;
    ncg_tuning best;
    rc = autotune_group(P, std::max(1, reps), best);
    gs.leave();
    if (rc) return rc;
    if (out) *out = best;
    return 0;
}

int ncg_solve_group(ncg_problem *const *probs, int32_t count, const ncg_solver_options *o, ncg_solve_stats *stats)
{
    if (!probs || !o || count < 1) return fail(NCG_E_ARG, "ncg_solve_group: null argument or empty group");
    if (count > kGroupMax) return fail(NCG_E_ARG, "ncg_solve_group: at most 64 problems per group");
    if (o->max_iterations < 0) return fail(NCG_E_ARG, "ncg_solve_group: negative iteration limit");
    std::vector<ncg_problem *> P(probs, probs + count);
    for (int i = 0; i < count; ++i) {
        ncg_problem *p = P[i];
        if (!p) return fail(NCG_E_ARG, "ncg_solve_group: null problem");
        if (!p->assembled) return fail(NCG_E_STATE, "ncg_solve_group: assemble every problem first");
        if (p->device != P[0]->device) return fail(NCG_E_ARG, "ncg_solve_group: the problems must share a device");
        if (p->precision != P[0]->precision) return fail(NCG_E_ARG, "ncg_solve_group: the problems must share a precision");
        if (p->nrhs != P[0]->nrhs) return fail(NCG_E_ARG, "ncg_solve_group: the problems must have equal n_rhs");
        for (int j = 0; j < i; ++j)
            if (P[j] == p) return fail(NCG_E_ARG, "ncg_solve_group: a problem appears twice");
    }
    int rc = group_options(P[0]->precision, o);
    if (rc) return rc;
    NCG_HIP(hipSetDevice(P[0]->device));
    for (ncg_problem *p : P)
        if ((rc = assemble_finish(p))) return rc;
    if (stats) memset(stats, 0, sizeof(*stats) * (size_t)count);
    return solve_group(P, o, stats);
}

int ncg_get_solution(ncg_problem *p, double *x)
{
    if (!p || !x) return fail(NCG_E_ARG, "ncg_get_solution: null argument");
    if (!p->solved) return fail(NCG_E_STATE, "ncg_get_solution: solve first");
    NCG_HIP(hipSetDevice(p->device));
    NCG_HIP(hipMemcpy(x, p->x, sizeof(cd) * (size_t)p->nrhs * p->n, hipMemcpyDeviceToHost));
    return 0;
}

int ncg_field(ncg_problem *p, double *phi, double *grad)
{
    if (!p || !phi) return fail(NCG_E_ARG, "ncg_field: null argument");
    if (!p->solved) return fail(NCG_E_STATE, "ncg_field: solve first");
    if (p->m == 0) return 0;
    NCG_HIP(hipSetDevice(p->device));
    int rc = field_enqueue(p, grad != nullptr);
    return rc ? rc : field_collect(p, phi, grad);
}

int ncg_get_system(ncg_problem *p, double *A, double *b)
{
    if (!p) return fail(NCG_E_ARG, "ncg_get_system: null argument");
    if (!p->assembled) return fail(NCG_E_STATE, "ncg_get_system: assemble first");
    if (p->plan.method != 0) return fail(NCG_E_STATE, "ncg_get_system: the FMM operator is not stored densely");
    NCG_HIP(hipSetDevice(p->device));
    int rcf = assemble_finish(p);
    if (rcf) return rcf;
    const size_t nn = (size_t)p->n * p->n;
    if (A) {
        if (p->precision == NCG_MIXED) {
            std::vector<float> tmp(2 * nn);
            NCG_HIP(hipMemcpy(tmp.data(), p->A, sizeof(float) * 2 * nn, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < 2 * nn; ++i) A[i] = tmp[i];
        } else {
            NCG_HIP(hipMemcpy(A, p->A, sizeof(double) * 2 * nn, hipMemcpyDeviceToHost));
        }
    }
    if (b) NCG_HIP(hipMemcpy(b, p->rhs, sizeof(cd) * (size_t)p->nrhs * p->n, hipMemcpyDeviceToHost));
    return 0;
}

int ncg_apply(ncg_problem *p, const double *x, double *y, int32_t n_rhs)
{
    if (!p || !x || !y) return fail(NCG_E_ARG, "ncg_apply: null argument");
    if (!p->assembled) return fail(NCG_E_STATE, "ncg_apply: assemble first");
    if (n_rhs < 1 || n_rhs > p->nrhs) return fail(NCG_E_ARG, "ncg_apply: n_rhs must be in [1, assembled n_rhs]");
    NCG_HIP(hipSetDevice(p->device));
    int rc = assemble_finish(p);
    if (rc) return rc;
    rc = ensure_work(p, 8);
    if (rc) return rc;
    const size_t bytes = sizeof(cd) * (size_t)n_rhs * p->n;
    const int keep = p->nrhs;
    p->nrhs = n_rhs;
    NCG_HIP(hipMemcpyAsync(p->work[0], x, bytes, hipMemcpyHostToDevice, p->stream));
    mv(p, 1, p->work[0], nullptr, p->work[1]);
    p->nrhs = keep;
    NCG_HIP(hipMemcpyAsync(y, p->work[1], bytes, hipMemcpyDeviceToHost, p->stream));
    harvest_mv(p);
    NCG_HIP(hipStreamSynchronize(p->stream));
    return 0;
}

// ---- allocation-free apply ------------------------------------------------------------
}  // extern "C"

namespace {

// Scratch of ncg_apply_noalloc: one array per device, held until the process ends.  In host
// code its name is only a stand-in for the device address, which hipGetSymbolAddress
// gives once per device (kept in scratch_ptr); calls on one device share it.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {nullptr};

int scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(NCG_E_ARG, "ncg_apply_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(NCG_E_DEVICE, std::string("hipGetSymbolAddress (scratch): ") + hipGetErrorString(e));
        scratch_ptr[device] = ptr;
    }
    *out = scratch_ptr[device];
    return 0;
}

}  // namespace

extern "C" {

int ncg_apply_noalloc(ncg_problem *p, const double *x, double *y, int32_t n_rhs)
{
    if (!p || !x || !y) return fail(NCG_E_ARG, "ncg_apply_noalloc: null argument");
    if (!p->assembled) return fail(NCG_E_STATE, "ncg_apply_noalloc: assemble first");
    if (n_rhs < 1 || n_rhs > p->nrhs)
        return fail(NCG_E_ARG, "ncg_apply_noalloc: n_rhs must be in [1, assembled n_rhs]");
    NCG_HIP(hipSetDevice(p->device));
    int rc = assemble_finish(p);
    if (rc) return rc;
    const size_t need = mv_scratch_bytes(p, n_rhs);
    if (need > kScratchBytes)
        return fail(NCG_E_ARG, "ncg_apply_noalloc: this apply needs " + std::to_string(need >> 20) +
                                   " MiB of scratch, more than the " + std::to_string(kScratchBytes >> 20) + " MiB it has");
    void *base = nullptr;
    if ((rc = scratch_address(p->device, &base))) return rc;
    MvScratch na;
    na.base = (char *)base;
    na.bytes = kScratchBytes;
    const int keep = p->nrhs;
    p->nrhs = n_rhs;
    mv(p, 1, x, nullptr, y, &na);
    p->nrhs = keep;
    NCG_HIP(hipGetLastError());
    NCG_HIP(hipStreamSynchronize(p->stream));
    return 0;
}

int ncg_fmm_plan(const ncg_geometry *g, double k, ncg_fmm_info *info)
{
    if (!g || !info || g->n_elements <= 0 || !g->nodes_per_element || !g->node_xyz)
        return fail(NCG_E_ARG, "ncg_fmm_plan: missing geometry");
    if (!(k > 0.0)) return fail(NCG_E_ARG, "ncg_fmm_plan: wave number must be positive");
    ncgf::Mesh M;
    build_fmm_mesh(g, M);
    ncgf::Plan P;
    if (!ncgf::make_plan(M, fmm_options(g), k, P)) return fail(NCG_E_ARG, P.error);
    fill_info(P, M.n, info);
    return 0;
}

int ncg_problem_fmm_info(ncg_problem *p, ncg_fmm_info *info)
{
    if (!p || !info) return fail(NCG_E_ARG, "ncg_problem_fmm_info: null argument");
    fill_info(p->plan, p->n, info);
    return 0;
}

int ncg_timings(ncg_problem *p, double *t, int32_t reset)
{
    if (!p || !t) return fail(NCG_E_ARG, "ncg_timings: null argument");
    t[0] = p->t_asm;
    t[1] = p->t_solve;
    t[2] = p->t_field;
    t[3] = p->far_calls ? p->far_ms_sum / p->far_calls : 0.0;
    t[4] = p->mv_calls ? p->mv_ms_sum / p->mv_calls : 0.0;
    t[5] = (double)p->mv_calls;
    t[6] = p->last_evals;
    const double se = p->precision == NCG_MIXED ? 8.0 : 16.0;
    if (p->plan.method == 0) t[7] = (double)p->n * p->n * se;
    else {
        ncg_fmm_info info;
        fill_info(p->plan, p->n, &info);
        double nnz = (double)info.nnz_near;
        for (int l = 0; l < info.levels; ++l) nnz += (double)(info.nnz_t[l] + info.nnz_d[l] + info.nnz_s[l]);
        t[7] = nnz * se;
    }
    if (reset) {
        p->far_ms_sum = p->mv_ms_sum = 0.0;
        p->far_calls = p->mv_calls = 0;
    }
    return 0;
}

}  // extern "C"

// ==================================================================================== internal
namespace ncgi {

int fail(int code, const std::string &msg) { return ::fail(code, msg); }

int assemble_async(ncg_problem *p, const ncg_frequency *f, int precision)
{
    return assemble_impl(p, f, precision, false);
}
int assemble_wait(ncg_problem *p)
{
    if (hipSetDevice(p->device) != hipSuccess) return ::fail(NCG_E_DEVICE, "hipSetDevice failed");
    return assemble_finish(p);
}
int field_begin(ncg_problem *p, bool want_grad)
{
    if (!p->solved) return ::fail(NCG_E_STATE, "ncg_field: solve first");
    if (p->m == 0) return 0;
    if (hipSetDevice(p->device) != hipSuccess) return ::fail(NCG_E_DEVICE, "hipSetDevice failed");
    return field_enqueue(p, want_grad);
}
int field_end(ncg_problem *p, double *phi, double *grad)
{
    if (p->m == 0) return 0;
    return field_collect(p, phi, grad);
}

void planner_mesh(const ncg_geometry *g, ncgf::Mesh &M, ncgf::Options &O)
{
    build_fmm_mesh(g, M);
    O = fmm_options(g);
}

size_t estimate_bytes(const ncgf::Mesh &M, const ncgf::Options &O, int m_eval, double k, int precision, int nrhs,
                      int restart)
{
    ncgf::Plan P;
    if (!ncgf::make_plan(M, O, k, P)) return 0;
    const double n = M.n, nr = nrhs, m = m_eval, se = precision == NCG_MIXED ? 8.0 : 16.0;
    double b = n * 180.0 + m * 24.0;                                   // geometry
    b += 16.0 * nr * n * (4 + 10 + std::max(restart, 2) + 1);          // vectors, work, GMRES basis
    b += 64.0 * nr * m;                                                // field outputs
    if (P.method == 0) {
        b += n * n * se + n * 12.0 * 8.0;                              // dense operator, near pairs
    } else {
        const ncgf::Level &F = P.lev.back();
        double nnz = 0.0;
        for (const ncgf::Cluster &c : F.cl) {
            double row = 0.0;
            for (int j : c.near) row += (double)F.cl[j].elems.size();
            nnz += row * (double)c.elems.size();
        }
        b += nnz * (se + 8.0) + n * 16.0;                              // near CSR
        b += n * 64.0 * 16.0;                                          // block-Jacobi blocks (bound)
        for (const ncgf::Level &L : P.lev) {
            double fan = 0.0;
            for (const ncgf::Cluster &c : L.cl) fan += (double)c.fan.size();
            const double sf = (precision == NCG_MIXED && L.L > O.min_terms) ? 8.0 : 16.0;
            const double ncl = (double)L.cl.size(), pts = (double)L.P;
            b += 1.5 * ((2.0 * n + fan) * pts * sf + 3.0 * 16.0 * nr * ncl * pts);   // T, S, D, work (grow slack)
            b += (4.0 * n + 3.0 * fan) * 4.0 + 32.0 * pts;
        }
        const ncgf::Level &L0 = P.lev.front();
        const double npair = (double)P.ev.size() * (double)L0.cl.size();
        b += 1.5 * (npair * (4.0 + 2.0 * L0.L) * 8.0 + 16.0 * nr * (double)P.ev.size() * L0.P) + 64.0 * nr * m;
    }
    return (size_t)(b * 1.05) + ((size_t)64 << 20);
}

}  // namespace ncgi

// ==================================================================================== kernel timing
extern "C" {

int ncg_set_kernel_timing(ncg_problem *p, int32_t on)
{
    if (!p) return fail(NCG_E_ARG, "ncg_set_kernel_timing: null argument");
    p->ktime = on ? 1 : 0;
    return 0;
}

int ncg_kernel_stats(ncg_problem *p, ncg_kernel_stat *out, int32_t cap, int32_t *count, int32_t reset)
{
    if (!p || !count || (cap > 0 && !out)) return fail(NCG_E_ARG, "ncg_kernel_stats: null argument");
    NCG_HIP(hipSetDevice(p->device));
    harvest_mv(p);
    int c = 0;
    for (int id = 0; id < kNumKernels; ++id) {
        if (p->k_calls[id] == 0) continue;
        if (c < cap) {
            memset(&out[c], 0, sizeof(out[c]));
            snprintf(out[c].name, sizeof(out[c].name), "%s", kKernelNames[id]);
            out[c].launches = p->k_calls[id];
            out[c].total_ms = p->k_ms[id];
            out[c].bytes = p->k_bytes[id];
            out[c].flops = p->k_flops[id];
        }
        ++c;
    }
    *count = c;
    if (reset)
        for (int id = 0; id < kNumKernels; ++id) {
            p->k_calls[id] = 0;
            p->k_ms[id] = p->k_bytes[id] = p->k_flops[id] = 0.0;
        }
    return 0;
}

}  // extern "C"
