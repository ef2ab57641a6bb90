/*
 * ncg.h — C ABI of libncg, the MI355X (gfx950) implementation of NumCalc's
 * Burton–Miller collocation BEM solve.
 *
 * The reference exposes no library API for this path: its boundary is the NumCalc
 * executable (mesh2hrtf/NumCalc/src/NC_Main.cpp:137-281) driven through files.  These
 * entry points are the seams inside that executable, one per phase of a frequency
 * step of NC_ControlProgram (src/NC_Main.cpp:370-957, src/ = mesh2hrtf/NumCalc/src/).
 * The drop-in `NumCalc` host program built beside this library calls them; a caller
 * that wants the solver without files (ctypes, a serving loop) calls them directly.
 *
 * Conventions: plain C types only; complex numbers are interleaved (re, im) doubles;
 * every function returns 0 on success or a negative NCG_E* code, with a message in
 * ncg_last_error().  No function calls exit().  A problem handle owns its device
 * memory and one HIP stream; calls on one handle must not overlap, handles on
 * different threads are independent.
 */
#ifndef NCG_H
#define NCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    NCG_OK = 0,
    NCG_E_ARG = -1,        /* invalid argument                                    */
    NCG_E_DEVICE = -2,     /* HIP runtime error or no device                      */
    NCG_E_MEMORY = -3,     /* device allocation failed                            */
    NCG_E_SUBDIV = -4,     /* adaptive subdivision limits exceeded (see message)  */
    NCG_E_STATE = -5,      /* call order violated (e.g. solve before assemble)    */
    NCG_E_NOCONV = -6      /* iterative solver reached its iteration limit        */
};

/* element boundary-condition kinds (ibval of src/NC_Input.cpp:1163-1406) */
enum { NCG_BC_VELOCITY = 0, NCG_BC_PRESSURE = 1, NCG_BC_ADMITTANCE = 2 };

/* arithmetic of the bulk quadrature and of the stored operator */
enum {
    NCG_FP64 = 0,          /* plain method: everything FP64                                   */
    NCG_MIXED = 1          /* FP32-complex bulk quadrature, operator and Krylov matvec;
                              FP64 singular/near integrals, Krylov dots and refinement residual;
                              FMM levels and the evaluation grid's translations stay FP64
                              where their series would cancel (k d < 1 or a large peak term) */
};

enum { NCG_SOLVER_CGS = 0, NCG_SOLVER_GMRES = 1, NCG_SOLVER_LU = 4 };
/* IPRECOND of NC.inp (src/NC_Input.cpp:298-323): 0 (and -1..-3, which choose the
 * scanning degree) incomplete LU, 1 row scaling, 2 none.  NCG_PRECOND_BLOCK has no
 * counterpart in NumCalc: block-Jacobi with the inverses of the finest clusters'
 * near-field self blocks (FMM plans only), for the served GMRES / mixed solves. */
enum { NCG_PRECOND_ILU = 0, NCG_PRECOND_SCALING = 1, NCG_PRECOND_NONE = 2, NCG_PRECOND_BLOCK = 3 };

/* Frequency-independent description of one boundary-element problem.
 * Replaces the geometry arrays NumCalc builds in NC_ReadMesh
 * (src/NC_Input.cpp:353-737): element node coordinates, the boundary-condition
 * kind per element (it decides the unknown: potential or normal velocity), and
 * the evaluation-grid nodes of NC_WriteResultsEvaluationGrid (src/NC_PostProcessing.cpp:581). */
/* Method of NC.inp "Main Parameters I" and the FMM inputs of "Main Parameters II"
 * (src/NC_Input.cpp:246-323); the per-frequency choice between dense BEM, SL-FMM and
 * ML-FMM then follows NC_DeclareArrays (src/NC_Addresses.cpp:58-287). */
typedef struct {
    int32_t method_bem;                 /* 0 dense (TBEM), 1 SL-FMM, 4 ML-FMM               */
    int32_t min_expansion_terms;        /* 5..10, 0 -> 8                                    */
    double cluster_edge;                /* ClusEdgL0; <= 0: computed as NumCalc does        */
} ncg_method;

typedef struct {
    int32_t n_elements;                 /* boundary elements = unknowns                    */
    const int32_t *nodes_per_element;   /* [n] 3 (triangle) or 4 (quadrilateral)          */
    const double *node_xyz;             /* [n][4][3]; 4th node ignored for triangles       */
    const int32_t *bc_kind;             /* [n] NCG_BC_*                                    */
    int32_t n_eval;                     /* evaluation points (may be 0)                    */
    const double *eval_xyz;             /* [n_eval][3]                                     */
    double harmonic_time_factor;        /* +1: e^{-i w t} (NumCalc default), -1: e^{+i w t} */
    int32_t interior;                   /* 0 exterior problem, 1 interior                  */
    const int32_t *node_index;          /* [n][4] mesh node ids (NULL: nodes with equal
                                           coordinates are one node); FMM cluster centres  */
    const int32_t *element_group;       /* [n] element group ids (NULL: one group)         */
    ncg_method method;
} ncg_geometry;

/* FMM plan summary: the cluster and expansion figures NumCalc writes to NC.out
 * (NC_FrequencyInformations src/NC_Main.cpp:1051-1153, src/NC_Main.cpp:491-522,
 * NC_AllocateSDTmtxs* src/NC_Addresses.cpp:1476-1654). */
typedef struct {
    int32_t method;                     /* 0 dense, 1 SL-FMM, 3 ML-FMM (methodFMM_)         */
    int32_t levels;
    int32_t clusters[10];
    double radius_max[10], radius_min[10];
    int32_t expansion_terms[10], points_theta[10], points_phi[10], points_sphere[10];
    int64_t nnz_near, nnz_t[10], nnz_d[10], nnz_s[10];
    int64_t canceled_pairs[10], far_pairs[10];
    int32_t eval_clusters;
    double eval_radius_max, eval_radius_min;
    char warning[256];
} ncg_fmm_info;

/* One frequency step with n_rhs load cases that share the operator.  Values are
 * the per-frequency ones of NC_UpdateFreqCurves (src/NC_EquationSystem.cpp:1815-1961):
 * prescribed pressures and source strengths already divided by i*rho*omega*h,
 * admittances already multiplied by i*h*k. */
typedef struct {
    double k;                           /* wave number omega/c                              */
    int32_t n_rhs;                      /* load cases (1 for a NumCalc step)                */
    const double *bc_value;             /* [n_rhs][n] complex: velocity or pressure per element */
    const double *admittance;           /* [n] complex, used where bc_kind == ADMITTANCE    */
    int32_t n_point_sources;
    const double *point_sources;        /* [ns][6]: x y z re im rhs_index                   */
    int32_t n_plane_waves;
    const double *plane_waves;          /* [nw][6]: unit direction, re im rhs_index         */
} ncg_frequency;

typedef struct {
    int32_t solver;                     /* NCG_SOLVER_*                                      */
    int32_t precond;                    /* NCG_PRECOND_*                                     */
    int32_t max_iterations;             /* NumCalc default 250 (src/NC_Main.cpp:83)          */
    double tolerance;                   /* NumCalc 1e-9 (src/NC_CommonFunctions.cpp:39)      */
    int32_t restart;                    /* GMRES restart length                              */
    int32_t refinement_steps;           /* NCG_MIXED: max outer FP64 refinement steps        */
    double refinement_tolerance;        /* NCG_MIXED: FP64 residual target                   */
    int32_t ilu_scanning_degree;        /* NCG_PRECOND_ILU: 0..3 (-IPRECOND), sets the drop
                                           threshold (src/NC_CommonFunctions.cpp:1367-1403) */
} ncg_solver_options;

typedef struct {
    int32_t iterations[64];             /* per load case (first 64)                          */
    double relative_residual[64];
    int32_t converged_all;
    int32_t refinement_steps;
    double ms_assemble, ms_solve, ms_field;   /* device time of the last calls (HIP events)  */
    int64_t ilu_nonzeros;               /* NCG_PRECOND_ILU: entries kept in L + U            */
    int32_t ilu_levels[2];              /* dependency depth of the L and U solves            */
} ncg_solve_stats;

typedef struct ncg_problem ncg_problem;

/* Library / device info.  ncg_device_count replaces nothing in the reference. */
const char *ncg_last_error(void);
int ncg_device_count(int32_t *count);

/* Method and FMM plan NumCalc would use at wave number k for this geometry, without a
 * device (host computation only).  Replaces the tree part of NC_DeclareArrays
 * (src/NC_Addresses.cpp:58-287, 290-1314, 1476-1937) and the expansion lengths of
 * src/NC_Main.cpp:421-489.  Returns NCG_E_ARG with NumCalc's message on its fatal cases. */
int ncg_fmm_plan(const ncg_geometry *geometry, double k, ncg_fmm_info *info);

/* Upload geometry, compute element data and the frequency-independent near-pair
 * table (the "which pairs need subdivision" plan) on the device.
 * Replaces the geometry part of NC_ReadMesh (src/NC_Input.cpp:522-605). */
int ncg_problem_create(const ncg_geometry *geometry, int32_t device, ncg_problem **out);
int ncg_problem_destroy(ncg_problem *p);

/* Dense Burton–Miller system and right-hand sides of one frequency on the device.
 * Replaces NC_UpdateFreqCurves' use sites + NC_SetupEquationSystem / NC_BuildSystemTBEM
 * (src/NC_EquationSystem.cpp:79-302, 1408-1812). precision: NCG_FP64 or NCG_MIXED. */
int ncg_assemble(ncg_problem *p, const ncg_frequency *freq, int32_t precision);

/* Solve the assembled system for all load cases.  An NCG_MIXED operator is solved by
 * FP64 iterative refinement around FP32 GMRES: it needs solver NCG_SOLVER_GMRES and
 * precond NCG_PRECOND_BLOCK or NCG_PRECOND_NONE (anything else: NCG_E_ARG).
 * Replaces NC_IterativeSolverCGS / Tfactor_usy+Tfbelim (src/NC_CommonFunctions.cpp:1066-1208,
 * src/NC_TypeDefinition.h:542-594, selected at src/NC_Main.cpp:833-866). */
int ncg_solve(ncg_problem *p, const ncg_solver_options *opt, ncg_solve_stats *stats);

/* Solve several assembled problems together (frequency steps of one mesh or of different
 * meshes, equal n_rhs and precision, one device) with GMRES or the mixed form: every Krylov
 * step is one launch per operation for the whole group, so the operator applies of the
 * group share each launch, and the host waits once per step for all of them.  Each
 * problem's result is the one ncg_solve gives it alone.  stats: [count] or NULL.  The
 * problems' streams are joined for the call; calls on them must not overlap it.
 * Replaces the per-frequency solves of NC_IterativeSolverCGS (src/NC_CommonFunctions.cpp:
 * 1066-1208) as run by manage_numcalc's one process per frequency
 * (mesh2hrtf/NumCalc/manage_numcalc.py:376-430): here the frequencies share launches. */
int ncg_solve_group(ncg_problem *const *problems, int32_t count, const ncg_solver_options *opt,
                    ncg_solve_stats *stats);

/* Launch shapes the operator applies and the dense assembly run with, chosen by measurement
 * on the problems at hand (ncg_autotune) rather than fixed.  The FMM fields change how the
 * grouped applies split their sums (hence the last bits of a result); the near-field mode,
 * the near-field assembly mapping and the assembly rows do not change any value.  A group
 * solve uses its first problem's
 * tuning.  No counterpart in the reference. */
typedef struct {
    int32_t t_waves, d_waves, s_waves;  /* waves per work-group of the grouped T, D, S applies:
                                           1, 2, 4 or 8 (default 8)                          */
    int32_t tuned;                      /* 1 once chosen by ncg_autotune                      */
    int32_t near_mode;                  /* near-field work-group: 0 one row chunk (default),
                                           1 one finest cluster (x gathered once), 2 and 3
                                           persistent work-groups over the row chunks of
                                           every member: 2 from work queues (one per XCD),
                                           3 by a static stride                              */
    int32_t mfma_splits;                /* dense, more than 4 load cases: column splits of the
                                           matrix-core product, 1..16 (0: from the size)     */
    int32_t assemble_rows;              /* dense: collocation rows per work-group of the
                                           far-field quadrature, 8..256 (0: 32)             */
    int32_t near_assembly;              /* FMM near-field quadrature: 0 a thread per column
                                           (geometry in registers, default), 1 a wave per
                                           collocation row (column geometry in LDS)         */
} ncg_tuning;
int ncg_set_tuning(ncg_problem *p, const ncg_tuning *t);
int ncg_get_tuning(ncg_problem *p, ncg_tuning *t);
/* Times the candidate launch shapes on assembled problems (a group as ncg_solve_group takes
 * it: equal n_rhs and precision, one device): near-field mode and T, D, S waves of the
 * grouped FMM applies, matrix-core splits of dense applies with more than 4 load cases, rows
 * per work-group of the dense far-field quadrature (re-run on the assembled frequency; the
 * operator's values do not change).  reps applies per candidate.  Sets the fastest choice on
 * every problem and returns it in *out (or NULL). */
int ncg_autotune(ncg_problem *const *problems, int32_t count, int32_t reps, ncg_tuning *out);

/* Arnoldi steps of the GMRES solves (ncg_solve, ncg_solve_group) that p leads: replayed as
 * HIP graphs (on, the default) or launched kernel by kernel (off).  A step's graph is built
 * once from its launches (kernel nodes in launch order; no stream capture, so other
 * threads' allocations and waits are unaffected).  Every launch of a step reads the step
 * index from device memory, so one graph serves every step of a solve; graphs are cached
 * per leader by a digest of all the launch arguments they hold and reused by later solves
 * whose launches are the same.  Not used while kernel timing
 * is on (its events time single launches) or when a member of the group is an FMM
 * operator with more than 4 load cases.  Results are the same either way, bit for bit.
 * Replaces the per-iteration host loop of NC_IterativeSolverCGS
 * (src/NC_CommonFunctions.cpp:1126-1176); no counterpart for the graphs themselves. */
int ncg_set_graphs(ncg_problem *p, int32_t on);
/* Arnoldi steps of the solves p led since it was created: out[0] graphs captured,
 * out[1] steps replayed from a graph, out[2] steps launched kernel by kernel, out[3]
 * graphs held in p's cache.  No counterpart in the reference. */
int ncg_graph_stats(ncg_problem *p, int64_t *out);

/* The method and FMM plan of the last ncg_assemble (what NumCalc logs per step). */
int ncg_problem_fmm_info(ncg_problem *p, ncg_fmm_info *info);

/* Solved unknowns to host: [n_rhs][n] complex (potential for velocity/admittance
 * elements, normal velocity for pressure elements).  Replaces reading zrhs after the
 * solve (src/NC_PostProcessing.cpp:367-578). */
int ncg_get_solution(ncg_problem *p, double *x);

/* Potential and its gradient at the evaluation points, incident field included:
 * phi [n_rhs][n_eval] complex, grad [n_rhs][n_eval][3] complex (grad may be NULL).
 * Replaces the TBEM branch of NC_WriteResultsEvaluationGrid
 * (src/NC_PostProcessing.cpp:581-866, 869-975, 1217-1322). */
int ncg_field(ncg_problem *p, double *phi, double *grad);

/* ---- inspection hooks used by the parity tests and benchmarks -------------------- */

/* Copy the assembled operator (row-major, complex128 even in NCG_MIXED) and the
 * right-hand sides [n_rhs][n] to host. */
int ncg_get_system(ncg_problem *p, double *A, double *b);

/* y = A x for n_rhs vectors [n_rhs][n] (host buffers), with the stored operator. */
int ncg_apply(ncg_problem *p, const double *x, double *y, int32_t n_rhs);

/* y = A x like ncg_apply, but on device buffers (x, y: [n_rhs][n] complex in device memory
 * of the problem's device) and without allocating anything: the apply's scratch (FMM
 * signatures, matrix-core operand and partial sums) comes from one 64 MiB array per device
 * that the library holds until the process ends; an apply that needs more fails with
 * NCG_E_ARG.  Calls on one device share that array: calls that overlap on one device race.
 * Returns after the apply has completed on the problem's stream.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid allocation). */
int ncg_apply_noalloc(ncg_problem *p, const double *x, double *y, int32_t n_rhs);

/* Per-phase device timings of the last assemble/solve/field calls, and the average
 * duration of the dominant kernel of each phase over the calls since the last reset
 * (HIP events on the problem's stream).  Layout of t[8]:
 * [0] assemble ms  [1] solve ms  [2] field ms  [3] far-field quadrature kernel avg ms
 * [4] matvec kernel avg ms  [5] matvec launches  [6] Green's-function evaluations of
 * the last assembly's quadrature (regular, subdivided and singular; counted once per plan
 * while kernel timing is on, else 0)
 * [7] bytes of the stored operator. */
int ncg_timings(ncg_problem *p, double *t, int32_t reset);

/* Per-kernel timing of the operator apply (near-field, T, D, S or dense matvec
 * launches, single-problem and grouped) and of the assembly and field kernels (regular,
 * subdivided and singular quadrature, FMM T / S / D tables, evaluation-grid field): HIP
 * events around every launch on the problem's stream while enabled, the algorithmic bytes
 * each apply launch must move (stored operator once per 4 load cases, vectors in and out)
 * and the algorithmic FLOPs of each assembly launch: 60 per Green's-function evaluation
 * (G and its three normal derivatives at one quadrature point; the evaluations are counted
 * on the device once per plan, in the first assembly and field pass with timing on), 16 per
 * plane-wave evaluation of T (element Gauss point x sphere point), 20 per S entry, 7 per
 * term of the D series.  No counterpart in the reference; for measurement.  Stats are
 * accumulated since the last reset; count receives the number of kernels with launches. */
typedef struct {
    char name[48];
    int64_t launches;
    double total_ms;
    double bytes;
    double flops;
} ncg_kernel_stat;
int ncg_set_kernel_timing(ncg_problem *p, int32_t on);
int ncg_kernel_stats(ncg_problem *p, ncg_kernel_stat *out, int32_t cap, int32_t *count, int32_t reset);

/* ---------------------------------------------------------------------------------
 * Serving runtime.  No counterpart in the reference, whose parallelism is one NumCalc
 * process per (source, frequency range) started by manage_numcalc
 * (mesh2hrtf/NumCalc/manage_numcalc.py:376-430, RAM admission from Memory.txt at
 * :236-290).  Here one process serves jobs (mesh, band of frequencies, load cases)
 * on one GPU with continuous batching: `lanes` worker threads each drive their own
 * problem handles and stream, take the next frequency step of any queued job as soon
 * as they are free and the step's device memory fits the budget, and keep geometry,
 * FMM trees and operators resident per mesh across frequencies.  Steps of different
 * jobs overlap on the device.  Results do not depend on what else is running.
 * --------------------------------------------------------------------------------- */
typedef struct ncg_server ncg_server;

typedef struct {
    int32_t device;                     /* HIP device                                        */
    int32_t lanes;                      /* steps in flight (threads + streams); 0 -> 4       */
    double memory_budget_bytes;         /* device memory resident problems may hold;
                                           0 -> 90% of the free memory at creation          */
    int32_t problems_per_lane;          /* meshes a lane keeps resident (LRU); 0 -> 2         */
    int32_t kernel_timing;              /* 1: per-kernel timing of every step (ncg_kernel_stats) */
    int32_t group;                      /* GMRES / mixed steps a lane solves together
                                           (ncg_solve_group); 0 -> 4, 1: one at a time       */
    int32_t autotune;                   /* 1: the first group of each operator shape (mesh
                                           size, method, clusters per level, n_rhs, precision)
                                           is timed with ncg_autotune and its choice used for
                                           every later group of that shape                  */
    const char *tuning_file;            /* text file of tuned shapes read at creation and
                                           written at ncg_server_destroy, or NULL            */
    int32_t eager;                      /* 1: Arnoldi steps launched kernel by kernel instead
                                           of replayed as captured graphs (ncg_set_graphs)  */
} ncg_server_config;

typedef struct {
    int32_t mesh;                       /* id from ncg_server_add_mesh                        */
    int32_t n_freq;
    const ncg_frequency *freq;          /* [n_freq], equal n_rhs; copied at submit           */
    int32_t precision;                  /* NCG_FP64 or NCG_MIXED                              */
    ncg_solver_options solver;
    double *solution;                   /* [n_freq][n_rhs][n] complex out, or NULL            */
    double *phi;                        /* [n_freq][n_rhs][m] complex out, or NULL            */
    double *grad;                       /* [n_freq][n_rhs][m][3] complex out, or NULL         */
} ncg_job;                              /* output buffers must stay valid until the job ends */

typedef struct {
    int32_t state;                      /* 0 queued, 1 running, 2 done, 3 failed             */
    int32_t error_code;
    int32_t freq_done;
    int32_t iterations;                 /* Krylov iterations summed over steps and load cases */
    double latency_ms;                  /* submit -> last step done                          */
    double queue_ms;                    /* submit -> first step started                      */
    int32_t refinement_steps_max;       /* mixed form: most refinement steps of a step       */
    double relative_residual_max;       /* worst final relative residual of a step and load
                                           case (the FP64 refinement residual, mixed form)   */
    char error[256];
} ncg_job_status;

typedef struct {
    int64_t jobs_done, jobs_failed, steps_done, solves_done;   /* solves = steps x load cases */
    double p50_ms, p99_ms, mean_ms;     /* job latency over finished jobs                    */
    double wall_s;                      /* since creation                                    */
    int64_t resident_bytes, peak_resident_bytes;
    int32_t lanes, queued_steps;
    double lane_ms[4];                  /* lane wall time summed over the lanes: assembly,
                                           solve, outputs (solution and grid), waiting       */
    int32_t tuned_shapes;               /* operator shapes with a tuning (ncg_server_tuning)  */
    int64_t graph_steps, eager_steps;   /* GMRES Arnoldi steps replayed from captured graphs,
                                           and launched kernel by kernel (ncg_set_graphs)    */
} ncg_server_info;

int ncg_server_create(const ncg_server_config *cfg, ncg_server **out);
/* Registers a mesh (the geometry is copied); problems are created on the lanes lazily. */
int ncg_server_add_mesh(ncg_server *s, const ncg_geometry *geometry, int32_t *mesh_id);
int ncg_server_submit(ncg_server *s, const ncg_job *job, int64_t *job_id);
/* Waits until the job is done or failed, at most timeout_s seconds (< 0: no limit);
 * returns NCG_E_STATE if it is still running at the limit. */
int ncg_server_wait(ncg_server *s, int64_t job_id, double timeout_s, ncg_job_status *status);
int ncg_server_query(ncg_server *s, int64_t job_id, ncg_job_status *status);
int ncg_server_stats(ncg_server *s, ncg_server_info *stats);
/* Per-kernel timing summed over the steps finished since the last reset (requires
 * kernel_timing in the configuration). */
int ncg_server_kernel_stats(ncg_server *s, ncg_kernel_stat *out, int32_t cap, int32_t *count, int32_t reset);
/* Turns per-kernel timing of the steps started from now on on or off (the configuration's
 * kernel_timing is the initial state). */
int ncg_server_set_kernel_timing(ncg_server *s, int32_t on);
/* The tuning of operator shape `index` (0 .. tuned_shapes - 1) and its key text. */
int ncg_server_tuning(ncg_server *s, int32_t index, char *key, int32_t key_cap, ncg_tuning *tuning);
/* Finishes the queued work, stops the lanes, frees every problem. */
int ncg_server_destroy(ncg_server *s);

#ifdef __cplusplus
}
#endif
#endif /* NCG_H */
