// Serving runtime of libncg (ncg_server_* in include/ncg.h).
//
// Jobs are (mesh, band of frequencies, load cases).  Each frequency of a job is one
// step: assemble, solve all load cases, evaluate the grid.  Steps wait in one queue in
// arrival order; `lanes` worker threads take the first step whose device memory fits
// the budget as soon as they are free, so steps of different jobs run side by side on
// the device, each on the stream of its problem handle.  Problem handles live in a
// pool shared by the lanes: a step reuses an idle problem of its mesh (geometry,
// cluster tree, structure and grow-only operator buffers stay resident across
// frequencies) or creates one, evicting idle problems of other meshes, least recently
// used first, when the budget requires it.  A step's footprint is estimated from the
// plan of the job's highest frequency (ncgi::estimate_bytes).
//
// The reference's counterpart is a pool of NumCalc processes started by
// manage_numcalc (mesh2hrtf/NumCalc/manage_numcalc.py:376-430) and admitted by their
// Memory.txt RAM estimates (:236-290).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <iterator>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ncg.h"
#include "ncg_internal.hpp"

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// a registered mesh: owned copies of the geometry arrays and the planner mesh used
// for footprint estimates
struct MeshRec {
    ncg_geometry g;
    std::vector<int32_t> nv, kind, node, group;
    std::vector<double> xyz, eval;
    ncgf::Mesh pm;
    ncgf::Options po;
    std::mutex pm_mu;   // pm caches cluster trees: one planner at a time
};

// owned copy of one ncg_frequency
struct FreqCopy {
    ncg_frequency f;
    std::vector<double> bc, adm, ps, pw;
};

struct Job {
    int64_t id = 0;
    int mesh = 0, precision = NCG_MIXED, nrhs = 1, n = 0, m = 0;
    int method_bem = 0;   // the mesh's NumCalc method: 0 dense, 1 / 4 FMM (chosen per frequency)
    ncg_solver_options solver;
    std::vector<FreqCopy> freq;
    double *solution = nullptr, *phi = nullptr, *grad = nullptr;
    size_t bytes = 0;   // device footprint of one step (highest frequency)
    clk::time_point t_submit, t_first, t_done;
    bool started = false;
    bool failed = false;   // a step failed: no further step starts, outputs are no longer written
    int inflight = 0;      // steps of this job running on the lanes
    int freq_done = 0, state = 0, error_code = 0;
    long long iterations = 0;
    int refine_max = 0;
    double relres_max = 0.0;
    std::string error;
};

struct Step {
    Job *job;
    int fi;
};

// a problem handle in the shared pool
struct Res {
    int mesh = -1;
    ncg_problem *p = nullptr;
    size_t bytes = 0;
    bool busy = false;
    unsigned long long last = 0;
};

}  // namespace

struct ncg_server {
    ncg_server_config cfg{};
    double budget = 0.0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<std::unique_ptr<MeshRec>> meshes;
    std::map<int64_t, std::unique_ptr<Job>> jobs;
    std::deque<Step> queue;
    std::vector<std::unique_ptr<Res>> pool;
    std::vector<std::thread> lanes;
    int64_t next_id = 1;
    unsigned long long tick = 0;
    bool stop = false;
    int running = 0;
    size_t resident = 0, peak = 0;
    std::vector<double> latencies;
    int64_t jobs_done = 0, jobs_failed = 0, steps_done = 0, solves_done = 0;
    clk::time_point t0;
    std::map<std::string, ncg_kernel_stat> kstats;   // per-kernel timing of finished steps
    double host_ms[4] = {0, 0, 0, 0};                 // lane wall time: assemble, solve, outputs, waiting
    // launch-shape tunings per operator shape (ncg_autotune), and the shapes being timed
    std::map<std::string, ncg_tuning> tunings;
    std::map<std::string, int> tuning_busy;
    std::condition_variable cv_tune;
    std::string tuning_file;
    std::atomic<int> ktime{0};   // per-kernel timing of the steps started from now on
    int64_t graph_steps = 0, eager_steps = 0;   // Arnoldi steps of the lanes' solves (replayed / one by one)
};

namespace {

int sfail(int code, const std::string &msg) { return ncgi::fail(code, msg); }

// Ends a job (lock held).  A failed step marks the job failed and drops its queued steps
// at once, but the job is reported ended (state 3, waiters woken) only when none of its
// steps is still running on another lane, so no lane writes into the caller's output
// buffers after ncg_server_wait has returned.
void end_job(ncg_server *S, Job *J)
{
    if (J->state >= 2) return;
    J->t_done = clk::now();
    if (J->failed) {
        J->state = 3;
        S->jobs_failed++;
    } else {
        J->state = 2;
        S->jobs_done++;
    }
    S->latencies.push_back(ms_between(J->t_submit, J->t_done));
    S->cv_done.notify_all();
}

void fail_job(ncg_server *S, Job *J, int code, const std::string &msg)
{
    if (J->state >= 2) return;
    if (!J->failed) {
        J->failed = true;
        J->error_code = code;
        J->error = msg;
        S->queue.erase(std::remove_if(S->queue.begin(), S->queue.end(), [&](const Step &s) { return s.job == J; }),
                       S->queue.end());
    }
    if (J->inflight == 0) end_job(S, J);
}

// Operator shape of an assembled problem, the key of its launch-shape tuning: unknowns,
// load cases, precision, method and clusters per level (they fix the launches' sizes).
std::string shape_key(ncg_problem *p, int n, int nrhs, int precision)
{
    ncg_fmm_info fi;
    if (ncg_problem_fmm_info(p, &fi) != 0) return std::string();
    std::string k = "n " + std::to_string(n) + " rhs " + std::to_string(nrhs) + " precision " +
                    std::to_string(precision) + " method " + std::to_string(fi.method) + " clusters";
    for (int l = 0; l < fi.levels && l < 10; ++l) k += " " + std::to_string(fi.clusters[l]);
    return k;
}

void load_tunings(ncg_server *S)
{
    FILE *f = fopen(S->tuning_file.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char *bar = strchr(line, '|');
        if (!bar) continue;
        *bar = 0;
        ncg_tuning t{};
        if (sscanf(bar + 1, "%d %d %d %d %d %d %d", &t.t_waves, &t.d_waves, &t.s_waves, &t.near_mode, &t.mfma_splits,
                   &t.assemble_rows, &t.near_assembly) >= 6) {
            t.tuned = 1;
            S->tunings[line] = t;
        }
    }
    fclose(f);
}

void save_tunings(ncg_server *S)
{
    FILE *f = fopen(S->tuning_file.c_str(), "w");
    if (!f) return;
    for (auto &kv : S->tunings)
        fprintf(f, "%s|%d %d %d %d %d %d %d\n", kv.first.c_str(), kv.second.t_waves, kv.second.d_waves,
                kv.second.s_waves, kv.second.near_mode, kv.second.mfma_splits, kv.second.assemble_rows,
                kv.second.near_assembly);
    fclose(f);
}

// The group's tuning (lock not held): the stored one of its shape, or (autotune) timed on
// this group while other lanes with the same shape wait for it.
void tune_group(ncg_server *S, const std::vector<ncg_problem *> &P, int n, int nrhs, int precision)
{
    const std::string key = shape_key(P[0], n, nrhs, precision);
    if (key.empty()) return;
    ncg_tuning t{};
    bool have = false, mine = false;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        S->cv_tune.wait(lk, [&] { return !S->tuning_busy.count(key); });
        auto it = S->tunings.find(key);
        if (it != S->tunings.end()) {
            t = it->second;
            have = true;
        } else if (S->cfg.autotune) {
            S->tuning_busy[key] = 1;
            mine = true;
        }
    }
    if (mine) {
        have = ncg_autotune(P.data(), (int32_t)P.size(), 3, &t) == 0;
        std::lock_guard<std::mutex> lk(S->mu);
        if (have) S->tunings[key] = t;
        S->tuning_busy.erase(key);
        S->cv_tune.notify_all();
    }
    if (have)
        for (ncg_problem *p : P) ncg_set_tuning(p, &t);
}

// Steps a lane may solve together (ncg_solve_group): GMRES or mixed-precision steps of equal
// precision, load-case count and solver options.
// Steps are grouped only where the group shares launches: FMM operators with at most 4
// load cases per pass.  A dense operator (methodBEM 0) or more load cases is applied by
// single-problem launches, one member after the other, so a group would only hold its
// members to the slowest one's iterations: on 256 dense meshes with 16 load cases, groups
// of 4 ran at 2980 frequency-solves/s and single steps at 4580.
// a GMRES step with the block-Jacobi or no preconditioner (ncg_solve_group's solves; their
// launch shapes are tuned per operator shape)
bool gmres_job(const Job *J)
{
    return J->solver.solver == NCG_SOLVER_GMRES &&
           (J->solver.precond == NCG_PRECOND_BLOCK || J->solver.precond == NCG_PRECOND_NONE);
}
bool groupable(const Job *J) { return gmres_job(J) && J->nrhs <= 4 && J->method_bem != 0; }
bool same_group(const Job *a, const Job *b)
{
    return groupable(a) && groupable(b) && a->precision == b->precision && a->nrhs == b->nrhs &&
           a->solver.precond == b->solver.precond && a->solver.max_iterations == b->solver.max_iterations &&
           a->solver.tolerance == b->solver.tolerance && a->solver.restart == b->solver.restart &&
           a->solver.refinement_steps == b->solver.refinement_steps &&
           a->solver.refinement_tolerance == b->solver.refinement_tolerance;
}

// Choice of the next step for a free lane (lock held).  Returns the queue index, the
// pool entry to reuse (or null: create) and the idle entries to evict, or -1.  With `with`
// set, only steps that may join that job's group are considered.
int choose(ncg_server *S, Res **reuse, std::vector<Res *> &evict, size_t &need, const Job *with = nullptr)
{
    const size_t cap = (size_t)S->cfg.lanes * (size_t)std::max(S->cfg.problems_per_lane, S->cfg.group);
    const size_t scan = std::min<size_t>(S->queue.size(), 64);
    for (size_t q = 0; q < scan; ++q) {
        const Step &st = S->queue[q];
        if (with && !same_group(with, st.job)) continue;
        const int mesh = st.job->mesh;
        Res *r = nullptr;
        for (auto &e : S->pool)
            if (!e->busy && e->mesh == mesh && (!r || e->last > r->last)) r = e.get();
        need = r ? (st.job->bytes > r->bytes ? st.job->bytes - r->bytes : 0) : st.job->bytes;
        // idle entries of other meshes, least recently used first
        std::vector<Res *> idle;
        for (auto &e : S->pool)
            if (!e->busy && e.get() != r) idle.push_back(e.get());
        std::sort(idle.begin(), idle.end(), [](const Res *a, const Res *b) { return a->last < b->last; });
        evict.clear();
        size_t res = S->resident;
        size_t count = S->pool.size();
        size_t k = 0;
        while ((double)(res + need) > S->budget || (!r && count >= cap)) {
            if (k == idle.size()) break;
            res -= idle[k]->bytes;
            evict.push_back(idle[k]);
            --count;
            ++k;
        }
        if ((double)(res + need) <= S->budget && (r || count < cap)) {
            *reuse = r;
            return (int)q;
        }
    }
    return -1;
}

// Takes step q of the queue for a lane (lock held): reserves its pool entry and memory,
// unlinks the entries to evict (the caller destroys them after unlocking).
struct Taken {
    Step st;
    Res *r = nullptr;
    int rc = 0;
    std::string msg;
    long long its = 0;
    int refine = 0;        // refinement steps of the mixed form
    double relres = 0.0;   // worst final relative residual over the load cases
};
Taken take(ncg_server *S, int q, Res *r, const std::vector<Res *> &evict, size_t need, std::vector<Res *> &dead)
{
    Taken t;
    t.st = S->queue[q];
    S->queue.erase(S->queue.begin() + q);
    Job *J = t.st.job;
    for (Res *e : evict) {
        S->resident -= e->bytes;
        for (size_t i = 0; i < S->pool.size(); ++i)
            if (S->pool[i].get() == e) {
                S->pool[i].release();
                S->pool.erase(S->pool.begin() + i);
                break;
            }
        dead.push_back(e);
    }
    if (!r) {
        S->pool.push_back(std::unique_ptr<Res>(new Res()));
        r = S->pool.back().get();
        r->mesh = J->mesh;
    }
    r->busy = true;
    S->resident += need;
    r->bytes += need;
    S->peak = std::max(S->peak, S->resident);
    S->running++;
    J->inflight++;
    if (!J->started) {
        J->started = true;
        J->t_first = clk::now();
        J->state = 1;
    }
    t.r = r;
    return t;
}

void run_lane(ncg_server *S)
{
    std::unique_lock<std::mutex> lk(S->mu);
    clk::time_point t_idle = clk::now();
    for (;;) {
        Res *r = nullptr;
        std::vector<Res *> evict;
        size_t need = 0;
        int q = -1;
        for (;;) {
            if (S->queue.empty()) {
                if (S->stop) return;
                S->cv_work.wait(lk);
                continue;
            }
            q = choose(S, &r, evict, need);
            if (q >= 0) break;
            if (S->running == 0) {
                // nothing runs that could free memory: the head step cannot fit at all
                Job *J = S->queue.front().job;
                fail_job(S, J, NCG_E_MEMORY,
                         "a step of this job needs " + std::to_string(J->bytes >> 20) + " MiB, beyond the " +
                             std::to_string((long long)S->budget >> 20) + " MiB device-memory budget");
                continue;
            }
            S->cv_work.wait(lk);
        }
        // the first step, then as many compatible steps as the group holds and memory allows
        std::vector<Res *> dead;
        std::vector<Taken> grp;
        grp.push_back(take(S, q, r, evict, need, dead));
        const Job *J0 = grp[0].st.job;
        while ((int)grp.size() < S->cfg.group && groupable(J0)) {
            Res *r2 = nullptr;
            std::vector<Res *> ev2;
            size_t need2 = 0;
            const int q2 = choose(S, &r2, ev2, need2, J0);
            if (q2 < 0) break;
            grp.push_back(take(S, q2, r2, ev2, need2, dead));
        }
        const clk::time_point t_start = clk::now();
        S->host_ms[3] += ms_between(t_idle, t_start);
        lk.unlock();

        for (Res *e : dead) {
            ncg_problem_destroy(e->p);
            delete e;
        }
        // problems and operators of every step of the group: every assembly is queued on its
        // problem's stream first, so they run side by side, then their error flags are read
        const int ktime = S->ktime.load();
        for (Taken &t : grp) {
            const Job *J = t.st.job;
            const MeshRec *M = S->meshes[J->mesh].get();
            if (!t.r->p) {
                t.rc = ncg_problem_create(&M->g, S->cfg.device, &t.r->p);
                if (t.rc) t.msg = ncg_last_error();
            }
            if (t.r->p) {
                ncg_set_kernel_timing(t.r->p, ktime);
                ncg_set_graphs(t.r->p, !S->cfg.eager);
            }
            if (!t.rc) {
                t.rc = ncgi::assemble_async(t.r->p, &J->freq[t.st.fi].f, J->precision);
                if (t.rc) t.msg = ncg_last_error();
            }
        }
        for (Taken &t : grp)
            if (!t.rc && (t.rc = ncgi::assemble_wait(t.r->p))) t.msg = ncg_last_error();
        // solve: the group's assembled steps together, or the one step alone
        std::vector<Taken *> ok;
        for (Taken &t : grp)
            if (!t.rc) ok.push_back(&t);
        if (!ok.empty() && gmres_job(J0)) {
            std::vector<ncg_problem *> P;
            for (Taken *t : ok) P.push_back(t->r->p);
            tune_group(S, P, ok[0]->st.job->n, J0->nrhs, J0->precision);
        }
        const clk::time_point t_asm = clk::now();
        int64_t g0[4] = {0, 0, 0, 0}, g1[4] = {0, 0, 0, 0};
        if (!ok.empty()) ncg_graph_stats(ok[0]->r->p, g0);
        if (ok.size() > 1) {
            std::vector<ncg_problem *> P;
            for (Taken *t : ok) P.push_back(t->r->p);
            std::vector<ncg_solve_stats> ss(P.size());
            const int rc = ncg_solve_group(P.data(), (int32_t)P.size(), &J0->solver, ss.data());
            const std::string msg = rc ? ncg_last_error() : "";
            for (size_t i = 0; i < ok.size(); ++i) {
                for (int c = 0; c < ok[i]->st.job->nrhs && c < 64; ++c) {
                    ok[i]->its += ss[i].iterations[c];
                    ok[i]->relres = std::max(ok[i]->relres, ss[i].relative_residual[c]);
                }
                ok[i]->refine = ss[i].refinement_steps;
                if (rc == NCG_E_NOCONV ? !ss[i].converged_all : rc != 0) {
                    ok[i]->rc = rc;
                    ok[i]->msg = msg;
                }
            }
        } else if (ok.size() == 1) {
            ncg_solve_stats ss;
            Taken &t = *ok[0];
            t.rc = ncg_solve(t.r->p, &t.st.job->solver, &ss);
            if (t.rc) t.msg = ncg_last_error();
            for (int c = 0; c < t.st.job->nrhs && c < 64; ++c) {
                t.its += ss.iterations[c];
                t.relres = std::max(t.relres, ss.relative_residual[c]);
            }
            t.refine = ss.refinement_steps;
        }
        if (!ok.empty()) ncg_graph_stats(ok[0]->r->p, g1);
        const clk::time_point t_solve = clk::now();
        // outputs: the solutions, then every grid field queued before the first is read
        std::vector<char> write(grp.size(), 1);
        {
            // a job with a step that failed meanwhile: its outputs are no longer wanted
            std::lock_guard<std::mutex> g(S->mu);
            for (size_t i = 0; i < grp.size(); ++i) write[i] = !grp[i].st.job->failed;
        }
        for (size_t i = 0; i < grp.size(); ++i) {
            Taken &t = grp[i];
            Job *J = t.st.job;
            if (!t.rc && write[i] && J->solution) {
                t.rc = ncg_get_solution(t.r->p, J->solution + 2 * (size_t)J->nrhs * J->n * t.st.fi);
                if (t.rc) t.msg = ncg_last_error();
            }
            if (!t.rc && write[i] && J->m > 0 && (J->phi || J->grad)) {
                t.rc = ncgi::field_begin(t.r->p, J->grad != nullptr);
                if (t.rc) t.msg = ncg_last_error();
            }
        }
        for (size_t i = 0; i < grp.size(); ++i) {
            Taken &t = grp[i];
            Job *J = t.st.job;
            if (t.rc || !write[i] || J->m == 0 || !(J->phi || J->grad)) continue;
            const size_t vec = 2 * (size_t)J->nrhs;
            std::vector<double> tmp;
            double *ph = J->phi ? J->phi + vec * J->m * t.st.fi : nullptr;
            if (!ph) {
                tmp.resize(vec * J->m);
                ph = tmp.data();
            }
            t.rc = ncgi::field_end(t.r->p, ph, J->grad ? J->grad + 3 * vec * J->m * t.st.fi : nullptr);
            if (t.rc) t.msg = ncg_last_error();
        }
        std::vector<std::vector<ncg_kernel_stat>> kst(grp.size());
        for (size_t i = 0; i < grp.size(); ++i) {
            ncg_kernel_stat ks[64];
            int32_t nks = 0;
            if (ktime && grp[i].r->p && ncg_kernel_stats(grp[i].r->p, ks, 64, &nks, 1) == 0)
                kst[i].assign(ks, ks + std::min(nks, 64));
        }

        const clk::time_point t_out = clk::now();
        lk.lock();
        S->host_ms[0] += ms_between(t_start, t_asm);
        S->host_ms[1] += ms_between(t_asm, t_solve);
        S->host_ms[2] += ms_between(t_solve, t_out);
        S->graph_steps += g1[1] - g0[1];
        S->eager_steps += g1[2] - g0[2];
        t_idle = clk::now();
        for (size_t g = 0; g < grp.size(); ++g) {
            Taken &t = grp[g];
            Job *J = t.st.job;
            for (const ncg_kernel_stat &k : kst[g]) {
                ncg_kernel_stat &a = S->kstats[k.name];
                if (a.launches == 0) snprintf(a.name, sizeof(a.name), "%s", k.name);
                a.launches += k.launches;
                a.total_ms += k.total_ms;
                a.bytes += k.bytes;
                a.flops += k.flops;
            }
            S->running--;
            t.r->busy = false;
            t.r->last = ++S->tick;
            if (t.rc && !t.r->p) {
                // the problem could not be created: give its reservation back
                S->resident -= t.r->bytes;
                for (size_t i = 0; i < S->pool.size(); ++i)
                    if (S->pool[i].get() == t.r) {
                        S->pool.erase(S->pool.begin() + i);
                        break;
                    }
            }
            S->steps_done++;
            S->solves_done += J->nrhs;
            J->iterations += t.its;
            J->refine_max = std::max(J->refine_max, t.refine);
            J->relres_max = std::max(J->relres_max, t.relres);
            J->inflight--;
            if (t.rc) {
                fail_job(S, J, t.rc, t.msg);
            } else if (J->failed) {
                if (J->inflight == 0) end_job(S, J);
            } else if (++J->freq_done == (int)J->freq.size()) {
                end_job(S, J);
            }
        }
        S->cv_work.notify_all();
    }
}

void fill_status(const Job *J, ncg_job_status *st)
{
    memset(st, 0, sizeof(*st));
    st->state = J->state;
    st->error_code = J->error_code;
    st->freq_done = J->freq_done;
    st->iterations = (int32_t)std::min<long long>(J->iterations, 0x7fffffff);
    st->refinement_steps_max = J->refine_max;
    st->relative_residual_max = J->relres_max;
    if (J->state >= 2) st->latency_ms = ms_between(J->t_submit, J->t_done);
    if (J->started) st->queue_ms = ms_between(J->t_submit, J->t_first);
    snprintf(st->error, sizeof(st->error), "%s", J->error.c_str());
}

}  // namespace

extern "C" {

int ncg_server_create(const ncg_server_config *cfg, ncg_server **out)
{
    if (!out) return sfail(NCG_E_ARG, "ncg_server_create: null argument");
    *out = nullptr;
    ncg_server_config c{};
    if (cfg) c = *cfg;
    if (c.lanes <= 0) c.lanes = 4;
    if (c.lanes > 16) return sfail(NCG_E_ARG, "ncg_server_create: at most 16 lanes");
    if (c.problems_per_lane <= 0) c.problems_per_lane = 2;
    if (c.group <= 0) c.group = 4;
    if (c.group > 64) return sfail(NCG_E_ARG, "ncg_server_create: at most 64 steps per group");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sfail(NCG_E_DEVICE, "no HIP device available");
    if (c.device < 0 || c.device >= ndev) return sfail(NCG_E_ARG, "device index out of range");
    if (hipSetDevice(c.device) != hipSuccess) return sfail(NCG_E_DEVICE, "hipSetDevice failed");
    double budget = c.memory_budget_bytes;
    if (budget <= 0.0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return sfail(NCG_E_DEVICE, "hipMemGetInfo failed");
        budget = 0.9 * (double)fr;
    }
    ncg_server *S = new ncg_server();
    S->cfg = c;
    S->cfg.tuning_file = nullptr;
    S->ktime = c.kernel_timing ? 1 : 0;
    if (c.tuning_file) S->tuning_file = c.tuning_file;
    if (!S->tuning_file.empty()) load_tunings(S);
    S->budget = budget;
    S->t0 = clk::now();
    for (int i = 0; i < c.lanes; ++i) S->lanes.emplace_back(run_lane, S);
    *out = S;
    return 0;
}

int ncg_server_add_mesh(ncg_server *S, const ncg_geometry *g, int32_t *mesh_id)
{
    if (!S || !g || !mesh_id) return sfail(NCG_E_ARG, "ncg_server_add_mesh: null argument");
    const int n = g->n_elements, m = g->n_eval > 0 ? g->n_eval : 0;
    if (n <= 0 || !g->nodes_per_element || !g->node_xyz || !g->bc_kind || (m > 0 && !g->eval_xyz))
        return sfail(NCG_E_ARG, "ncg_server_add_mesh: missing geometry");
    std::unique_ptr<MeshRec> M(new MeshRec());
    M->nv.assign(g->nodes_per_element, g->nodes_per_element + n);
    M->kind.assign(g->bc_kind, g->bc_kind + n);
    M->xyz.assign(g->node_xyz, g->node_xyz + 12 * (size_t)n);
    if (m) M->eval.assign(g->eval_xyz, g->eval_xyz + 3 * (size_t)m);
    if (g->node_index) M->node.assign(g->node_index, g->node_index + 4 * (size_t)n);
    if (g->element_group) M->group.assign(g->element_group, g->element_group + n);
    M->g = *g;
    M->g.nodes_per_element = M->nv.data();
    M->g.bc_kind = M->kind.data();
    M->g.node_xyz = M->xyz.data();
    M->g.eval_xyz = m ? M->eval.data() : nullptr;
    M->g.node_index = g->node_index ? M->node.data() : nullptr;
    M->g.element_group = g->element_group ? M->group.data() : nullptr;
    ncgi::planner_mesh(&M->g, M->pm, M->po);
    std::lock_guard<std::mutex> lk(S->mu);
    *mesh_id = (int32_t)S->meshes.size();
    S->meshes.push_back(std::move(M));
    return 0;
}

int ncg_server_submit(ncg_server *S, const ncg_job *job, int64_t *job_id)
{
    if (!S || !job || !job_id) return sfail(NCG_E_ARG, "ncg_server_submit: null argument");
    if (job->n_freq <= 0 || !job->freq) return sfail(NCG_E_ARG, "ncg_server_submit: no frequencies");
    if (job->precision != NCG_FP64 && job->precision != NCG_MIXED)
        return sfail(NCG_E_ARG, "ncg_server_submit: precision must be NCG_FP64 or NCG_MIXED");
    MeshRec *M;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (job->mesh < 0 || job->mesh >= (int)S->meshes.size()) return sfail(NCG_E_ARG, "ncg_server_submit: unknown mesh");
        M = S->meshes[job->mesh].get();
    }
    const int n = M->g.n_elements, m = M->g.n_eval > 0 ? M->g.n_eval : 0;
    const int nr = job->freq[0].n_rhs;
    if (nr <= 0 || nr > 64) return sfail(NCG_E_ARG, "ncg_server_submit: n_rhs must be 1..64");
    std::unique_ptr<Job> J(new Job());
    J->mesh = job->mesh;
    J->precision = job->precision;
    J->solver = job->solver;
    J->nrhs = nr;
    J->method_bem = M->g.method.method_bem;
    J->n = n;
    J->m = m;
    J->solution = job->solution;
    J->phi = job->phi;
    J->grad = job->grad;
    double kmax = 0.0;
    J->freq.resize(job->n_freq);
    for (int i = 0; i < job->n_freq; ++i) {
        const ncg_frequency &f = job->freq[i];
        if (f.n_rhs != nr) return sfail(NCG_E_ARG, "ncg_server_submit: every frequency needs the same n_rhs");
        if (!(f.k > 0.0)) return sfail(NCG_E_ARG, "ncg_server_submit: wave numbers must be positive");
        FreqCopy &c = J->freq[i];
        c.f = f;
        if (f.bc_value) c.bc.assign(f.bc_value, f.bc_value + 2 * (size_t)nr * n);
        if (f.admittance) c.adm.assign(f.admittance, f.admittance + 2 * (size_t)n);
        if (f.n_point_sources > 0) c.ps.assign(f.point_sources, f.point_sources + 6 * (size_t)f.n_point_sources);
        if (f.n_plane_waves > 0) c.pw.assign(f.plane_waves, f.plane_waves + 6 * (size_t)f.n_plane_waves);
        c.f.bc_value = f.bc_value ? c.bc.data() : nullptr;
        c.f.admittance = f.admittance ? c.adm.data() : nullptr;
        c.f.point_sources = f.n_point_sources > 0 ? c.ps.data() : nullptr;
        c.f.plane_waves = f.n_plane_waves > 0 ? c.pw.data() : nullptr;
        kmax = std::max(kmax, f.k);
    }
    {
        // the footprint estimate plans the highest frequency; lanes keep running meanwhile
        std::lock_guard<std::mutex> lp(M->pm_mu);
        J->bytes = ncgi::estimate_bytes(M->pm, M->po, m, kmax, job->precision, nr,
                                        job->solver.restart > 0 ? job->solver.restart : 50);
    }
    if (J->bytes == 0) return sfail(NCG_E_ARG, "ncg_server_submit: no FMM plan at the highest frequency");
    std::lock_guard<std::mutex> lk(S->mu);
    J->id = S->next_id++;
    J->t_submit = clk::now();
    Job *jp = J.get();
    S->jobs[J->id] = std::move(J);
    for (int i = 0; i < job->n_freq; ++i) S->queue.push_back(Step{jp, i});
    *job_id = jp->id;
    S->cv_work.notify_all();
    return 0;
}

int ncg_server_query(ncg_server *S, int64_t job_id, ncg_job_status *st)
{
    if (!S || !st) return sfail(NCG_E_ARG, "ncg_server_query: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return sfail(NCG_E_ARG, "ncg_server_query: unknown job");
    fill_status(it->second.get(), st);
    return 0;
}

int ncg_server_wait(ncg_server *S, int64_t job_id, double timeout_s, ncg_job_status *st)
{
    if (!S) return sfail(NCG_E_ARG, "ncg_server_wait: null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->jobs.find(job_id);
    if (it == S->jobs.end()) return sfail(NCG_E_ARG, "ncg_server_wait: unknown job");
    Job *J = it->second.get();
    auto done = [&] { return J->state >= 2; };
    if (timeout_s < 0) S->cv_done.wait(lk, done);
    else S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    if (st) fill_status(J, st);
    if (!done()) return sfail(NCG_E_STATE, "ncg_server_wait: job still running");
    return J->state == 3 ? sfail(J->error_code, J->error) : 0;
}

int ncg_server_stats(ncg_server *S, ncg_server_info *st)
{
    if (!S || !st) return sfail(NCG_E_ARG, "ncg_server_stats: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    memset(st, 0, sizeof(*st));
    st->jobs_done = S->jobs_done;
    st->jobs_failed = S->jobs_failed;
    st->steps_done = S->steps_done;
    st->solves_done = S->solves_done;
    std::vector<double> l = S->latencies;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) {
            const double pos = q * (double)(l.size() - 1);
            const size_t i = (size_t)pos;
            const double f = pos - (double)i;
            return i + 1 < l.size() ? l[i] * (1.0 - f) + l[i + 1] * f : l[i];
        };
        st->p50_ms = pct(0.5);
        st->p99_ms = pct(0.99);
        double s = 0.0;
        for (double v : l) s += v;
        st->mean_ms = s / (double)l.size();
    }
    st->wall_s = ms_between(S->t0, clk::now()) / 1000.0;
    st->resident_bytes = (int64_t)S->resident;
    st->peak_resident_bytes = (int64_t)S->peak;
    st->lanes = S->cfg.lanes;
    st->queued_steps = (int32_t)S->queue.size();
    for (int i = 0; i < 4; ++i) st->lane_ms[i] = S->host_ms[i];
    st->tuned_shapes = (int32_t)S->tunings.size();
    st->graph_steps = S->graph_steps;
    st->eager_steps = S->eager_steps;
    return 0;
}

int ncg_server_kernel_stats(ncg_server *S, ncg_kernel_stat *out, int32_t cap, int32_t *count, int32_t reset)
{
    if (!S || !count || (cap > 0 && !out)) return sfail(NCG_E_ARG, "ncg_server_kernel_stats: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    int c = 0;
    for (auto &kv : S->kstats) {
        if (c < cap) out[c] = kv.second;
        ++c;
    }
    *count = c;
    if (reset) S->kstats.clear();
    return 0;
}

int ncg_server_set_kernel_timing(ncg_server *S, int32_t on)
{
    if (!S) return sfail(NCG_E_ARG, "ncg_server_set_kernel_timing: null server");
    S->ktime = on ? 1 : 0;
    return 0;
}

int ncg_server_destroy(ncg_server *S)
{
    if (!S) return 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    for (std::thread &t : S->lanes) t.join();
    for (auto &e : S->pool)
        if (e->p) ncg_problem_destroy(e->p);
    if (!S->tuning_file.empty()) save_tunings(S);
    delete S;
    return 0;
}

int ncg_server_tuning(ncg_server *S, int32_t index, char *key, int32_t key_cap, ncg_tuning *tuning)
{
    if (!S || !tuning) return sfail(NCG_E_ARG, "ncg_server_tuning: null argument");
    std::lock_guard<std::mutex> lk(S->mu);
    if (index < 0 || index >= (int32_t)S->tunings.size()) return sfail(NCG_E_ARG, "ncg_server_tuning: no such shape");
    auto it = S->tunings.begin();
    std::advance(it, index);
    if (key && key_cap > 0) snprintf(key, (size_t)key_cap, "%s", it->first.c_str());
    *tuning = it->second;
    return 0;
}

}  // extern "C"
