// Drop-in NumCalc: same command line, inputs and outputs as the reference executable
// (mesh2hrtf/NumCalc/src/NC_Main.cpp:137-1048), with every frequency step assembled,
// solved and post-processed on the GPU through libncg (include/ncg.h).
//
// Environment knobs (the command line stays the reference's):
//   NCG_DEVICE=<i>           HIP device (default 0)
//   NCG_PRECISION=fp64|mixed FP64 plain method (default) or FP32 bulk + FP64 refinement
//   NCG_SOLVER=cgs|gmres     Krylov method for the FP64 path (default cgs, as NumCalc)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <filesystem>
#include <iostream>

#include "nc_host.h"
#include "ncg.h"

using nc::cplx;

namespace {

void stamp(std::ostream &os, const char *what)
{
    time_t t = time(nullptr);
    tm *d = localtime(&t);
    os << what << d->tm_mday << "/" << d->tm_mon + 1 << "/" << d->tm_year + 1900 << " " << d->tm_hour << ":"
       << d->tm_min << ":" << d->tm_sec;
}

[[noreturn]] void gpu_fatal(std::ofstream &log, const char *where, int rc)
{
    nc::fatal(log, std::string(where) + " failed (" + std::to_string(rc) + "): " + ncg_last_error());
}

// Which analysis blocks a step writes.  NumCalc rebuilds (and reports) its cluster
// structure at the first step, when the method or the number of levels changes, or
// when the previous structure had cluster pairs cancelled into the near field
// (NC_DeclareArrays src/NC_Addresses.cpp:119-137, 222-223).
struct PlanLog {
    int method = -1, levels = -1;
    long long cancelled = 0;

    bool rebuilt(const ncg_fmm_info &fi, bool first)
    {
        const bool r = first || fi.method != method || fi.levels != levels || cancelled > 0;
        method = fi.method;
        levels = fi.levels;
        if (r) {
            cancelled = 0;
            if (fi.method)
                for (int l = 0; l < fi.levels; ++l) cancelled += fi.canceled_pairs[l];
        }
        return r;
    }
};

// cancelled-pair table (src/NC_Addresses.cpp:225-255), analysis type and cluster table
// (NC_FrequencyInformations src/NC_Main.cpp:1051-1153), written when the structure is rebuilt
void plan_info(std::ostream &os, const ncg_fmm_info &fi, bool first, int n)
{
    static const char *names[4] = {"Traditional BEM", "Single level fast multipole BEM", "",
                                   "Direct multilevel fast multipole BEM"};
    if (fi.method && (fi.canceled_pairs[0] || fi.canceled_pairs[fi.levels - 1])) {
        os << "\nInformations about the canceled interacting cluster pairs:" << std::endl;
        os << "   Level   N. canceled pairs   N. interacting pairs   Cancel ratio (%)" << std::endl;
        for (int l = 0; l < fi.levels; ++l) {
            const double r = fi.canceled_pairs[l] > 0 ? (double)fi.canceled_pairs[l] / (double)fi.far_pairs[l] : 0.0;
            os << "   " << std::setw(3) << l << "     " << std::setw(10) << fi.canceled_pairs[l] << "           "
               << std::setw(10) << fi.far_pairs[l] << "               " << std::setw(5) << r * 100.0 << std::endl;
        }
    }
    os << "\n" << names[fi.method] << std::endl;
    if (first) {
        os << "\nNumber of equations = " << n << std::endl;
        if (fi.method == 0) os << "Number of entries of the coefficient matrix = " << (long long)n * n << std::endl;
    }
    if (fi.method == 0) return;
    os << "\nInformations about clusters:" << std::endl;
    os << "   Level    Num. of clusters    Maximum radius   Minimum radius" << std::endl;
    for (int l = 0; l < fi.levels; ++l)
        os << "   " << std::setw(3) << l << "       " << std::setw(6) << fi.clusters[l] << "                " << std::setw(6)
           << fi.radius_max[l] << "         " << std::setw(6) << fi.radius_min[l] << std::endl;
    if (fi.eval_clusters)
        os << " resu. nets" << "  " << std::setw(6) << fi.eval_clusters << "                " << std::setw(6)
           << fi.eval_radius_max << "         " << std::setw(6) << fi.eval_radius_min << std::endl;
}

// FMM expansion table of every FMM step (src/NC_Main.cpp:491-522); the multilevel method
// starts it with one more empty line (src/NC_Main.cpp:448-449)
void expansion_info(std::ostream &os, const ncg_fmm_info &fi)
{
    if (fi.method == 0) return;
    if (fi.method == 3) os << std::endl;
    os << "\nInformations about the FMM expansion:" << std::endl;
    os << "   Level    N. expa. terms    N. p. theta    N. p. phi    N. p. sphere" << std::endl;
    for (int l = 0; l < fi.levels; ++l)
        os << "   " << std::setw(3) << l << "       " << std::setw(5) << fi.expansion_terms[l] << "           "
           << std::setw(5) << fi.points_theta[l] << "           " << std::setw(5) << fi.points_phi[l] << "         "
           << std::setw(5) << fi.points_sphere[l] << std::endl;
}

// near/T/D/S nonzero summary written when the FMM matrices are set up
// (NC_AllocateSDTmtxsSLFMM src/NC_Addresses.cpp:1580-1595, ...MLFMM :1611-1653)
void fmm_nonzeros(std::ostream &os, const ncg_fmm_info &fi)
{
    long long st = 0, sd = 0, ss = 0;
    for (int l = 0; l < fi.levels; ++l) {
        st += fi.nnz_t[l];
        sd += fi.nnz_d[l];
        ss += fi.nnz_s[l];
    }
    os << "\nNonzeros of the near field matrix ..................... = " << fi.nnz_near << std::endl;
    os << "Nonzeros of the T-matrix .............................. = " << st << std::endl;
    if (fi.method == 3)
        for (int l = 0; l < fi.levels; ++l)
            os << "    Nonzeros of T_lev_" << l << " ............................... = " << fi.nnz_t[l] << std::endl;
    os << (fi.method == 3 ? "Nonzeros of the D-matrices ............................ = "
                          : "Nonzeros of the D-matrix .............................. = ")
       << sd << std::endl;
    if (fi.method == 3)
        for (int l = 0; l < fi.levels; ++l)
            os << "    Nonzeros of D_lev_" << l << " ............................... = " << fi.nnz_d[l] << std::endl;
    os << "Nonzeros of the S-matrix .............................. = " << ss << std::endl;
    if (fi.method == 3)
        for (int l = 0; l < fi.levels; ++l)
            os << "    Nonzeros of S_lev_" << l << " ............................... = " << fi.nnz_s[l] << std::endl;
    os << "Sum of nonzeros of the FMBEM matrices ................. = " << fi.nnz_near + st + sd + ss << std::endl;
}

// RAM estimate of the -estimate_ram run (src/NC_Main.cpp:524-731), with the
// reference's 32-bit counters
double ram_estimate(const ncg_fmm_info &fi, int n)
{
    unsigned int nint = 0, ncmplx = (unsigned int)n;
    if (fi.method == 0) {
        std::cout << "TBEM: \n";
        std::cout << "Number of entries right hand side: " << ncmplx << "\n";
        std::cout << "Number of Matrix entries: " << ncmplx * ncmplx << "\n";
        ncmplx += (unsigned int)n * (unsigned int)n;
    } else {
        const int fin = fi.levels - 1;
        ncmplx += (unsigned int)fi.points_sphere[fin] * 2u;
        nint += (unsigned int)fi.nnz_near + (unsigned int)n + 1u;
        ncmplx += (unsigned int)fi.nnz_near;
        std::cout << (fi.method == 1 ? "SLFMM:\nNearfield: " : "MLFMM: \nNearfield ") << fi.nnz_near << "\n";
        for (int l = 0; l < fi.levels; ++l) {
            const unsigned int P = (unsigned int)fi.points_sphere[l], nc = (unsigned int)fi.clusters[l];
            nint += (unsigned int)fi.nnz_t[l] + nc * P + 1u;
            ncmplx += (unsigned int)fi.nnz_t[l];
            std::cout << (fi.method == 1 ? "T Matrix: " : "Tmatrix level " + std::to_string(l) + ": ") << fi.nnz_t[l] << "\n";
        }
        for (int l = 0; l < fi.levels; ++l) {
            const unsigned int P = (unsigned int)fi.points_sphere[l], nc = (unsigned int)fi.clusters[l];
            nint += (unsigned int)fi.nnz_d[l] + nc * P + (fi.method == 3 ? 1u : 0u);
            ncmplx += (unsigned int)fi.nnz_d[l];
            std::cout << (fi.method == 1 ? "D Matrix: " : "D Matrix level " + std::to_string(l) + ": ") << fi.nnz_d[l] << "\n";
        }
        for (int l = 0; l < fi.levels; ++l) {
            nint += (unsigned int)fi.nnz_s[l] + (unsigned int)n + 1u;
            ncmplx += (unsigned int)fi.nnz_s[l];
            std::cout << (fi.method == 1 ? "S Matrix: " : "S Matrix level " + std::to_string(l) + ": ") << fi.nnz_s[l] << "\n";
        }
    }
    std::cout << "Total Nr. of Integers: " << nint << "\n";
    std::cout << "Total Nr. of Cmplxdble: " << ncmplx << "\n";
    return (double)(nint * sizeof(int) + ncmplx * 16) / 1000.0 / 1000.0 / 1000.0;
}

}  // namespace

int main(int argc, char **argv)
{
    int istart = 0, iend = 0, nitermax = 250;
    bool estimate_ram = false, check_normals = false;
    // command line (src/NC_Main.cpp:150-194)
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "-h")) {
            printf("-istart   int : start index of iteration\n");
            printf("-iend     int : end index of iteration\n");
            printf("-nitermax int : max number of CGS iterations\n");
            printf("-estimate_ram : estimation the RAM consumption of ML-FMM-BEM and write estimate to Memory.txt. "
                   "Estimate is obtained from the number of non-zeros in the FMM matrices.\n");
            printf("-check_normals : check if all normals point to the same domain\n");
            printf("-h            : this message\n");
            return 0;
        } else if (!strcmp(argv[i], "-istart")) {
            if (++i < argc) istart = atoi(argv[i]);
        } else if (!strcmp(argv[i], "-iend")) {
            if (++i < argc) iend = atoi(argv[i]);
        } else if (!strcmp(argv[i], "-nitermax")) {
            if (++i < argc) nitermax = atoi(argv[i]);
        } else if (!strcmp(argv[i], "-estimate_ram")) {
            estimate_ram = true;
        } else if (!strcmp(argv[i], "-check_normals")) {
            check_normals = true;
        } else {
            std::cerr << "\nNumCalc was called with an unknown parameter or flag. Use NumCalc -h for help.\n";
            exit(-1);
        }
    }
    if (istart < 0) {
        std::cerr << "I need to start with a step bigger 0\n";
        exit(-1);
    }
    printf("istart= %d iend = %d\n", istart, iend);
    std::cout << "\n";
    stamp(std::cout, "---------- NumCalc started: ");
    std::cout << " ----------\n" << std::endl;
    char host[150] = {0};
    gethostname(host, sizeof(host) - 1);
    std::cout << "Running on:         " << host << "\n" << std::endl;
    // log file name (src/NC_Main.cpp:218-233)
    char fname[200];
    if (istart > 0) {
        if (iend >= istart) snprintf(fname, sizeof(fname), "NC%d-%d.out", istart, iend);
        else if (iend > 0) {
            std::cerr << "istart is bigger than iend. Stopping now!\n";
            exit(-1);
        } else snprintf(fname, sizeof(fname), "NCfrom%d.out", istart);
    } else {
        if (iend > 0) snprintf(fname, sizeof(fname), "NCuntil%d.out", iend);
        else snprintf(fname, sizeof(fname), "NC.out");
    }
    std::ofstream log(fname);
    if (!log) nc::fatal(log, "can not open the output stream NCout!");
    log << "\n";
    stamp(log, "Start time: ");
    log << std::endl;
    if (istart == 0 && iend == 0 && !estimate_ram) {
        std::error_code ec;
        std::filesystem::remove_all("be.out", ec);
    }
    mkdir("be.out", S_IRWXU | S_IRGRP | S_IWGRP | S_IXGRP | S_IROTH | S_IXOTH);

    const time_t t0 = time(nullptr);
    nc::Project P;
    nc::read_project(log, P, check_normals);
    std::ofstream estim;
    if (estimate_ram) estim.open("Memory.txt");
    const time_t t1 = time(nullptr);

    const char *env = getenv("NCG_PRECISION");
    const int precision = env && !strcmp(env, "mixed") ? NCG_MIXED : NCG_FP64;
    env = getenv("NCG_SOLVER");
    const int gmres = env && !strcmp(env, "gmres");
    env = getenv("NCG_DEVICE");
    const int device = env ? atoi(env) : 0;

    // boundary-element order: internal element order, BE elements only
    std::vector<int> be;
    for (int e = 0; e < P.nelem; ++e)
        if (P.elem_prop[e] == 0) be.push_back(e);
    const int n = (int)be.size();
    std::vector<double> evalp;
    for (int k = 0; k < P.nnode; ++k)
        if (P.node_class[k] == -1)
            for (int d = 0; d < 3; ++d) evalp.push_back(P.node_xyz[3 * (size_t)k + d]);
    const int m = (int)evalp.size() / 3;

    // boundary elements unfolded across the symmetry planes (nc::Symmetry)
    const nc::Symmetry sym(P);
    const int nref = sym.nref, N = n * nref;
    ncg_problem *prob = nullptr;
    std::vector<int> nvv, kind, nodes, groups;
    std::vector<double> xyz;
    nc::unfold_mesh(P, be, sym, nvv, kind, nodes, groups, xyz);
    ncg_geometry g;
    memset(&g, 0, sizeof(g));
    g.node_index = nodes.data();
    g.element_group = groups.data();
    g.method.method_bem = P.method_bem;
    g.method.min_expansion_terms = P.min_terms;
    g.method.cluster_edge = P.clus_edge;
    g.n_elements = N;
    g.nodes_per_element = nvv.data();
    g.node_xyz = xyz.data();
    g.bc_kind = kind.data();
    g.n_eval = m;
    g.eval_xyz = m ? evalp.data() : nullptr;
    g.harmonic_time_factor = P.h;
    g.interior = P.interior;
    if (!estimate_ram) {
        const int rc = ncg_problem_create(&g, device, &prob);
        if (rc) gpu_fatal(log, "ncg_problem_create", rc);
    }
    const time_t t2 = time(nullptr);

    if (iend > 0 && iend > P.nfreq) iend = P.nfreq;
    if (iend == 0) iend = P.nfreq;
    if (istart > 0) istart--;
    long lti_eqa = 0, lti_sol = 0, lti_pos = 0;
    time_t tlast = t2;
    PlanLog plan_log;
    auto log_plan = [&](const ncg_fmm_info &fi, bool first) {
        const bool rebuilt = plan_log.rebuilt(fi, first);
        for (std::ostream *os : {(std::ostream *)&std::cout, (std::ostream *)&log}) {
            if (rebuilt) plan_info(*os, fi, first, n);
            expansion_info(*os, fi);
        }
        if (fi.warning[0]) nc::warning(log, fi.warning);
    };
    for (int step = istart; step < iend; ++step) {
        const time_t ts = time(nullptr);
        const double freq = P.freqs[step];
        const double omega = 2.0 * 3.1415926535897932385 * freq;
        const double rpfact = P.rho * omega * P.h;
        const double k = omega / P.c;
        std::cout << "\nStep " << step + 1 << ", Frequency = " << freq << " Hz" << std::endl;
        log << "\n\n\n>> S T E P   N U M B E R   A N D   F R E Q U E N C Y <<\n" << std::endl;
        log << "Step " << step + 1 << ", Frequency = " << freq << " Hz" << std::endl;
        if (estimate_ram) {
            ncg_fmm_info fi;
            const int rc = ncg_fmm_plan(&g, k, &fi);
            if (rc) gpu_fatal(log, "ncg_fmm_plan", rc);
            log_plan(fi, step == istart);
            const double gb = ram_estimate(fi, N);
            std::cout << "RAM Estimation: " << gb << " GByte\n";
            estim << step + 1 << " " << freq << " " << gb << "\n";
            continue;
        }
        // per-frequency values (NC_UpdateFreqCurves, src/NC_EquationSystem.cpp:1815-1961)
        std::vector<double> bcv(2 * (size_t)N), adm(2 * (size_t)N, 0.0);
        for (int q = 0; q < n; ++q) {
            const int e = be[q];
            cplx v = P.bc0[e];
            if (P.ibval[e] == 1) v = v / cplx(0.0, rpfact);
            bcv[2 * q] = v.real();
            bcv[2 * q + 1] = v.imag();
            if (P.ibval[e] == 2) {
                cplx a = P.adm1[e];
                const int cr = P.curve_of[6 * (size_t)e + 2], ci = P.curve_of[6 * (size_t)e + 3];
                if (cr >= 0) a = cplx(a.real() * nc::curve_factor(log, P, cr, freq), a.imag());
                if (ci >= 0) a = cplx(a.real(), a.imag() * nc::curve_factor(log, P, ci, freq));
                a *= cplx(0.0, k * P.h);
                adm[2 * q] = a.real();
                adm[2 * q + 1] = a.imag();
            }
        }
        // mirror images: prescribed velocity and pressure times the image's sign, the same
        // admittance
        for (int r = 1; r < nref; ++r)
            for (int q = 0; q < n; ++q) {
                const size_t Q = (size_t)r * n + q;
                bcv[2 * Q] = sym.sign(r) * bcv[2 * q];
                bcv[2 * Q + 1] = sym.sign(r) * bcv[2 * q + 1];
                adm[2 * Q] = adm[2 * q];
                adm[2 * Q + 1] = adm[2 * q + 1];
            }
        auto scaled = [&](const nc::Source &s) {
            cplx z = s.amp0;
            if (s.curve[0] >= 0) z = cplx(z.real() * nc::curve_factor(log, P, s.curve[0], freq), z.imag());
            if (s.curve[1] >= 0) z = cplx(z.real(), z.imag() * nc::curve_factor(log, P, s.curve[1], freq));
            return z / cplx(0.0, rpfact);
        };
        // sources and their mirror images (NC_IncidentWaveRHS src/NC_EquationSystem.cpp:1735-1810):
        // an image carries the sign of its mirrors; with symmetry planes the plane-wave phase
        // is taken relative to the planes, exp(i h k d_r . (x - c)) with c the plane coordinates
        std::vector<double> ps, pw;
        for (const nc::Source &s : P.ps)
            for (int r = 0; r < nref; ++r) {
                const cplx z = scaled(s) * sym.sign(r);
                double x[3] = {s.xyz[0], s.xyz[1], s.xyz[2]};
                sym.mirror(r, x);
                ps.insert(ps.end(), {x[0], x[1], x[2], z.real(), z.imag(), 0.0});
            }
        for (const nc::Source &s : P.pw)
            for (int r = 0; r < nref; ++r) {
                double dir[3] = {s.xyz[0], s.xyz[1], s.xyz[2]}, dc = 0.0;
                for (size_t b = 0; b < sym.dims.size(); ++b) {
                    if (r >> b & 1) dir[sym.dims[b]] = -dir[sym.dims[b]];
                    dc += dir[sym.dims[b]] * P.sym_coord[sym.dims[b]];
                }
                const cplx z = scaled(s) * sym.sign(r) * std::exp(cplx(0.0, -P.h * k * dc));
                pw.insert(pw.end(), {dir[0], dir[1], dir[2], z.real(), z.imag(), 0.0});
            }
        // NumCalc adds the incident waves to the right-hand side only when
        // numIncidentPlaneWaves_ + numPointSources_ > 0 (src/NC_EquationSystem.cpp:247,1286),
        // but to the evaluation grid whenever either count is nonzero
        // (src/NC_PostProcessing.cpp:640,664): with the test point source (-1) and one plane
        // wave, the wave reaches the grid only.
        const bool rhs_waves = P.npw + P.nps > 0;
        ncg_frequency fr;
        fr.k = k;
        fr.n_rhs = 1;
        fr.bc_value = bcv.data();
        fr.admittance = adm.data();
        fr.n_point_sources = (int)ps.size() / 6;
        fr.point_sources = ps.empty() ? nullptr : ps.data();
        fr.n_plane_waves = rhs_waves ? (int)pw.size() / 6 : 0;
        fr.plane_waves = pw.empty() ? nullptr : pw.data();
        int rc = ncg_assemble(prob, &fr, precision);
        if (rc) gpu_fatal(log, "ncg_assemble", rc);
        ncg_fmm_info fi;
        rc = ncg_problem_fmm_info(prob, &fi);
        if (rc) gpu_fatal(log, "ncg_problem_fmm_info", rc);
        log_plan(fi, step == istart);
        if (fi.method) fmm_nonzeros(log, fi);
        const time_t ta = time(nullptr);
        // solve (src/NC_Main.cpp:833-866, NC_IterativeSolverCGS)
        ncg_solver_options so;
        memset(&so, 0, sizeof(so));
        so.solver = P.method_solver == 4 ? NCG_SOLVER_LU : (gmres ? NCG_SOLVER_GMRES : NCG_SOLVER_CGS);
        so.precond = P.precond == 0 ? NCG_PRECOND_ILU : P.precond == 1 ? NCG_PRECOND_SCALING : NCG_PRECOND_NONE;
        if ((precision == NCG_MIXED || so.solver == NCG_SOLVER_GMRES) && so.solver != NCG_SOLVER_LU) {
            // GMRES and the mixed form (FP64 refinement around FP32 GMRES) run with the
            // block-Jacobi preconditioner, whatever IPRECOND says
            so.solver = NCG_SOLVER_GMRES;
            so.precond = NCG_PRECOND_BLOCK;
        }
        so.ilu_scanning_degree = P.scan_degree;
        so.max_iterations = nitermax;
        so.tolerance = 1.0e-9;
        so.restart = 50;
        so.refinement_steps = 8;
        so.refinement_tolerance = 1.0e-6;
        ncg_solve_stats st;
        rc = ncg_solve(prob, &so, &st);
        if (so.solver != NCG_SOLVER_LU) {
            if (rc == NCG_E_NOCONV) nc::warning(log, "Warning: Maximum number of iterations is reached!");
            else if (rc) gpu_fatal(log, "ncg_solve", rc);
            log << "\n" << (so.solver == NCG_SOLVER_CGS ? "CGS" : "GMRES")
                << " solver: number of iterations = " << st.iterations[0] << ", relative error = " << st.relative_residual[0]
                << std::endl;
            if (rc == NCG_E_NOCONV && precision == NCG_FP64 && fi.method == 0) {
                // TBEM fallback of src/NC_CommonFunctions.cpp:1181-1191, solved with the
                // assembled right-hand side (the reference passes the unconverged iterate)
                nc::warning(log, "Iteration method CGS failed, direct method is used!");
                so.solver = NCG_SOLVER_LU;
                rc = ncg_assemble(prob, &fr, precision);
                if (!rc) rc = ncg_solve(prob, &so, &st);
                if (rc) gpu_fatal(log, "ncg_solve (direct)", rc);
            }
        } else if (rc) {
            gpu_fatal(log, "ncg_solve (direct)", rc);
        }
        std::vector<cplx> x(N);
        rc = ncg_get_solution(prob, (double *)x.data());
        if (rc) gpu_fatal(log, "ncg_get_solution", rc);
        const time_t tsol = time(nullptr);
        // surface values (src/NC_PostProcessing.cpp:399-421)
        std::vector<cplx> pot(P.nelem), vel(P.nelem);
        for (int q = 0; q < n; ++q) {
            const int e = be[q];
            const cplx b(bcv[2 * q], bcv[2 * q + 1]);
            if (P.ibval[e] == 1) { pot[e] = b; vel[e] = x[q]; }
            else {
                pot[e] = x[q];
                vel[e] = b;
                if (P.ibval[e] == 2) vel[e] -= x[q] * cplx(adm[2 * q], adm[2 * q + 1]) * (P.interior ? -1.0 : 1.0);
            }
        }
        nc::write_boundary(log, P, step + 1, pot, vel, rpfact);
        if (m > 0) {
            std::vector<cplx> phi(m), grad(3 * (size_t)m);
            rc = ncg_field(prob, (double *)phi.data(), (double *)grad.data());
            if (rc) gpu_fatal(log, "ncg_field", rc);
            if (!rhs_waves)   // the grid-only plane waves: exp(i h k d.x) and its gradient
                for (size_t w = 0; w < pw.size(); w += 6)
                    for (int q = 0; q < m; ++q) {
                        const double *x = &evalp[3 * (size_t)q];
                        const cplx g = cplx(pw[w + 3], pw[w + 4]) *
                                       std::exp(cplx(0.0, P.h * k * (pw[w] * x[0] + pw[w + 1] * x[1] + pw[w + 2] * x[2])));
                        phi[q] += g;
                        for (int d = 0; d < 3; ++d) grad[3 * (size_t)q + d] += cplx(0.0, P.h * k) * g * pw[w + d];
                    }
            nc::write_evalgrid(log, P, step + 1, phi, grad, rpfact);
        }
        const time_t tp = time(nullptr);
        const long e1 = (long)(ta - ts), e2 = (long)(tsol - ta), e3 = (long)(tp - tsol);
        lti_eqa += e1;
        lti_sol += e2;
        lti_pos += e3;
        double t[8];
        ncg_timings(prob, t, 0);
        for (std::ostream *os : {(std::ostream *)&log, (std::ostream *)&std::cout}) {
            *os << "\nTime Statistic For The Current Step (In Second)\n" << std::endl;
            *os << "Assembling the equation system         : " << e1 << std::endl;
            *os << "Solving the equation system            : " << e2 << std::endl;
            *os << "Post processing                        : " << e3 << std::endl;
            *os << "Total                                  : " << e1 + e2 + e3 << std::endl;
        }
        // device times go to the console only: NC.out keeps NumCalc's text line for line
        std::cout << "GPU time (ms): assembly " << t[0] << ", solve " << t[1] << ", evaluation grid " << t[2]
                  << std::endl;
        tlast = tp;
    }
    if (!estimate_ram) {
        log << "\n\n\n>> T I M E   S T A T I S T I C   (In   Second) <<\n" << std::endl;
        log << "Input                                  : " << t1 - t0 << std::endl;
        log << "Address computation                    : " << t2 - t1 << std::endl;
        log << "Assembling the equation system         : " << lti_eqa << std::endl;
        log << "Solving the equation system            : " << lti_sol << std::endl;
        log << "Post processing                        : " << lti_pos << std::endl;
        log << "The whole job                          : " << tlast - t0 << std::endl;
    }
    if (prob) ncg_problem_destroy(prob);
    log << "\n";
    stamp(log, "End time: ");
    log << std::endl;
    std::cout << "\n";
    stamp(std::cout, "---------- NumCalc ended: ");
    std::cout << " ----------" << std::endl;
    return 0;
}
