// NC.inp and mesh-file reader of the drop-in NumCalc.  Record order, tokenizer and
// error behaviour follow mesh2hrtf/NumCalc/src/NC_Input.cpp (cited per function).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <iomanip>
#include <iostream>
#include <unordered_map>

#include "nc_host.h"

namespace nc {

namespace {

constexpr int kLine = 300;   // SIZE, src/NC_ConstantsVariables.h:34
constexpr int kTerms = 30;   // NTRM

// NC_GetLine (src/NC_Input.cpp:1596-1631): split on ' ' only, the last character of the
// line is taken to be the newline, a first token containing '#' makes the line empty,
// end of file is fatal.
int get_line(FILE *f, std::vector<std::string> &tok)
{
    char buf[kLine];
    if (!fgets(buf, kLine, f)) {
        std::cout << "\nError parsing the input file\n";
        exit(-1);
    }
    std::string line(buf);
    tok.assign(kTerms, std::string());
    int n = 0;
    const int len = (int)line.size() - 1;
    int ss = 0, nn = 0;
    while (nn < len) {
        size_t pos = line.find(' ', ss);
        nn = pos == std::string::npos ? len : (int)pos;
        if (nn - ss > 0 && n < kTerms) tok[n++] = line.substr(ss, nn - ss);
        ss = nn + 1;
    }
    if (n > 0 && tok[0].find('#') != std::string::npos) n = 0;
    return n;
}

int next(FILE *f, std::vector<std::string> &tok)
{
    int n;
    while (!(n = get_line(f, tok))) {
    }
    return n;
}

// NC_String2Integer / NC_String2Double (src/NC_Input.cpp:1635-1683)
int to_int(const std::string &s)
{
    if (s.size() > 15) {
        std::cout << "Length of the string > n_chwk!" << std::endl << "The string: " << s << std::endl;
        exit(1);
    }
    return atoi(s.c_str());
}

double to_double(const std::string &s)
{
    if (s.size() > 25) {
        std::cout << "Length of the string > n_chwk!" << std::endl << "The string: " << s << std::endl;
        exit(1);
    }
    return atof(s.c_str());
}

std::string str(int v) { return std::to_string(v); }

// NC_ComputeFrequencies (src/NC_Input.cpp:1686-1732): linear interpolation of the
// (time, frequency) curve at t = first + (i+1)*increment
void compute_frequencies(Project &P, const std::vector<double> &ct, const std::vector<double> &cf)
{
    const int np = (int)ct.size();
    P.freqs.assign(P.nfreq, 0.0);
    for (int i = 0; i < P.nfreq; ++i) {
        const double t = P.freq_first + (double)(i + 1) * P.freq_inc;
        if (ct[0] >= t) { P.freqs[i] = cf[0]; continue; }
        int j1 = -1;
        for (int j = 1; j < np; ++j) {
            if (ct[j] > t) { j1 = j; break; }
            if (ct[j] == t) { P.freqs[i] = cf[j]; j1 = np; break; }
        }
        if (j1 == -1) P.freqs[i] = cf[np - 1];
        else if (j1 < np)
            P.freqs[i] = cf[j1 - 1] + (t - ct[j1 - 1]) * (cf[j1] - cf[j1 - 1]) / (ct[j1] - ct[j1 - 1]);
    }
}

// element centre normal, Jacobian at (s,t) (NC_ComputeUnitVectorOnElement)
double jacobian(const Project &P, int e, double s, double t, double *nrm)
{
    const int nv = P.elem_nv[e];
    double ds[4], dt[4], xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
    if (nv == 3) {
        ds[0] = dt[1] = 1.0; ds[1] = dt[0] = 0.0; ds[2] = dt[2] = -1.0;
    } else {
        ds[0] = 0.25 * (t + 1.0); ds[2] = 0.25 * (t - 1.0); ds[1] = -ds[0]; ds[3] = -ds[2];
        dt[0] = 0.25 * (s + 1.0); dt[2] = 0.25 * (s - 1.0); dt[1] = -dt[2]; dt[3] = -dt[0];
    }
    for (int a = 0; a < nv; ++a) {
        const double *x = &P.node_xyz[3 * (size_t)P.elem_conn[4 * (size_t)e + a]];
        for (int d = 0; d < 3; ++d) { xs[d] += ds[a] * x[d]; xt[d] += dt[a] * x[d]; }
    }
    double n0 = xs[1] * xt[2] - xs[2] * xt[1], n1 = xs[2] * xt[0] - xs[0] * xt[2], n2 = xs[0] * xt[1] - xs[1] * xt[0];
    double L = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    if (nrm) { nrm[0] = n0 / L; nrm[1] = n1 / L; nrm[2] = n2 / L; }
    return L;
}

void read_mesh_files(std::ofstream &log, FILE *in, Project &P, bool check_normals)
{
    std::vector<std::string> tok;
    next(in, tok);
    if (tok[0] != "NODES") fatal(log, "Key word NODES expected!");
    P.node_ext.assign(P.nnode, 0);
    P.node_xyz.assign(3 * (size_t)P.nnode, 0.0);
    int i = 0;
    while (i < P.nnode) {
        next(in, tok);
        FILE *f = fopen(tok[0].c_str(), "r");
        if (!f) {
            std::cout << "No Nodes file found!" << std::endl;
            log << "No Nodes file found!" << std::endl;
            exit(-1);
        }
        std::vector<std::string> t;
        next(f, t);
        const int cnt = to_int(t[0]);
        for (int k = 0; k < cnt; ++k) {
            next(f, t);
            if (i >= P.nnode) fatal(log, "More nodes in the node files than declared in NC.inp!");
            P.node_ext[i] = to_int(t[0]);
            for (int d = 0; d < 3; ++d) P.node_xyz[3 * (size_t)i + d] = to_double(t[1 + d]);
            ++i;
        }
        fclose(f);
    }
    next(in, tok);
    if (tok[0] != "ELEMENTS") fatal(log, "Key word ELEMENTS expected!");
    P.elem_ext.assign(P.nelem, 0);
    P.elem_nv.assign(P.nelem, 0);
    P.elem_prop.assign(P.nelem, 0);
    P.elem_group.assign(P.nelem, 0);
    P.elem_conn.assign(4 * (size_t)P.nelem, 0);
    i = 0;
    while (i < P.nelem) {
        next(in, tok);
        FILE *f = fopen(tok[0].c_str(), "r");
        if (!f) {
            std::cout << "No Elements file found!" << std::endl;
            log << "No Elements file found!" << std::endl;
            exit(-1);
        }
        std::vector<std::string> t;
        next(f, t);
        const int cnt = to_int(t[0]);
        for (int k = 0; k < cnt; ++k) {
            const int nt = next(f, t);
            if (nt < 7 || nt > 8)
                fatal(log, "number of entris of a input line must be 7 or 8!",
                      "number of the element under the key word ELEMENTS = " + str(to_int(t[0])) +
                          "\nnumber of entrie of the line = " + str(nt));
            if (i >= P.nelem) fatal(log, "More elements in the element files than declared in NC.inp!");
            P.elem_ext[i] = to_int(t[0]);
            P.elem_nv[i] = nt - 4;
            for (int a = 0; a < P.elem_nv[i]; ++a) P.elem_conn[4 * (size_t)i + a] = to_int(t[1 + a]);
            P.elem_prop[i] = to_int(t[nt - 3]);
            P.elem_group[i] = to_int(t[nt - 1]);
            ++i;
        }
        fclose(f);
    }
    // BE / evaluation split (src/NC_Input.cpp:437-457)
    int eval_group = 0, first_eval = -1;
    for (int e = 0; e < P.nelem; ++e) {
        if (P.elem_prop[e] == 0) P.n_be++;
        else if (P.elem_prop[e] == 2) {
            if (first_eval < 0) { first_eval = e; eval_group = P.elem_group[e]; }
            else if (P.elem_group[e] != eval_group)
                fatal(log, "Elements of the evaluation mesh must be of the same el. group!");
            P.n_eval_elem++;
        } else
            fatal(log, "listElementProperty[i] must be 0 or 2!");
    }
    // symmetry planes (src/NC_Input.cpp:459-495)
    if (P.nsym > 0) {
        next(in, tok);
        if (tok[0] != "SYMMETRY") fatal(log, "Key word SYMMETRY expected!");
        next(in, tok);
        for (int d = 0; d < 3; ++d) P.sym_planes[d] = to_int(tok[d]);
        next(in, tok);
        for (int d = 0; d < 3; ++d) P.sym_coord[d] = to_double(tok[d]);
        // plane types 0 none, 1 symmetric, 2 antisymmetric, 3 and 4 read as +-2 (only the
        // sign is used by the solve: negative = antisymmetric)
        int cnt = 0;
        for (int d = 0; d < 3; ++d) {
            int &t = P.sym_planes[d];
            if (t) cnt++;
            if (t == 2) t = -1;
            else if (t == 3) t = 2;
            else if (t == 4) t = -2;
            else if (t != 0 && t != 1)
                fatal(log, "listSymmetryPlanes[i] must be equal to 0, 1, 2, 3, or 4!",
                      "i = " + str(d) + "\nlistSymmetryPlanes[i] = " + str(t));
        }
        if (cnt != P.nsym)
            fatal(log, "j must equal to numSymmetricPlanes_", "j = " + str(cnt) + "\nnumSymmetricPlanes_ = " + str(P.nsym));
    }
    // external node numbers -> internal indices (src/NC_Input.cpp:497-511)
    std::unordered_map<int, int> idx;
    idx.reserve(P.nnode * 2);
    for (int k = P.nnode - 1; k >= 0; --k) idx[P.node_ext[k]] = k;   // first occurrence wins
    for (int e = 0; e < P.nelem; ++e)
        for (int a = 0; a < P.elem_nv[e]; ++a) {
            auto it = idx.find(P.elem_conn[4 * (size_t)e + a]);
            if (it == idx.end())
                fatal(log, "Node of element does not exist in the connectivity list!",
                      "Number of element = " + str(P.elem_ext[e]) + "\nLocal number of the node in the element = " + str(a));
            P.elem_conn[4 * (size_t)e + a] = it->second;
        }
    if (check_normals) {
        // NC_checkedges (src/NC_Input.cpp:741-789): every shared edge of the BE mesh must
        // be traversed in opposite directions by its two elements
        std::unordered_map<long long, int> dir;
        bool ok = true;
        for (int e = 0; e < P.nelem && ok; ++e) {
            if (P.elem_prop[e] != 0) continue;
            const int nv = P.elem_nv[e];
            for (int a = 0; a < nv; ++a) {
                const long long u = P.elem_conn[4 * (size_t)e + a], v = P.elem_conn[4 * (size_t)e + (a + 1) % nv];
                const long long key = u * (long long)P.nnode + v;
                if (dir.count(key)) { ok = false; break; }
                dir[key] = e;
            }
        }
        if (!ok) fatal(log, "There seems to be at least one element that is differently ordered to the rest");
    }
    // areas, average BE area (src/NC_Input.cpp:522-605)
    P.area.assign(P.nelem, 0.0);
    int nbe = 0;
    for (int e = 0; e < P.nelem; ++e) {
        const bool tri = P.elem_nv[e] == 3;
        const double s = tri ? 1.0 / 3.0 : 0.0;
        P.area[e] = jacobian(P, e, s, s, nullptr) * (tri ? 0.5 : 4.0);
        if (P.elem_prop[e] == 0) { P.avg_area += P.area[e]; nbe++; }
    }
    P.avg_area /= (double)nbe;
    // node classes (src/NC_Input.cpp:607-656)
    P.node_class.assign(P.nnode, -3);
    for (int e = 0; e < P.nelem; ++e)
        if (P.elem_prop[e] == 0)
            for (int a = 0; a < P.elem_nv[e]; ++a) P.node_class[P.elem_conn[4 * (size_t)e + a]] = 1;
    for (int e = 0; e < P.nelem; ++e)
        if (P.elem_prop[e] == 2)
            for (int a = 0; a < P.elem_nv[e]; ++a) {
                int &c = P.node_class[P.elem_conn[4 * (size_t)e + a]];
                if (c == 1)
                    fatal(log, "BE mesh and evaluation mesh have common node!",
                          "Number of the common node = " + str(P.node_ext[P.elem_conn[4 * (size_t)e + a]]));
                c = -1;
            }
    for (int k = 0; k < P.nnode; ++k) {
        if (P.node_class[k] == 1) P.node_class[k] = P.n_be_nodes++;
        else if (P.node_class[k] == -1) P.n_eval_nodes++;
        else {
            std::cout << "\nnumber of the isolated node: " << P.node_ext[k] << std::endl;
            log << "\nnumber of the isolated node: " << P.node_ext[k] << std::endl;
            fatal(log, "Isolated node found!");
        }
    }
    if (P.n_be_nodes == 0) fatal(log, "Number of boundary elements is zero!");
    // BE bounding-box diagonal (src/NC_Input.cpp:659-703); across a symmetry plane the box
    // is the one of the mesh and its mirror image, and the mesh must lie on one side
    double mx[3] = {-9.9e20, -9.9e20, -9.9e20}, mn[3] = {9.9e20, 9.9e20, 9.9e20};
    for (int k = 0; k < P.nnode; ++k) {
        if (P.node_class[k] < 0) continue;
        for (int d = 0; d < 3; ++d) {
            mx[d] = std::max(mx[d], P.node_xyz[3 * (size_t)k + d]);
            mn[d] = std::min(mn[d], P.node_xyz[3 * (size_t)k + d]);
        }
    }
    const double dtl = sqrt(P.avg_area) / 100.0;
    P.diameter = 0.0;
    for (int d = 0; d < 3; ++d) {
        double ext = mx[d] - mn[d];
        if (P.nsym > 0 && P.sym_planes[d] != 0) {
            const double c = P.sym_coord[d];
            if (mx[d] >= c - dtl && mn[d] >= c - dtl) ext = 2.0 * (mx[d] - c);
            else if (mx[d] <= c + dtl && mn[d] <= c + dtl) ext = 2.0 * (c - mn[d]);
            else {
                warning(log, "", "Number of the symmetric plane = " + str(d));
                fatal(log, "All BE nodes must be located on the same side of a symmetric plane!",
                      "crdmaxmin[j][0] - listSymmetryPlanesCoordinates[j] = " + std::to_string(mx[d] - c) +
                          "\ncrdmaxmin[j][1] - listSymmetryPlanesCoordinates[j] =  = " + std::to_string(mn[d] - c));
            }
        }
        P.diameter += ext * ext;
    }
    P.diameter = sqrt(P.diameter);
    // element groups (src/NC_Input.cpp:705-736)
    for (int e = 0; e < P.nelem; ++e) {
        bool found = false;
        for (size_t g = 0; g < P.group_index.size(); ++g)
            if (P.group_index[g] == P.elem_group[e]) { found = true; break; }
        if (!found) { P.group_index.push_back(P.elem_group[e]); P.group_prop.push_back(P.elem_prop[e]); }
    }
    if ((int)P.group_index.size() < P.ngroups) P.ngroups = (int)P.group_index.size();
    else if ((int)P.group_index.size() > P.ngroups)
        fatal(log, "ielgrp must be less than or equal to nelgrp",
              "ielgrp = " + str((int)P.group_index.size()) + "\nnelgrp = " + str(P.ngroups));
    P.group_count.assign(P.ngroups, 0);
    for (int e = 0; e < P.nelem; ++e)
        for (int g = 0; g < P.ngroups; ++g)
            if (P.elem_group[e] == P.group_index[g]) { P.group_count[g]++; break; }
}

struct BcLine {
    int lo, hi;
    std::string key;
    double re, im;
    bool ref;
    int cre, cim;
};

// NC_StoreBoundaryConditions (src/NC_Input.cpp:1163-1406), no transfer admittances
void store_bc(std::ofstream &log, Project &P, const std::vector<BcLine> &lines)
{
    std::unordered_map<int, int> loc;
    for (int e = P.nelem - 1; e >= 0; --e)
        if (P.elem_prop[e] != 2) loc[P.elem_ext[e]] = e;
    P.ibval.assign(P.nelem, 0);
    P.bc0.assign(P.nelem, cplx(0, 0));
    P.adm1.assign(P.nelem, cplx(0, 0));
    P.curve_of.assign(6 * (size_t)P.nelem, -1);
    std::vector<int> impe_curve(P.curves.size(), 0);
    for (size_t l = 0; l < lines.size(); ++l) {
        const BcLine &b = lines[l];
        int typ;
        bool impe = false;
        if (b.key == "VELO") typ = 0;
        else if (b.key == "PRES") typ = 1;
        else if (b.key == "ADMI") typ = 2;
        else if (b.key == "IMPE") { typ = 2; impe = true; }
        else if (b.key == "TRAD") fatal(log, "Transfer admittance boundary condition is not supported!");
        else if (b.key == "TRIM") fatal(log, "Transfer impedance boundary condition is not supported!");
        else {
            std::cout << "The false key word = " << b.key << std::endl;
            log << "The false key word = " << b.key << std::endl;
            fatal(log, "False key word in a input line for boundary conditions!", "Input line number = " + str((int)l + 1));
        }
        double re = b.re, im = b.im;
        if (impe && re == 0 && im == 0) { impe = false; typ = 1; }
        if (b.ref && impe) {
            std::cerr << "Sorry Impedance boundary conditions in CURVES is not implemented yet. Please use ADMI "
                         "conditions. They work with CURVES. Sorry for the inconvinience.\n";
            exit(-1);
        }
        if (impe) {
            const double d = re * re + im * im;
            re /= d;
            im /= -d;
        }
        for (int j = b.lo; j <= b.hi; ++j) {
            auto it = loc.find(j);
            if (it == loc.end()) continue;
            const int e = it->second;
            switch (typ) {
            case 0: P.bc0[e] = cplx(re, im); break;
            case 1: P.bc0[e] = cplx(re, im); P.ibval[e] = 1; break;
            case 2:
                P.adm1[e] = cplx(re, im);
                if (b.ref) { P.curve_of[6 * (size_t)e + 2] = b.cre; P.curve_of[6 * (size_t)e + 3] = b.cim; }
                if (P.ibval[e] == 0) P.ibval[e] = 2;
                else if (P.ibval[e] == 1)
                    std::cout << "Warning; for el. " << j << " the pressure boundary condition is "
                              << "prescribed, the additional admittance boundary condition is ignored!" << std::endl;
                break;
            }
        }
    }
}

}  // namespace

[[noreturn]] void fatal(std::ofstream &log, const std::string &msg, const std::string &extra)
{
    std::cout << "\nError: " << msg << std::endl;
    if (!extra.empty()) std::cout << extra << std::endl;
    if (log) {
        log << "\nError: " << msg << std::endl;
        if (!extra.empty()) log << extra << std::endl;
        log.flush();
    }
    exit(1);
}

void warning(std::ofstream &log, const std::string &msg, const std::string &extra)
{
    std::cout << "\n" << msg << std::endl;
    if (!extra.empty()) std::cout << extra << std::endl;
    if (log) {
        log << "\n" << msg << std::endl;
        if (!extra.empty()) log << extra << std::endl;
    }
}

// NC_FreqCurveInterpolation (src/NC_EquationSystem.cpp:1963-2009)
double curve_factor(std::ofstream &log, const Project &P, int id, double f)
{
    for (const Curve &c : P.curves) {
        if (c.id != id) continue;
        if (f <= c.f[0]) return c.fac[0];
        for (size_t l = 1; l < c.f.size(); ++l)
            if (f < c.f[l]) return c.fac[l - 1] + (c.fac[l] - c.fac[l - 1]) / (c.f[l] - c.f[l - 1]) * (f - c.f[l - 1]);
        return c.fac.back();
    }
    fatal(log, "A curve can not be found!", "Number of the curve = " + str(id));
}

void read_project(std::ofstream &log, Project &P, bool check_normals)
{
    FILE *in = fopen("NC.inp", "r");
    if (!in) fatal(log, "No input file found!");
    std::vector<std::string> tok;
    log << "\n\n>> G E N E R A L   P A R A M E T E R S <<\n" << std::endl;
    std::cout << "3D analysis" << std::endl;
    log << "\n3D analysis" << std::endl;
    // NC_ReadBasicParametersA (src/NC_Input.cpp:151-210)
    int n = next(in, tok);
    if (tok[0] != "Mesh2HRTF") fatal(log, "Key word Mesh2HRTF expected!");
    P.version = tok[0];
    for (int i = 1; i < n; ++i) P.version += " " + tok[i];
    n = next(in, tok);
    P.title = tok[0];
    for (int i = 1; i < n; ++i) P.title += " " + tok[i];
    std::cout << "Job title: " << P.title << std::endl;
    next(in, tok);   // control parameter I
    next(in, tok);   // control parameter II
    P.nfreq = to_int(tok[1]);
    P.freq_inc = to_double(tok[2]);
    P.freq_first = to_double(tok[3]);
    // NC_ReadBasicParametersB (src/NC_Input.cpp:213-347)
    next(in, tok);
    const int npc = to_int(tok[1]);
    std::vector<double> ct(npc), cf(npc);
    for (int i = 0; i < npc; ++i) {
        next(in, tok);
        ct[i] = to_double(tok[0]);
        cf[i] = to_double(tok[1]);
    }
    if (npc > 0) compute_frequencies(P, ct, cf);
    else P.freqs.assign(P.nfreq, 0.0);
    next(in, tok);
    P.ngroups = to_int(tok[0]);
    P.nelem = to_int(tok[1]);
    P.nnode = to_int(tok[2]);
    P.min_terms = to_int(tok[3]);
    P.nsym = to_int(tok[4]);
    P.method_bem = to_int(tok[7]);
    P.method_solver = to_int(tok[8]);
    if (P.nelem < 1) fatal(log, "Number of elements must > 0!", "Number of elements = " + str(P.nelem));
    if (P.nnode < 1) fatal(log, "Number of nodes must > 0!", "Number of nodes = " + str(P.nnode));
    if (P.min_terms == 0) P.min_terms = 8;
    if (P.min_terms < 5 || P.min_terms > 10)
        fatal(log, "minExpansionTermsFMM_  must be in [5 10]!", "minExpansionTermsFMM_ = " + str(P.min_terms));
    if (P.nsym < 0 || P.nsym > 3)
        fatal(log, "Number of symmetry planes must be in [0 3]!", "Number of symmetry planes = " + str(P.nsym));
    if (P.method_solver < 0 || P.method_solver > 4 || (P.method_solver > 0 && P.method_solver < 4))
        fatal(log, "ISOLVER must be 0 or 4!", "ISOLVER = " + str(P.method_solver));
    if (P.method_bem != 0 && P.method_bem != 1 && P.method_bem != 4)
        fatal(log, "methodBEM_ must be in {0,1,4}!", "methodBEM_ = " + str(P.method_bem));
    if (P.method_bem && P.method_solver == 4) P.method_solver = 0;
    next(in, tok);
    P.npw = to_int(tok[0]);
    P.nps = to_int(tok[1]);
    P.interior = to_int(tok[2]);
    P.clus_edge = to_double(tok[3]);
    P.precond = to_int(tok[4]);
    if (P.npw < 0) fatal(log, "Number of the incident plane waves must >= 0!");
    if (P.nps < -1) fatal(log, "Number of the point sources must >= -1!");
    // -1 ("test point source", src/NC_Main.cpp:342-343) reads no point source; the count
    // still enters NumCalc's tests for incident waves (see the driver)
    if (P.interior != 0 && P.interior != 1) fatal(log, "IFINTE must = 0 or 1!", "IFINTE = " + str(P.interior));
    if (P.precond < -3 || P.precond > 2) fatal(log, "IPRECOND must be in [-3 2]!", "IPRECOND = " + str(P.precond));
    if (P.precond <= -3) { P.scan_degree = 3; P.precond = 0; }
    else if (P.precond < 0) { P.scan_degree = -P.precond; P.precond = 0; }
    if (P.method_solver == 4) P.precond = 2;
    next(in, tok);   // main parameters III
    next(in, tok);   // main parameters IV
    P.c = to_double(tok[0]);
    P.rho = to_double(tok[1]);
    P.h = to_double(tok[2]);
    if (P.h != -1.0) P.h = 1.0;
    if (P.c < 300.0) warning(log, "speedOfSound_ perhaps too small!");
    if (P.rho < 1.0) warning(log, "densityOfMedium_ perhaps too small!");
    // NC_Read (src/NC_Input.cpp:45-148)
    read_mesh_files(log, in, P, check_normals);
    next(in, tok);
    if (tok[0] != "BOUNDARY") fatal(log, "Key word BOUNDARY expected!");
    std::vector<BcLine> lines;
    const int maxl = 2 * P.n_be;
    bool retu = false;
    for (int i = 0; i < maxl; ++i) {
        const int nt = next(in, tok);
        if (tok[0] == "RETU") { retu = true; break; }
        if (nt < 9) fatal(log, "A input line under BOUNDARY is too short!", "Number of the input line = " + str(i));
        if (tok[0] != "ELEM") fatal(log, "Key word ELEM expected!");
        if (tok[2] != "TO") fatal(log, "Key word TO expected!");
        if ((tok[6] == "-1") != (tok[8] == "-1"))
            fatal(log, "ERROR: Curves have to be specified for all parts of a Boundary or Source. I.e., ALL curve numbers "
                       "in one line must be -1 or ALL curve numbers must not be -1");
        BcLine b;
        b.lo = to_int(tok[1]);
        b.hi = to_int(tok[3]);
        b.key = tok[4];
        b.re = to_double(tok[5]);
        b.im = to_double(tok[7]);
        b.ref = tok[6] != "-1";
        b.cre = b.ref ? to_int(tok[6]) : -1;
        b.cim = b.ref ? to_int(tok[8]) : -1;
        lines.push_back(b);
    }
    if (!retu) fatal(log, "Key word RETU expected!");
    // NC_ReadSoundSources (src/NC_Input.cpp:1002-1070)
    auto read_sources = [&](int count, const char *key, std::vector<Source> &dst, bool normalize) {
        next(in, tok);
        if (tok[0] != key)
            fatal(log, std::string("Key word ") + (normalize ? "PLANE WAVE" : "POINT SOURCES") + " expected!");
        for (int i = 0; i < count; ++i) {
            const int nt = next(in, tok);
            if (nt != 8)
                fatal(log, std::string("Number of terms of an input line for ") + (normalize ? "plane wave" : "pont source") +
                               " must be 8!",
                      "Number of terms = " + str(nt));
            Source s;
            s.id = to_int(tok[0]);
            for (int d = 0; d < 3; ++d) s.xyz[d] = to_double(tok[1 + d]);
            if (normalize) {
                const double L = sqrt(s.xyz[0] * s.xyz[0] + s.xyz[1] * s.xyz[1] + s.xyz[2] * s.xyz[2]);
                for (int d = 0; d < 3; ++d) s.xyz[d] /= L;
            }
            s.amp0 = cplx(to_double(tok[4]), to_double(tok[6]));
            s.curve[0] = to_int(tok[5]);
            s.curve[1] = to_int(tok[7]);
            dst.push_back(s);
        }
    };
    if (P.npw > 0) read_sources(P.npw, "PLANE", P.pw, true);
    if (P.nps > 0) read_sources(P.nps, "POINT", P.ps, false);
    // NC_ReadFrequencyCurves (src/NC_Input.cpp:1073-1160)
    next(in, tok);
    bool curves = false;
    if (tok[0] == "CURVES") curves = true;
    else if (tok[0] != "POST") fatal(log, "Key word CURVES or POST PROCESS expected!");
    if (curves) {
        const int nt = next(in, tok);
        if (nt != 2) fatal(log, "Number of terms of the input line direct under the key word CURVE must be 2!");
        const int ncv = std::max(0, to_int(tok[0]));
        for (int c = 0; c < ncv; ++c) {
            if (next(in, tok) != 2) fatal(log, "Number of terms of the firs input line for a curve must be 2!");
            Curve cv;
            cv.id = to_int(tok[0]);
            const int np = to_int(tok[1]);
            for (int k = 0; k < np; ++k) {
                next(in, tok);
                cv.f.push_back(to_double(tok[0]));
                cv.fac.push_back(to_double(tok[1]));
            }
            P.curves.push_back(cv);
        }
        next(in, tok);
        if (tok[0] != "POST") fatal(log, "Key word POST PROCESS expected!");
    }
    for (;;) {
        const int nt = get_line(in, tok);
        if (nt == 1 && tok[0] == "END") break;
    }
    fclose(in);
    store_bc(log, P, lines);
    write_parameters(log, P);
}

// NC_WriteParameters (src/NC_Input.cpp:1409-1593)
Symmetry::Symmetry(const Project &P)
{
    coord.assign(P.sym_coord, P.sym_coord + 3);
    type.assign(3, 0.0);
    for (int d = 0; d < 3; ++d)
        if (P.nsym > 0 && P.sym_planes[d] != 0) {
            dims.push_back(d);
            type[d] = P.sym_planes[d];
        }
    nref = 1 << dims.size();
}

double Symmetry::sign(int r) const
{
    double s = 1.0;
    for (size_t b = 0; b < dims.size(); ++b)
        if ((r >> b & 1) && type[dims[b]] < 0) s = -s;
    return s;
}

void Symmetry::mirror(int r, double *x) const
{
    for (size_t b = 0; b < dims.size(); ++b)
        if (r >> b & 1) x[dims[b]] = 2.0 * coord[dims[b]] - x[dims[b]];
}

void unfold_mesh(const Project &P, const std::vector<int> &be, const Symmetry &S, std::vector<int> &nv,
                 std::vector<int> &kind, std::vector<int> &nodes, std::vector<int> &groups, std::vector<double> &xyz)
{
    const int n = (int)be.size(), N = n * S.nref;
    nv.assign(N, 0);
    kind.assign(N, 0);
    groups.assign(N, 0);
    nodes.assign(4 * (size_t)N, -1);
    xyz.assign(12 * (size_t)N, 0.0);
    for (int r = 0; r < S.nref; ++r)
        for (int q = 0; q < n; ++q) {
            const int e = be[q], Q = r * n + q, m = P.elem_nv[e];
            const bool odd = __builtin_popcount(r) & 1;
            nv[Q] = m;
            kind[Q] = P.ibval[e];
            groups[Q] = P.elem_group[e];
            for (int a = 0; a < m; ++a) {
                const int src = odd && (a < 2 || m == 4) ? (a ^ 1) : a;
                const int v = P.elem_conn[4 * (size_t)e + src];
                double x[3] = {P.node_xyz[3 * (size_t)v], P.node_xyz[3 * (size_t)v + 1], P.node_xyz[3 * (size_t)v + 2]};
                int moved = 0;
                for (size_t b = 0; b < S.dims.size(); ++b)
                    if ((r >> b & 1) && 2.0 * S.coord[S.dims[b]] - x[S.dims[b]] != x[S.dims[b]]) moved |= 1 << b;
                S.mirror(r, x);
                nodes[4 * (size_t)Q + a] = v + P.nnode * moved;
                for (int d = 0; d < 3; ++d) xyz[12 * (size_t)Q + 3 * a + d] = x[d];
            }
        }
}

void write_parameters(std::ofstream &log, const Project &Pc)
{
    Project &P = const_cast<Project &>(Pc);
    double dmax = 0, dmin = 1.0e20, dave = 0, deqi = 0;
    int k1 = 0;
    for (int e = 0; e < P.nelem; ++e) {
        if (P.elem_prop[e] == 2) continue;
        const int nv = P.elem_nv[e];
        if (P.area[e] < 0.0) fatal(log, "Element with negative area is found!");
        double ref = sqrt(P.area[e]);
        if (nv == 3) ref *= 1.51967;
        for (int a = 0; a < nv; ++a) {
            const double *x = &P.node_xyz[3 * (size_t)P.elem_conn[4 * (size_t)e + a]];
            const double *y = &P.node_xyz[3 * (size_t)P.elem_conn[4 * (size_t)e + (a + 1) % nv]];
            const double d = sqrt((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]) + (x[2] - y[2]) * (x[2] - y[2]));
            dmax = std::max(dmax, d);
            dmin = std::min(dmin, d);
            dave += d;
            deqi += nv == 3 ? d * 0.748 : d;
            if (d > ref * 10.0 && d < ref * 100.0)
                warning(log, "Element with very long side found!", "external elemen number = " + str(P.elem_ext[e]));
            if (d > ref * 100.0) fatal(log, "Element of very distorted form is found!", "external elemen number = " + str(P.elem_ext[e]));
            k1++;
        }
    }
    P.dmesh_ave = dave / k1;
    P.dmesh_min = dmin;
    P.dmesh_max = dmax;
    P.delta = dmin / 100.0;
    P.freq_max = P.c / (6.0 * (deqi / k1));
    int i;
    for (i = 0; i < P.nfreq; ++i)
        if (P.freqs[i] > P.freq_max) break;
    if (P.nfreq > i) {
        log << std::endl;
        warning(log, "There are some frequencies higher than the highest allowable frequency!",
                "Number of frequencies outside the range= " + str(P.nfreq - i));
    }
    log << P.version << std::endl << P.title << std::endl << std::endl;
    log << "Number of frequency steps ................... = " << P.nfreq << std::endl << std::endl;
    log << "First step .................................. = " << P.freq_first << std::endl;
    log << "Number of steps ............................. = " << P.nfreq << std::endl;
    log << "Step increment .............................. = " << P.freq_inc << std::endl;
    log << "Highest allowable frequency .................. = " << P.freq_max << std::endl;
    log << "\n\n\n>> B A S I C   P A R A M E T E R S   F O R   B E M <<\n" << std::endl;
    log << "Number of element groups .................... = " << P.ngroups << std::endl;
    log << "Total number of elements .................... = " << P.nelem << std::endl;
    log << "Number of boundary elements ................. = " << P.n_be << std::endl;
    log << "Number of evaluation elements ............... = " << P.n_eval_elem << std::endl;
    log << "Maximum edge length of boundary elements .... = " << dmax << std::endl;
    log << "Average edge length of boundary elements .... = " << P.dmesh_ave << std::endl;
    log << "Minimum edge length of boundary elements .... = " << dmin << std::endl << std::endl;
    log << "Total number of nodes ....................... = " << P.nnode << std::endl;
    log << "Number of nodes of the boundary elements .... = " << P.n_be_nodes << std::endl;
    log << "Number of nodes of the evaluation elements .. = " << P.n_eval_nodes << std::endl << std::endl;
    log << "Number of symmetric planes .................. = " << P.nsym << std::endl << std::endl;
    log << "Number of incident plane waves .............. = " << P.npw << std::endl;
    log << "Number of incident spharic waves ............ = " << P.nps << std::endl << std::endl;
    log << "Sound speed ................................. = " << P.c << std::endl;
    log << "Density of the medium ....................... = " << P.rho << std::endl << std::endl;
    log << "Diameter of the BE-mesh ..................... = " << P.diameter << std::endl;
    if (P.method_bem > 0) log << std::endl << "Minimum number of terms in the FMM expansion. = " << P.min_terms << std::endl;
    log << std::endl << (P.interior ? "Interior domain problem" : "Exterior domain problem") << std::endl << std::endl;
    log << (P.h == 1.0 ? "Harmonic time factor = EXP(-i*omega*t)" : "Harmonic time factor = EXP(i*omega*t)") << std::endl
        << std::endl;
    if (P.method_solver == 0) log << "methodSolver_  = 0: CGS iterative solver" << std::endl;
    else log << "methodSolver_  = 4: Gaussean elimination (direct) solver" << std::endl;
    if (P.method_solver < 4) {
        if (P.precond == 0) log << "methodPreconditioner_ = 0: incomplete LU-decomposition preconditioner" << std::endl;
        else if (P.precond == 1) log << "methodPreconditioner_ = 1: scaling preconditioner" << std::endl;
        else log << "methodPreconditioner_ = 2: no preconditioner" << std::endl;
    }
    log << std::endl << " No.   Freq.   Lambda/L_BE_Mesh  Lambda/L_Element" << std::endl;
    const double ell = sqrt(P.avg_area);
    for (int q = 0; q < P.nfreq; ++q) {
        const double wl = P.c / P.freqs[q];
        log << std::setw(3) << q + 1 << " " << std::setw(7) << P.freqs[q] << "   " << std::setw(10) << wl / P.diameter
            << "        " << std::setw(10) << wl / ell << std::endl;
    }
    log << std::endl;
}

}  // namespace nc
