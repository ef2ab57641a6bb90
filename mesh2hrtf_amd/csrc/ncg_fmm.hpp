// Fast-multipole plan of libncg: NumCalc's cluster tree, near/far lists, method choice
// per frequency and expansion lengths, restated on the host (they are integer and
// geometric decisions taken once per mesh and per frequency; the FMM operators
// themselves are built and applied on the device).
//
// Reference: mesh2hrtf/NumCalc/src/NC_Addresses.cpp (clustering, lists, CSR layouts),
// src/NC_Main.cpp:348-489 (expansion lengths), src/NC_3dFunctions.cpp:786-818 (sphere).
#pragma once

#include <string>
#include <vector>

namespace ncgf {

struct Cluster {
    std::vector<int> elems;     // element indices, NumCalc's order
    double cen[3] = {0, 0, 0};
    double radius = 0.0;
    int father = -1;
    std::vector<int> near;      // clusters of this level not in the far field (incl. itself), ascending
    std::vector<int> fan;       // interaction list (far here, fathers near), ascending
};

struct Level {
    std::vector<Cluster> cl;
    double rmax = 0.0, rmin = 0.0, ravg = 0.0;
    int L = 0, ntheta = 0, nphi = 0, P = 0;
    std::vector<double> u;      // [P][3] unit vectors
    std::vector<double> w;      // [P] weights
    long long n_canceled = 0, n_far_geom = 0;
    // digest of the cancelled pairs: with the tree fixed, the near and interaction lists
    // are a function of this set
    unsigned long long cancel_hash = 0;
};

struct EvalCluster {
    std::vector<int> pts;
    double cen[3] = {0, 0, 0};
    double radius = 0.0;
};

// Cluster tree of one (method, number of levels): clusters with their elements, centres,
// radii and fathers, before the frequency-dependent near/far split, and the evaluation
// clusters once computed (they depend on the cluster edge only).
struct Tree {
    int method = -1, nlev = 0;
    std::vector<Level> lev;
    // the levels with their near and interaction lists when no pair is cancelled (min_dist
    // below every level's shortest geometrically far centre distance, dfar): the lists of
    // every such frequency, built once
    bool geo_ready = false;
    std::vector<Level> geo;
    std::vector<double> dfar;
    bool ev_ready = false;
    double ev_edge = 0.0, ev_rmax = 0.0, ev_rmin = 0.0;
    std::vector<EvalCluster> ev;
};

// Frequency-independent mesh data the plan needs (boundary elements only).
struct Mesh {
    int n = 0;
    std::vector<int> nv;             // 3 or 4
    std::vector<int> node;           // [n][4] node ids (dedupe key)
    std::vector<double> X;           // [n][4][3]
    std::vector<int> group;          // element group of each element
    std::vector<int> group_order;    // BE groups in NumCalc's group order
    std::vector<double> cen;         // [n][3] centres (NumCalc's shape-function sum)
    std::vector<double> area;        // [n]
    double avg_area = 0.0, delta = 0.0, diameter = 0.0;
    std::vector<double> eval;        // [m][3]
    int m = 0;
    mutable std::vector<Tree> trees; // trees built so far (they depend on the mesh only)
};

struct Options {
    int method_bem = 0;          // 0 TBEM, 1 SL-FMM, 4 ML-FMM (NC.inp Main Parameters I)
    int min_terms = 8;           // minimum expansion length (5..10)
    double clus_edge = 0.0;      // ClusEdgL0 (<= 0: computed)
};

struct Plan {
    int method = 0;              // 0 TBEM, 1 SL, 3 ML (NumCalc's methodFMM_)
    std::vector<Level> lev;      // 1 level for SL, numClusterLevels for ML
    std::vector<EvalCluster> ev; // evaluation-point clusters
    double ev_rmax = 0.0, ev_rmin = 0.0;
    double clus_edge = 0.0;      // ClusEdgL1 in force
    double min_dist = 0.0;       // minClusterDistance = H_Argum_Min / k
    std::vector<double> cancel_ratio;
    std::string error;           // NumCalc's fatal message when the plan cannot be made
    std::string warning;
};

void prepare_mesh(Mesh &M);
// Plan for wave number k; returns false with P.error set on NumCalc's fatal conditions.
bool make_plan(const Mesh &M, const Options &O, double k, Plan &P);
// Gauss-Legendre rule of n points (n <= 80), ascending
void gauss_legendre(int n, double *x, double *w);

}  // namespace ncgf
