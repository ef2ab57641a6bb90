// Host side of the drop-in NumCalc: the project data read from NC.inp and the mesh
// files, and the helpers shared by the reader, the writers and the driver.
// Reference: mesh2hrtf/NumCalc/src/NC_Input.cpp (reader), NC_PostProcessing.cpp
// (be.out writers), NC_Main.cpp (driver and log).
#pragma once

#include <complex>
#include <fstream>
#include <string>
#include <vector>

namespace nc {

using cplx = std::complex<double>;

struct Curve {
    int id = 0;
    std::vector<double> f, fac;
};

struct Source {            // plane wave (direction) or point source (position)
    int id = 0;
    double xyz[3] = {0, 0, 0};
    cplx amp0;             // as read (before per-frequency scaling)
    int curve[2] = {-1, -1};
};

struct Project {
    std::string version, title;
    int nfreq = 0;
    double freq_inc = 0.0, freq_first = 0.0;
    std::vector<double> freqs;
    int ngroups = 0, nelem = 0, nnode = 0, min_terms = 8, nsym = 0, method_bem = 0, method_solver = 0;
    int npw = 0, nps = 0, interior = 0, precond = 0, scan_degree = 0;
    double clus_edge = 0.0, c = 343.0, rho = 1.21, h = 1.0;
    // nodes and elements (internal order = file order)
    std::vector<int> node_ext;
    std::vector<double> node_xyz;       // [nnode][3]
    std::vector<int> elem_ext, elem_nv, elem_prop, elem_group;
    std::vector<int> elem_conn;         // [nelem][4] internal node indices
    int sym_planes[3] = {0, 0, 0};
    double sym_coord[3] = {0, 0, 0};
    // boundary conditions per element (NC_StoreBoundaryConditions)
    std::vector<int> ibval;             // 0 velo, 1 pres, 2 velo+adm
    std::vector<cplx> bc0;              // prescribed value (all nodes alike)
    std::vector<cplx> adm1;             // admittance
    std::vector<int> curve_of;          // [nelem][6] curve ids, -1 none
    std::vector<Source> pw, ps;
    std::vector<Curve> curves;
    // derived
    int n_be = 0, n_eval_elem = 0, n_be_nodes = 0, n_eval_nodes = 0;
    std::vector<int> node_class;        // >=0 BE node, -1 evaluation node
    std::vector<int> group_index, group_prop, group_count;
    double avg_area = 0.0, diameter = 0.0, delta = 0.0;
    double dmesh_min = 0.0, dmesh_max = 0.0, dmesh_ave = 0.0, freq_max = 0.0;
    std::vector<double> area;           // per element
};

// Fatal input/consistency error: the message goes to stdout and NC.out, then exit(1)
// (NC_Error_Exit_*, src/NC_CommonFunctions.cpp:48-335).
[[noreturn]] void fatal(std::ofstream &log, const std::string &msg, const std::string &extra = "");
void warning(std::ofstream &log, const std::string &msg, const std::string &extra = "");

// NC.inp + mesh files from the current directory (NC_ReadBasicParametersA/B, NC_Read,
// NC_ReadMesh, NC_ReadBoundaryConditions, NC_ReadSoundSources, NC_ReadFrequencyCurves,
// NC_StoreBoundaryConditions, NC_WriteParameters).
void read_project(std::ofstream &log, Project &P, bool check_normals);
void write_parameters(std::ofstream &log, const Project &P);

// Symmetry planes (src/NC_Input.cpp:459-495, NC_ReflectElement src/NC_3dFunctions.cpp:357-452).
// NumCalc adds, for every element, the integrals over its mirror images with the factor -1
// per antisymmetric plane.  This build unfolds the mesh instead: image r (bit b set =
// mirrored in the b-th plane) of boundary element q becomes unknown r n + q.  The unfolded
// problem with mirrored boundary values and sources has the symmetric solution
// x[r n + q] = sign(r) x[q]; its first n unknowns are NumCalc's.
struct Symmetry {
    std::vector<int> dims;   // axes that carry a plane, ascending
    int nref = 1;            // 2^planes images, the identity first
    std::vector<double> coord, type;
    explicit Symmetry(const Project &P);
    double sign(int r) const;                  // -1 per antisymmetric plane mirrored in r
    void mirror(int r, double *x) const;       // point x -> its image r
};

// element arrays of the ABI geometry for the boundary elements be[] and their images:
// nodes mirrored, nodes 0<->1 (and 2<->3 of a quad) swapped after an odd number of
// mirrors so the normal keeps pointing outwards, as NumCalc reorders them; a node on a
// plane keeps its id across that plane
void unfold_mesh(const Project &P, const std::vector<int> &be, const Symmetry &S, std::vector<int> &nv,
                 std::vector<int> &kind, std::vector<int> &nodes, std::vector<int> &groups, std::vector<double> &xyz);

// per-frequency scaling of NC_UpdateFreqCurves (src/NC_EquationSystem.cpp:1815-2009)
double curve_factor(std::ofstream &log, const Project &P, int curve_id, double freq);

// be.out/be.<step>/{pBoundary,vBoundary,pEvalGrid,vEvalGrid}
void write_boundary(std::ofstream &log, const Project &P, int step, const std::vector<cplx> &pot,
                    const std::vector<cplx> &vel, double rpfact);
void write_evalgrid(std::ofstream &log, const Project &P, int step, const std::vector<cplx> &phi,
                    const std::vector<cplx> &grad, double rpfact);

}  // namespace nc
