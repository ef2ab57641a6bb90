// Internal C++ seams of libncg shared by the C ABI (ncg_api.cpp) and the serving
// runtime (ncg_server.cpp); not part of the public interface.
#pragma once

#include <stddef.h>

#include <string>

#include "ncg.h"
#include "ncg_fmm.hpp"

namespace ncgi {

// records the message ncg_last_error() returns on this thread and passes the code on
int fail(int code, const std::string &msg);

// ncg_assemble without the wait: the kernels are queued on the problem's stream and the
// assembly's error flags are read by assemble_wait (or the next solve / apply)
int assemble_async(ncg_problem *p, const ncg_frequency *f, int precision);
int assemble_wait(ncg_problem *p);
// ncg_field in two halves: queue the field kernels, then copy the results out (waits)
int field_begin(ncg_problem *p, bool want_grad);
int field_end(ncg_problem *p, double *phi, double *grad);

// planner mesh and options of an ABI geometry
void planner_mesh(const ncg_geometry *g, ncgf::Mesh &M, ncgf::Options &O);

// Device bytes a problem of this mesh holds after a step at wave number k: geometry,
// the dense or FMM operator with its grow-only slack, Krylov and evaluation-grid
// workspace.  Returns 0 when NumCalc's plan fails at this k.
size_t estimate_bytes(const ncgf::Mesh &M, const ncgf::Options &O, int m_eval, double k, int precision, int nrhs,
                      int restart);

}  // namespace ncgi
