// be.out writers of the drop-in NumCalc; formats of NC_WriteResultsObjectMesh and
// NC_WriteResultsEvaluationGrid (mesh2hrtf/NumCalc/src/NC_PostProcessing.cpp:447-546, 688-749).
#include <stdio.h>
#include <sys/stat.h>

#include "nc_host.h"

namespace nc {

static FILE *open_out(std::ofstream &log, int step, const char *name)
{
    char path[256];
    snprintf(path, sizeof(path), "be.out/be.%d", step);
    mkdir(path, S_IRWXU | S_IRGRP | S_IWGRP | S_IXGRP | S_IROTH | S_IXOTH);
    snprintf(path, sizeof(path), "be.out/be.%d/%s", step, name);
    FILE *f = fopen(path, "w");
    if (!f) fatal(log, std::string("Can not open the file ") + path);
    return f;
}

// pot/vel: per element (BE elements only meaningful), p = i * rpfact * pot
void write_boundary(std::ofstream &log, const Project &P, int step, const std::vector<cplx> &pot,
                    const std::vector<cplx> &vel, double rpfact)
{
    FILE *fp = open_out(log, step, "pBoundary");
    FILE *fv = open_out(log, step, "vBoundary");
    fprintf(fp, "%s\n", P.version.c_str());
    fprintf(fv, "%s\n", P.version.c_str());
    fprintf(fp, "%5d\n", P.ngroups);
    fprintf(fv, "%5d\n", P.ngroups);
    for (int g = 0; g < P.ngroups; ++g) {
        bool has_be = false;
        for (int e = 0; e < P.nelem && !has_be; ++e)
            if (P.elem_prop[e] != 2 && P.elem_group[e] == P.group_index[g]) has_be = true;
        if (!has_be) continue;
        fprintf(fp, "%5d %5d\n", P.group_index[g], P.group_count[g]);
        fprintf(fv, "%5d %5d\n", P.group_index[g], P.group_count[g]);
        for (int e = 0; e < P.nelem; ++e) {
            if (P.elem_prop[e] == 2 || P.elem_group[e] != P.group_index[g]) continue;
            const cplx p = cplx(0.0, rpfact) * pot[e];
            fprintf(fp, "%5d % E % E\n", P.elem_ext[e], p.real(), p.imag());
            fprintf(fv, "%5d % E % E\n", P.elem_ext[e], vel[e].real(), vel[e].imag());
        }
    }
    fclose(fp);
    fclose(fv);
}

// phi: per evaluation node, grad: [node][3]
void write_evalgrid(std::ofstream &log, const Project &P, int step, const std::vector<cplx> &phi,
                    const std::vector<cplx> &grad, double rpfact)
{
    FILE *fp = open_out(log, step, "pEvalGrid");
    FILE *fv = open_out(log, step, "vEvalGrid");
    fprintf(fp, "%s\n", P.version.c_str());
    fprintf(fv, "%s\n", P.version.c_str());
    int group = -1;
    for (int e = 0; e < P.nelem; ++e)
        if (P.elem_prop[e] == 2) { group = P.elem_group[e]; break; }
    const int ngrp = P.n_eval_nodes > 0 ? 1 : 0;
    fprintf(fv, "%5d\n", ngrp);
    if (ngrp) fprintf(fv, "%5d %5d\n", group, P.n_eval_nodes);
    int q = 0;
    for (int k = 0; k < P.nnode; ++k) {
        if (P.node_class[k] != -1) continue;
        const cplx *g = &grad[3 * (size_t)q];
        fprintf(fv, "%5d %E %E %E %E %E %E\n", P.node_ext[k], g[0].real(), g[0].imag(), g[1].real(), g[1].imag(),
                g[2].real(), g[2].imag());
        ++q;
    }
    fprintf(fp, "%5d\n", ngrp);
    if (ngrp) fprintf(fp, "%5d %5d\n", group, P.n_eval_nodes);
    q = 0;
    for (int k = 0; k < P.nnode; ++k) {
        if (P.node_class[k] != -1) continue;
        const cplx p = cplx(0.0, rpfact) * phi[q];
        fprintf(fp, "%5d % E % E\n", P.node_ext[k], p.real(), p.imag());
        ++q;
    }
    fclose(fp);
    fclose(fv);
}

}  // namespace nc
