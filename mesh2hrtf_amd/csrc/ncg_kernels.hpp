// Host-side launchers of the libncg kernels (implemented in ncg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

namespace ncgk {

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

void assemble_far(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, void *A, int mixed,
                  int rows_per_block, unsigned long long *evals);
void assemble_near(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const int *rows,
                   const int *cols, long long npairs, void *A, int mixed, int *err);
void rhs_sources(hipStream_t s, const DevElements &D, const BMParams &b, const int *src, int nsrc, const void *bcval,
                 int nrhs, void *rhs_src, int *err);
void assemble_diag(hipStream_t s, const DevElements &D, const BMParams &b, const double *adm, const void *bcval,
                   int nrhs, const double *ps, int nps, const double *pw, int npw, const void *rhs_src, int has_src,
                   void *A, int mixed, void *rhs);

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
void multidot(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *w, void *h);
void multiaxpy(hipStream_t s, int n, int nrhs, int kcount, const void *V, const void *h, void *w, int add);
void lu_factor(hipStream_t s, void *A, int n);
void lu_solve(hipStream_t s, const void *A, int n, int nrhs, void *b);
void field_prep(hipStream_t s, int n, int nrhs, const int8_t *kind, const double *adm, double tau, const void *bcval,
                const void *x, void *pot, void *vel);
void field(hipStream_t s, const DevElements &D, const BMParams &b, const double *pts, int m, int nrhs, const void *pot,
           const void *vel, const double *ps, int nps, const double *pw, int npw, void *phi, void *grad, int fp32,
           int *err);

}  // namespace ncgk
