/* Internal glue of the LinA drop-in (C host code over liblina_hip's C ABI). */
#ifndef LINALG_HIP_DROPIN_INTERNAL_H
#define LINALG_HIP_DROPIN_INTERNAL_H

#include "../../include/lina_hip.h"
#include "../../include/linalg/linear_algebra.h"

/* CSR copy of a row-list matrix */
typedef struct {
    int n;
    int* rp;
    int* ci;
    double* v;
} dropin_csr;
dropin_csr dropin_to_csr(sparse_matrix* A);
void dropin_free_csr(dropin_csr* c);
lina_matrix* dropin_device_matrix(sparse_matrix* A);

/* an operator y = A x: a device matrix, or a host callback staged through
 * host buffers */
typedef struct {
    lina_matrix* dev;
    void (*host)(int n, double* x, double* y);
} dropin_op;
/* a preconditioner z = M^-1 r (z is also the initial guess): AMG, ILU, or a
 * host callback; returns the callback's count */
typedef struct {
    lina_solver* amg;
    lina_ilu* ilu;
    int (*host)(int n, double* r, double* z);
} dropin_prec;

typedef struct {
    int n;
    double* hx;
    double* hy;
} dropin_stage;
void dropin_stage_init(dropin_stage* s, int n);
void dropin_stage_free(dropin_stage* s);
void dropin_op_apply(const dropin_op* o, dropin_stage* s, const double* dx, double* dy);
int dropin_prec_apply(const dropin_prec* p, dropin_stage* s, const double* dr, double* dz);

double* dropin_dalloc(int n);       /* device vector, zeroed */
void dropin_dfree(double* d);
void dropin_h2d(double* d, const double* h, int n);
void dropin_d2h(double* h, const double* d, int n);
void dropin_d2d(double* dst, const double* src, int n);
double dropin_dot(const double* a, const double* b, int n);
void dropin_check(int rc, const char* what);

/* the AMG settings of Multigrid.c's globals for the next solve */
void dropin_amg_sync(void* amg_info);
lina_matrix* dropin_pcg_default_matrix(void);
lina_ilu* dropin_pcg_ilu(char which); /* 'B' both sweeps, 'L' forward, 'U' backward */
lina_matrix* dropin_rn_default_matrix(void);

#endif
