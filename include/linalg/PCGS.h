/* PCGS.h of the liblinalg_hip drop-in: LinA's preconditioned CG with its
 * global configuration (LinA/PCGS.h:13-22), run on the GPU.
 *
 * pcg keeps the reference's semantics (PCGS.c:163-233): B and X scaled by
 * 1/|B| in place and back, stop on |R|/|M^-1 R_0|, the preconditioner called
 * on the previous Z (warm start), get_pcg_residuum = |M^-1 R_end|/|M^-1 R_0|.
 * The vectors live in HBM; the operator and preconditioner run on the device
 * when they are the library's own (PCG_mult with PCG_default's matrix,
 * PCG_preconlr with PCG_default's factors, PCG_AMG_precon), and through
 * host staging buffers when a caller registers its own callbacks. */
#ifndef LINALG_HIP_PCGS_H
#define LINALG_HIP_PCGS_H

#include "linear_algebra.h"

#ifdef __cplusplus
extern "C" {
#endif

int pcg(double* X, double* B, int n, double tol, int max_iter);
double get_pcg_residuum(void);

void PCG_set_precon(int (*Precon)(int n, double* R, double* Z));
void PCG_set_mat_mult(void (*Mat_mult)(int n, double* x, double* y));
void* PCG_get_precon(void);
void* PCG_get_mat_mult(void);
void PCG_set_default_precon(void);
void PCG_default(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U);
void PCG_set_AMG_preconditioner(sparse_matrix* A, int num_fields, int depth);

/* the library's own operator / preconditioners (PCGS.c:36-116), usable as
 * callbacks and callable directly on host arrays */
void PCG_mult(int n, double* y, double* z);
int PCG_preconlr(int n, double* z, double* w);
int PCG_AMG_precon(int n, double* Y, double* X);
/* the two sweeps of PCG_preconlr on their own (PCGS.c:61-100): w = L^-1 z
 * with PCG_default's L (diagonal last in each row), w = U^-1 z with its U
 * (diagonal first) */
void PCG_preconl(int n, double* z, double* w);
void PCG_preconr(int n, double* z, double* w);

#ifdef __cplusplus
}
#endif
#endif
