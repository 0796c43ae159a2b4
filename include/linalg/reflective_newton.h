/* reflective_newton.h of the liblinalg_hip drop-in
 * (LinA/reflective_newton.h:41-79): the box-constrained trust-region
 * reflective Newton method on q(x) = x.H.x/2 + x.b (the reference's default
 * functional and gradient), run on the GPU by liblina_hip.  Inner solvers:
 * RN_PCG (ILU(0)-PCG), RN_BICGSTAB (BiCGStab + AMG), RN_AMG (AMG_solve); the
 * RN_UMF slot takes the library's GMRES(30) + ILU(0) (RN_GMRES).  User
 * functionals, gradients, Hessian callbacks and local solvers are not run
 * on the device: reflective_newton returns -1 for them. */
#ifndef LINALG_HIP_REFLECTIVE_NEWTON_H
#define LINALG_HIP_REFLECTIVE_NEWTON_H

#include "BiCGStab.h"
#include "Multigrid.h"
#include "PCGS.h"
#include "linear_algebra.h"

#ifdef __cplusplus
extern "C" {
#endif

#define POSITIVE_DEFINITE 1
#define INDEFINITE 0
#define RN_NO 0
#define RN_YES 1
#define RN_FAILED 0
#define RN_BICGSTAB 1
#define RN_AMG 2
#define RN_PCG 3
#define RN_CHOLESKY 4
#define RN_UMF 5
#define RN_GMRES 6

typedef struct RN_STATISTICS {
    int calls;
    int average_eig_iter;
    int number_neg_curv;
    int line_search_count;
    int line_search_iter;
    int total_iter;
    double res;
} RN_statistics;

void RN_set_params(int R_Newton_max_iter, double R_Newton_eps, double R_Newton_sigma_l, double R_Newton_sigma_u,
                   double R_Newton_rho, double R_Newton_Delta, double R_Newton_Xi, double R_Newton_trust_lower,
                   double R_Newton_trust_upper, double R_Newton_trust_radius_min, double R_Newton_trust_radius_max);
void set_RN_params(int max_iter, double eps, double sigma_l, double sigma_u);
void set_RN_constraints(int* Index_set, double* Lower, double* Upper, int n);
void set_RN_Hesse(sparse_matrix* A);
void set_RN_linear(double* B);
void set_RN_functional(double (*F)(double* X, int n));
void RN_set_gradient(double* (*Gradient_func)(double* X));
void RN_set_Hesse_function(sparse_matrix* (*Hesse_func)(double* X), int dim_Hesse);
void RN_set_default_matrix(sparse_matrix* M);
void RN_default_mat_mult(int n, double* y, double* z);
void RN_set_AMG_data(amg_system_info* AMG_info);
void RN_set_solver(int type);
void RN_set_local_solver(int (*Solver)(sparse_matrix* A, double* X, double* b));
void RN_set_trust_region(double R_Newton_Delta);
void RN_set_trust_limits(double min, double max);
void RN_print_info(int mode);
int RN_info_set(void);
void RN_set_debug_mode(int mode);
void RN_set_kernel(double* K, int n);
void RN_init_statistics(void);
void print_RN_statistics(FILE* file);
int reflective_newton(double* X);
/* reflective_newton.c:394-450: X = c V1 + s V2, c = Delta cos t,
 * s = Delta sin t, for the t of least q = X.A.X + X.b on the circle of the
 * current trust radius Delta (RN_set_trust_region, RN_set_params), found
 * among the stationary points of q(t).  When there is none the reference
 * exits; here X is filled with NaN and lina_last_error() says why. */
void subspace_minimum(sparse_matrix* A, double* b, double* V1, double* V2, double* X);
/* reflective_newton.c:64-69: the list 0, 1, ..., size-1 */
int* get_all_indices(int size);
/* :376-392: n samples of F over [min, max] into the file "func" */
void RN_print_function(double (*F)(double x), int n, double min, double max);
/* :719-757: random search on set_RN_functional's functional within the
 * bounds of set_RN_constraints; -1 unless every component is bounded */
int simulated_annealing_min(double* X, int n);

#ifdef __cplusplus
}
#endif
#endif
