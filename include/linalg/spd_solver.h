/* spd_solver.h of the liblinalg_hip drop-in: GPU AMG-PCG solvers with the
 * signatures of Crack's UMFPACK drivers (fem_amd/host/spd_solver.c).  No
 * counterpart in LinA: they replace multifrontal_solver
 * (Crack/crack.c:5424-5525, the elastic step at :9340) and
 * multifrontal_solver_RN (:5664, RN_set_local_solver at :9207) on the SPD
 * systems of the staggered loop.  Rows whose only nonzero is the diagonal
 * (Set_dirichlet_BC's unit rows, :6445-6459) are constraints; their columns
 * are eliminated symmetrically before the solve.  X is the initial guess and
 * receives the solution. */
#ifndef LINALG_HIP_SPD_SOLVER_H
#define LINALG_HIP_SPD_SOLVER_H

#include "linear_algebra.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LINA_SPD_ELASTIC 0
#define LINA_SPD_LOCAL 1

/* multifrontal_solver(A, X, b, Caller, time): SUCCESS (1) when the relative
 * residual reached the tolerance, FAILED (0) otherwise; the wall time is
 * added to *time (may be NULL) */
int lina_spd_solver(sparse_matrix* A, double* X, double* b, char* Caller, double* time);
/* the RN_set_local_solver slot (multifrontal_solver_RN): same, own hierarchy */
int lina_spd_local_solver(sparse_matrix* A, double* X, double* b);
/* relative-residual target, iteration limit, AMG depth, unknowns per node
 * and LINA_PREC_FP64 / LINA_PREC_MIXED of one of the two solvers
 * (defaults: elastic 1e-12, 20000, 8, 2, mixed; local 1e-12, 20000, 6, 1,
 * FP64); drops its cached hierarchy */
void lina_spd_solver_settings(int which, double tol, int max_iter, int depth, int dof, int precision);
/* iterations and final relative residual of the last solve */
int lina_spd_last_iterations(int which, double* relres);
/* frees both cached hierarchies */
void lina_spd_solver_free(void);

#ifdef __cplusplus
}
#endif
#endif
