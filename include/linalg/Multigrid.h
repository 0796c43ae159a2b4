/* Multigrid.h of the liblinalg_hip drop-in: LinA's AMG (LinA/Multigrid.h).
 * AMG_setup builds the reference's hierarchy bit for bit (on the host) and
 * keeps it in HBM; AMG_solve runs the V-cycles on the GPU.  The global
 * settings (AMG_set_*) apply to the next solve as in the reference. */
#ifndef LINALG_HIP_MULTIGRID_H
#define LINALG_HIP_MULTIGRID_H

#include "linear_algebra.h"

#ifdef __cplusplus
extern "C" {
#endif

#define NO 0
#define YES 1
#define UNDECIDED 0
#define COARSE 1
#define FINE 2
#define ISOLATED 3
#define DIRECT 1
#define STANDARD 2
#define AGGRESSIVE 1
#define STAND_ALONE 0
#define FULL_AMG 2

/* amg_system_info (Multigrid.h:45-70): the reference's fields first, in its
 * order, then the device handle.  AMG_setup fills every field as the
 * reference does -- Coarse_A, Prolongation, Restriction, Coarsening and the
 * diagonal Signs as row lists copied from the hierarchy the GPU holds, L/U
 * the ILU(0) factors of each level above the coarsest and the dense LU of
 * the coarsest -- and AMG_recompute_coarse_matrices refreshes Coarse_A, L
 * and U.  The solves read the device copy only. */
typedef struct AMG_SYSTEM_INFO {
    sparse_matrix** Coarse_A;
    sparse_matrix** Prolongation;
    sparse_matrix** Restriction;
    sparse_matrix** L;
    sparse_matrix** U;
    sparse_matrix* Signs;
    int** Coarsening;
    int* Fine_sizes;
    int* Coarse_sizes;
    int depth;
    int AMG_cycle_structure;
    int AMG_prolongation_method;
    int AMG_smoothing_iterations;
    int AMG_aggr_smoothing_iterations;
    int AMG_coarsening;
    int AMG_info;
    int AMG_deg_of_freedom;
    int (*F_Prolong)(sparse_matrix* A, sparse_matrix* B, int* Label, int i, int** Indices, double** Values,
                     int** Status, int** Strong, int** Positions);
    void (*AMG_coarsest_solver)(sparse_matrix* L, sparse_matrix* U, double* F, double* Sol, int size);
    int* (*Get_iter)(int depth);
    void* device;      /* lina_solver* of liblina_hip */
    void* settings;    /* its lina_amg_params, refreshed from the AMG_set_* globals before each solve */
    int main_size;
} amg_system_info;

/* amg_statistics (Multigrid.h:72-79) */
typedef struct AMG_STATISTICS {
    int calls;
    int max_iter;
    double max_res;
    int* Iterations;
    double* Residuals;
    char File_name[512];
} amg_statistics;

amg_system_info* AMG_setup(sparse_matrix* A, int AMG_deg_of_freedom, int depth, int AMG_smoothing_iterations,
                           int AMG_aggr_smoothing_iterations, int AMG_coarsening, int AMG_cycle_structure,
                           int AMG_prolongation_method,
                           void (*AMG_coarsest_solver)(sparse_matrix* L, sparse_matrix* U, double* F, double* Sol,
                                                       int size),
                           int* (*Get_iter)(int depth));
int AMG_solve(amg_system_info* Data, double* F, double* Fine_Sol, amg_statistics* Stat);
void AMG_recompute_coarse_matrices(amg_system_info* Info, sparse_matrix* New_matrix);
void AMG_free_data(amg_system_info* Data);
int* AMG_default_iterations(int depth);

/* smoothers selectable with AMG_set_smoother (identified by address) */
void AMG_SOR_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix, int iter,
                      int deg_freedom);
void AMG_parallel_Jacobi_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix,
                                  int iter, int deg_freedom);

void AMG_print_info(int s);
int AMG_info_set(void);
void AMG_set_SOR_relaxation_coeff(double alpha);
void set_connection_thresholds(double negative, double positive);
void AMG_set_eps(double eps);
void AMG_set_max_iter(int iter);
int AMG_get_max_iter(void);
void AMG_set_deg_of_freedom(int n);
void AMG_set_matrix_tolerance(double tol);
void AMG_set_smoother(void (*Smoother)(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix,
                                       int iter, int deg_freedom));

amg_statistics AMG_init_stat(const char* Out);
void AMG_add_stat(amg_statistics* Stat, int iter, double res);
void AMG_print_stat(amg_statistics* Stat);

#ifdef __cplusplus
}
#endif
#endif
