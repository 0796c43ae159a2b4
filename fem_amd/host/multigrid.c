/* AMG entry points of the LinA drop-in (LinA/Multigrid.c): the hierarchy is
 * built by liblina_hip (bit-identical to AMG_setup) and kept in HBM;
 * AMG_solve runs on the GPU with the settings of the global AMG_set_*
 * calls in effect at the time of the solve, as in the reference. */
#include <string.h>

#include "../../include/linalg/Multigrid.h"
#include "dropin.h"

/* the reference's global defaults (Multigrid.c:35-59) */
static double g_eps_minus = 0.3, g_eps_plus = 0.5, g_matrix_tol = 0.1;
static int g_max_iter = 20;
static double g_eps = 1e-8;
static double g_alpha = 1.0;
static int g_info = YES;
static int g_smoother = LINA_SMOOTHER_SOR;
static int g_dof = 1;

/* The two smoothers of the reference's V-cycle.  Passed to AMG_set_smoother
 * they select the GPU V-cycle's smoother (by address); called directly they
 * sweep the caller's host row list as the reference does: forward SOR over
 * the rows whose label differs from `fix` (Multigrid.c:1345-1365), or the
 * undamped Jacobi sweep over the first (n/dof)*dof rows in node blocks
 * (Multigrid.c:1367-1400; alpha is not used there). */
void AMG_SOR_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix, int iter,
                      int deg_freedom) {
    int k, i, j;
    (void)deg_freedom;
    for (k = 0; k < iter; k++)
        for (i = 0; i < A->size; i++) {
            double s = F[i], d = 0;
            int has = 0;
            if (Label[i] == fix) continue;
            for (j = 0; j < A->Len[i]; j++) {
                const int c = A->Indices[i][j];
                if (c != i) s -= A->Values[i][j] * Sol[c];
                else {
                    d = A->Values[i][j];
                    has = 1;
                }
            }
            if (!has) {
                fprintf(stderr, "AMG_SOR_smoother: row %d has no diagonal entry\n", i);
                return;
            }
            Sol[i] = (1. - alpha) * Sol[i] + alpha * s / d;
        }
}

void AMG_parallel_Jacobi_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix,
                                  int iter, int deg_freedom) {
    const int n = A->size, N = n / deg_freedom;
    double* old = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int k, i, d, j;
    (void)alpha;
    for (k = 0; k < iter; k++) {
        memcpy(old, Sol, (size_t)n * sizeof(double));
        for (i = 0; i < N; i++)
            for (d = 0; d < deg_freedom; d++) {
                const int r = i + N * d;
                double s = F[r], a = 0;
                if (Label[r] == fix) continue;
                for (j = 0; j < A->Len[r]; j++) {
                    const int c = A->Indices[r][j];
                    if (c != r) s -= A->Values[r][j] * old[c];
                    else a = A->Values[r][j];
                }
                Sol[r] = s / a;
            }
    }
    free(old);
}

void AMG_set_smoother(void (*Smoother)(sparse_matrix*, double*, double*, int*, double, int, int, int)) {
    if (Smoother == NULL || Smoother == AMG_SOR_smoother) g_smoother = LINA_SMOOTHER_SOR;
    else if (Smoother == AMG_parallel_Jacobi_smoother) g_smoother = LINA_SMOOTHER_JACOBI;
    else fprintf(stderr, "liblinalg_hip: a caller-defined AMG smoother is not run on the GPU; keeping SOR\n");
}

void AMG_print_info(int s) { g_info = s; }
int AMG_info_set(void) { return g_info; }
void AMG_set_SOR_relaxation_coeff(double alpha) { g_alpha = alpha; }
void set_connection_thresholds(double negative, double positive) {
    g_eps_minus = negative;
    g_eps_plus = positive;
}
void AMG_set_eps(double eps) { g_eps = eps; }
void AMG_set_max_iter(int iter) { g_max_iter = iter; }
int AMG_get_max_iter(void) { return g_max_iter; }
void AMG_set_deg_of_freedom(int n) { g_dof = n; }
void AMG_set_matrix_tolerance(double tol) { g_matrix_tol = tol; }

int* AMG_default_iterations(int depth) {
    int i;
    int* it = (int*)malloc((depth > 0 ? depth : 1) * sizeof(int));
    for (i = 0; i < depth; i++) it[i] = 1;
    return it;
}

/* ------------------------------------------- the level row lists */

static void drop(sparse_matrix** M) {
    if (*M) {
        Free_sparse(*M);
        free(*M);
        *M = NULL;
    }
}

static sparse_matrix* rowlist(int n, const int* rp, const int* ci, const double* v) {
    sparse_matrix* M = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    int i;
    M->size = n;
    M->Len = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    M->Indices = (int**)malloc((size_t)(n > 0 ? n : 1) * sizeof(int*));
    M->Values = (double**)malloc((size_t)(n > 0 ? n : 1) * sizeof(double*));
    for (i = 0; i < n; i++) {
        const int m = rp[i + 1] - rp[i];
        M->Len[i] = m;
        M->Indices[i] = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
        M->Values[i] = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
        memcpy(M->Indices[i], ci + rp[i], (size_t)m * sizeof(int));
        memcpy(M->Values[i], v + rp[i], (size_t)m * sizeof(double));
    }
    return M;
}

/* level l's A, P or R of the device hierarchy's host copy */
static sparse_matrix* level_matrix(lina_solver* s, int l, char which, int rows) {
    const int nnz = lina_amg_level_nnz(s, l, which);
    int* rp = (int*)malloc((size_t)(rows + 1) * sizeof(int));
    int* ci = (int*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    double* v = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    sparse_matrix* M = NULL;
    if (nnz >= 0 && lina_amg_level_copy(s, l, which, rp, ci, v) == 0) M = rowlist(rows, rp, ci, v);
    else dropin_check(-1, "AMG level copy");
    free(rp);
    free(ci);
    free(v);
    return M;
}

/* Coarse_A[l] of every level, then the factors AMG_setup keeps beside them
 * (Multigrid.c:1676-1680, 1741-1743): ILU(0) of every level above the
 * coarsest, the dense LU of the coarsest */
static void materialise_matrices(amg_system_info* I) {
    lina_solver* s = (lina_solver*)I->device;
    int l;
    for (l = 0; l < I->depth; l++) {
        drop(&I->Coarse_A[l]);
        drop(&I->L[l]);
        drop(&I->U[l]);
        I->Coarse_A[l] = level_matrix(s, l, 'A', I->Fine_sizes[l]);
        I->L[l] = sparse_zero(I->Fine_sizes[l]);
        I->U[l] = sparse_zero(I->Fine_sizes[l]);
        if (!I->Coarse_A[l]) continue;
        if (l > 0) incomplete_LU_factorization(I->Coarse_A[l], I->L[l], I->U[l]);
        else sparse_LU_factorization(I->Coarse_A[0], I->L[0], I->U[0]);
    }
}

/* the hierarchy as the reference's amg_system_info holds it: level
 * matrices, transfer operators, C/F labels and the finest row signs
 * (levels indexed as there: depth-1 finest, 0 coarsest; level 0 has no
 * P, R or labels) */
static void materialise(amg_system_info* I) {
    lina_solver* s = (lina_solver*)I->device;
    const int d = I->depth, n = I->Fine_sizes[d - 1];
    double* sg = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int l, i;
    I->Coarse_A = (sparse_matrix**)calloc((size_t)d, sizeof(sparse_matrix*));
    I->Prolongation = (sparse_matrix**)calloc((size_t)d, sizeof(sparse_matrix*));
    I->Restriction = (sparse_matrix**)calloc((size_t)d, sizeof(sparse_matrix*));
    I->L = (sparse_matrix**)calloc((size_t)d, sizeof(sparse_matrix*));
    I->U = (sparse_matrix**)calloc((size_t)d, sizeof(sparse_matrix*));
    I->Coarsening = (int**)calloc((size_t)d, sizeof(int*));
    for (l = 1; l < d; l++) {
        I->Prolongation[l] = level_matrix(s, l, 'P', I->Fine_sizes[l]);
        I->Restriction[l] = level_matrix(s, l, 'R', I->Coarse_sizes[l]);
        I->Coarsening[l] = (int*)malloc((size_t)(I->Fine_sizes[l] > 0 ? I->Fine_sizes[l] : 1) * sizeof(int));
        dropin_check(lina_amg_labels(s, l, I->Coarsening[l]), "AMG labels");
    }
    materialise_matrices(I);
    I->Signs = sparse_zero(n);
    lina_amg_signs(s, sg);
    for (i = 0; i < n; i++) I->Signs->Values[i][0] = sg[i];
    free(sg);
}

static void release_levels(amg_system_info* I) {
    int l;
    for (l = 0; l < I->depth; l++) {
        if (I->Coarse_A) drop(&I->Coarse_A[l]);
        if (I->Prolongation) drop(&I->Prolongation[l]);
        if (I->Restriction) drop(&I->Restriction[l]);
        if (I->L) drop(&I->L[l]);
        if (I->U) drop(&I->U[l]);
        if (I->Coarsening) free(I->Coarsening[l]);
    }
    free(I->Coarse_A);
    free(I->Prolongation);
    free(I->Restriction);
    free(I->L);
    free(I->U);
    free(I->Coarsening);
    drop(&I->Signs);
    I->Coarse_A = I->Prolongation = I->Restriction = I->L = I->U = NULL;
    I->Coarsening = NULL;
    I->Signs = NULL;
}

/* params of a hierarchy, refreshed from the globals before each solve */
typedef struct {
    lina_amg_params p;
} amg_side;

void dropin_amg_sync(void* info) {
    amg_system_info* I = (amg_system_info*)info;
    amg_side* sd = (amg_side*)I->settings;
    sd->p.smoother = g_smoother;
    sd->p.omega = g_alpha;
    sd->p.amg_eps = g_eps;
    sd->p.amg_max_iter = g_max_iter;
    dropin_check(lina_amg_set_params((lina_solver*)I->device, &sd->p), "AMG settings");
}

amg_system_info* AMG_setup(sparse_matrix* A, int dof, int depth, int smoothing_iterations,
                           int aggr_smoothing_iterations, int coarsening, int cycle_structure,
                           int prolongation_method,
                           void (*coarsest_solver)(sparse_matrix*, sparse_matrix*, double*, double*, int),
                           int* (*Get_iter)(int)) {
    amg_system_info* I;
    amg_side* sd;
    dropin_csr c;
    int l;
    if (cycle_structure != STAND_ALONE && cycle_structure != FULL_AMG)
        fprintf(stderr, "liblinalg_hip: unknown cycle structure %d; using STAND_ALONE\n", cycle_structure);
    if (coarsest_solver || Get_iter)
        fprintf(stderr, "liblinalg_hip: caller-defined coarse solvers and cycle counts are not used\n");
    sd = (amg_side*)calloc(1, sizeof(amg_side));
    lina_params_reference(&sd->p, dof, depth);
    sd->p.smoothing_iterations = smoothing_iterations;
    sd->p.aggr_smoothing_iterations = aggr_smoothing_iterations;
    sd->p.coarsening = coarsening;
    sd->p.eps_minus = g_eps_minus;
    sd->p.eps_plus = g_eps_plus;
    sd->p.matrix_tol = g_matrix_tol;
    sd->p.smoother = g_smoother;
    sd->p.omega = g_alpha;
    sd->p.amg_eps = g_eps;
    sd->p.amg_max_iter = g_max_iter;
    sd->p.cycle = cycle_structure == FULL_AMG ? LINA_CYCLE_FULL : LINA_CYCLE_V;
    c = dropin_to_csr(A);
    I = (amg_system_info*)calloc(1, sizeof(amg_system_info));
    I->device = lina_amg_create(c.n, c.rp, c.ci, c.v, &sd->p, 1);
    dropin_free_csr(&c);
    if (!I->device) {
        dropin_check(-1, "AMG_setup");
        free(sd);
        free(I);
        return NULL;
    }
    I->settings = sd;
    I->depth = lina_amg_depth((lina_solver*)I->device);
    I->Fine_sizes = (int*)calloc(I->depth, sizeof(int));
    I->Coarse_sizes = (int*)calloc(I->depth, sizeof(int));
    for (l = 0; l < I->depth; l++) {
        int f, co, na, np, nr;
        lina_amg_level_info((lina_solver*)I->device, l, &f, &co, &na, &np, &nr);
        I->Fine_sizes[l] = f;
        I->Coarse_sizes[l] = co;
    }
    I->AMG_cycle_structure = sd->p.cycle == LINA_CYCLE_FULL ? FULL_AMG : STAND_ALONE;
    I->AMG_prolongation_method = prolongation_method;
    I->AMG_smoothing_iterations = smoothing_iterations;
    I->AMG_aggr_smoothing_iterations = aggr_smoothing_iterations;
    I->AMG_coarsening = coarsening;
    I->AMG_info = g_info;
    I->AMG_deg_of_freedom = dof;
    I->Get_iter = AMG_default_iterations;
    I->main_size = A->size;
    materialise(I);
    return I;
}

int AMG_solve(amg_system_info* Data, double* F, double* Fine_Sol, amg_statistics* Stat) {
    int it;
    dropin_amg_sync(Data);
    if (Stat) {
        double rel = 0;
        it = lina_amg_solve_stat((lina_solver*)Data->device, F, Fine_Sol, &rel);
        dropin_check(it, "AMG_solve");
        if (it >= 0) AMG_add_stat(Stat, it, rel);
    } else {
        it = lina_amg_solve((lina_solver*)Data->device, F, Fine_Sol);
        dropin_check(it, "AMG_solve");
    }
    return it;
}

void AMG_recompute_coarse_matrices(amg_system_info* Info, sparse_matrix* New_matrix) {
    dropin_csr c = dropin_to_csr(New_matrix);
    dropin_check(lina_amg_recompute_csr((lina_solver*)Info->device, c.n, c.rp, c.ci, c.v),
                 "AMG_recompute_coarse_matrices");
    dropin_free_csr(&c);
    if (Info->Coarse_A) materialise_matrices(Info);
}

void AMG_free_data(amg_system_info* Data) {
    if (!Data) return;
    release_levels(Data);
    lina_amg_destroy((lina_solver*)Data->device);
    free(Data->settings);
    free(Data->Fine_sizes);
    free(Data->Coarse_sizes);
    Data->device = NULL;
    Data->settings = NULL;
    Data->Fine_sizes = Data->Coarse_sizes = NULL;
}

/* statistics (Multigrid.c AMG_init_stat / AMG_add_stat / AMG_print_stat) */
amg_statistics AMG_init_stat(const char* Out) {
    amg_statistics s;
    memset(&s, 0, sizeof(s));
    if (Out) strncpy(s.File_name, Out, sizeof(s.File_name) - 1);
    return s;
}

void AMG_add_stat(amg_statistics* Stat, int iter, double res) {
    Stat->Iterations = (int*)realloc(Stat->Iterations, (Stat->calls + 1) * sizeof(int));
    Stat->Residuals = (double*)realloc(Stat->Residuals, (Stat->calls + 1) * sizeof(double));
    Stat->Iterations[Stat->calls] = iter;
    Stat->Residuals[Stat->calls] = res;
    Stat->calls++;
    if (iter > Stat->max_iter) Stat->max_iter = iter;
    if (res > Stat->max_res) Stat->max_res = res;
}

void AMG_print_stat(amg_statistics* Stat) {
    FILE* f = Stat->File_name[0] ? fopen(Stat->File_name, "w") : NULL;
    int i;
    if (!f) return;
    for (i = 0; i < Stat->calls; i++) fprintf(f, "%d\t%e\n", Stat->Iterations[i], Stat->Residuals[i]);
    fclose(f);
}
