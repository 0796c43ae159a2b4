/*
 * ref_harness.c -- ctypes glue around the reference LinA C, compiled together
 * with the reference's own sources into oracle/_ref/libreflina.so.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or loads
 * this file.  It converts flat CSR arrays to the reference's row-list
 * sparse_matrix (LinA/linear_algebra.h:29-34), drives the reference's own
 * entry points with the configuration a test asks for, and copies results
 * (hierarchy, labels, iterates, counters) back into caller-owned arrays.
 *
 * Determinism: aggressive_coarsening (LinA/Multigrid.c:566) reads the
 * uninitialised array Importance_2nd as RB-tree keys.  refh_init() sets
 * glibc's M_PERTURB so every malloc'd block starts zero-filled, which is the
 * value a fresh mmap gives and the value the oracle and the product assume.
 */
#include <malloc.h>
#include <string.h>
#include <time.h>
#include <omp.h>
#include "linear_algebra.h"
#include "PCGS.h"
#include "Multigrid.h"
#include "BiCGStab.h"
#include "reflective_newton.h"

void refh_sym_reset(void);

/* ---------------------------------------------------------------- memory */

/* 1: zero-fill every malloc (parity runs); 0: glibc defaults; 2: no
 * perturbation but a fixed mmap threshold, so blocks >= 128 KiB are always
 * fresh zero pages (timing runs in a fresh process) */
void refh_init(int mode) {
    if (mode == 1) mallopt(M_PERTURB, 0xFF);   /* malloc'd bytes = ~0xFF = 0 */
    else mallopt(M_PERTURB, 0);
    if (mode == 2) mallopt(M_MMAP_THRESHOLD, 128 * 1024);
}

/* build a row-list matrix from CSR (rows must be sorted, >= 1 entry each) */
sparse_matrix* refh_from_csr(int n, const int* rowptr, const int* col, const double* val) {
    int i, j;
    sparse_matrix* A = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    A->size = n;
    A->Len = (int*)malloc(n * sizeof(int));
    A->Values = (double**)malloc(n * sizeof(double*));
    A->Indices = (int**)malloc(n * sizeof(int*));
    for (i = 0; i < n; i++) {
        int len = rowptr[i + 1] - rowptr[i];
        A->Len[i] = len;
        A->Values[i] = (double*)malloc((len > 0 ? len : 1) * sizeof(double));
        A->Indices[i] = (int*)malloc((len > 0 ? len : 1) * sizeof(int));
        for (j = 0; j < len; j++) {
            A->Values[i][j] = val[rowptr[i] + j];
            A->Indices[i][j] = col[rowptr[i] + j];
        }
    }
    return A;
}

void refh_free(sparse_matrix* A) {
    if (A == NULL) return;
    Free_sparse(A);
    free(A);
}

int refh_size(sparse_matrix* A) { return A ? A->size : 0; }

int refh_nnz(sparse_matrix* A) {
    int i, s = 0;
    if (A == NULL) return 0;
    for (i = 0; i < A->size; i++) s += A->Len[i];
    return s;
}

/* copy out as CSR into caller arrays sized (n+1, nnz, nnz) */
void refh_to_csr(sparse_matrix* A, int* rowptr, int* col, double* val) {
    int i, j, k = 0;
    rowptr[0] = 0;
    for (i = 0; i < A->size; i++) {
        for (j = 0; j < A->Len[i]; j++) {
            col[k] = A->Indices[i][j];
            val[k] = A->Values[i][j];
            k++;
        }
        rowptr[i + 1] = k;
    }
}

/* ------------------------------------------------ plugged-in smoothers
 * The reference takes its smoother through AMG_set_smoother
 * (LinA/Multigrid.c:1967-1969, signature Multigrid.h:88).  Besides its own
 * AMG_SOR_smoother / AMG_parallel_Jacobi_smoother, these two GPU-friendly
 * smoothers are plugged into the reference's V-cycle so that the reference
 * itself produces the iterates the product's multicolour-SOR and l1-Jacobi
 * paths must reproduce.  Colouring: greedy in natural row order over the
 * symmetrised pattern of nonzero off-diagonal values.
 */
static int* greedy_colouring(sparse_matrix* A, int* ncol_out) {
    int n = A->size, i, j, c, ncol = 0;
    int* color = (int*)malloc(n * sizeof(int));
    /* symmetrised adjacency: count transposed entries */
    int* tlen = (int*)calloc(n, sizeof(int));
    for (i = 0; i < n; i++)
        for (j = 0; j < A->Len[i]; j++)
            if (A->Indices[i][j] != i && A->Values[i][j] != 0) tlen[A->Indices[i][j]]++;
    int** tadj = (int**)malloc(n * sizeof(int*));
    int* tfill = (int*)calloc(n, sizeof(int));
    for (i = 0; i < n; i++) tadj[i] = (int*)malloc((tlen[i] + 1) * sizeof(int));
    for (i = 0; i < n; i++)
        for (j = 0; j < A->Len[i]; j++) {
            int k = A->Indices[i][j];
            if (k != i && A->Values[i][j] != 0) tadj[k][tfill[k]++] = i;
        }
    int cap = 64;
    int* mark = (int*)malloc(cap * sizeof(int));
    for (c = 0; c < cap; c++) mark[c] = -1;
    for (i = 0; i < n; i++) {
        for (j = 0; j < A->Len[i]; j++) {
            int k = A->Indices[i][j];
            if (k < i && A->Values[i][j] != 0) mark[color[k]] = i;
        }
        for (j = 0; j < tlen[i]; j++) {
            int k = tadj[i][j];
            if (k < i) mark[color[k]] = i;
        }
        c = 0;
        while (c < cap && mark[c] == i) c++;
        if (c == cap) {
            int old = cap, q;
            cap *= 2;
            mark = (int*)realloc(mark, cap * sizeof(int));
            for (q = old; q < cap; q++) mark[q] = -1;
        }
        color[i] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    for (i = 0; i < n; i++) free(tadj[i]);
    free(tadj); free(tlen); free(tfill); free(mark);
    *ncol_out = ncol;
    return color;
}

void refh_mc_sor_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha,
                          int fix, int iter, int deg_freedom) {
    int n = A->size, ncol, k, c, i, j, ind, diag;
    double sum;
    int* color = greedy_colouring(A, &ncol);
    for (k = 0; k < iter; k++)
        for (c = 0; c < ncol; c++)
            for (i = 0; i < n; i++)
                if (color[i] == c && Label[i] != fix) {
                    sum = F[i];
                    diag = -1;
                    for (j = 0; j < A->Len[i]; j++) {
                        ind = A->Indices[i][j];
                        if (ind != i) sum -= A->Values[i][j] * Sol[ind];
                        else diag = j;
                    }
                    Sol[i] = (1. - alpha) * Sol[i] + alpha * sum / A->Values[i][diag];
                }
    free(color);
}

void refh_l1_jacobi_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha,
                             int fix, int iter, int deg_freedom) {
    int n = A->size, k, i, j;
    double sum, d;
    double* Old = (double*)malloc(n * sizeof(double));
    for (k = 0; k < iter; k++) {
        memcpy(Old, Sol, n * sizeof(double));
        for (i = 0; i < n; i++)
            if (Label[i] != fix) {
                sum = F[i];
                d = 0;
                for (j = 0; j < A->Len[i]; j++) {
                    sum -= A->Values[i][j] * Old[A->Indices[i][j]];
                    d += fabs(A->Values[i][j]);
                }
                Sol[i] = Old[i] + sum / d;
            }
    }
    free(Old);
}

/* Symmetric variants: the V-cycle (Multigrid.c:1549,1565) calls the smoother
 * twice per level visit (pre, post).  These smoothers alternate direction on
 * successive calls for the same matrix -- forward on the pre-smoothing call,
 * reversed (colours / rows in reverse order) on the post-smoothing call -- so
 * one V-cycle is a symmetric operator.  refh_sym_reset() restarts the
 * alternation (called before every preconditioner application). */
#define SYM_SLOTS 64
static sparse_matrix* sym_key[SYM_SLOTS];
static int sym_phase[SYM_SLOTS];
void refh_sym_reset(void) { memset(sym_key, 0, sizeof(sym_key)); memset(sym_phase, 0, sizeof(sym_phase)); }
static int sym_next_phase(sparse_matrix* A) {
    int s;
    for (s = 0; s < SYM_SLOTS; s++) {
        if (sym_key[s] == A) { int p = sym_phase[s]; sym_phase[s] ^= 1; return p; }
        if (sym_key[s] == NULL) { sym_key[s] = A; sym_phase[s] = 1; return 0; }
    }
    return 0;
}

void refh_mc_ssor_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha,
                           int fix, int iter, int deg_freedom) {
    int n = A->size, ncol, k, c, cc, i, j, ind, diag;
    double sum;
    int backward = sym_next_phase(A);
    int* color = greedy_colouring(A, &ncol);
    for (k = 0; k < iter; k++)
        for (cc = 0; cc < ncol; cc++) {
            c = backward ? ncol - 1 - cc : cc;
            for (i = 0; i < n; i++)
                if (color[i] == c && Label[i] != fix) {
                    sum = F[i];
                    diag = -1;
                    for (j = 0; j < A->Len[i]; j++) {
                        ind = A->Indices[i][j];
                        if (ind != i) sum -= A->Values[i][j] * Sol[ind];
                        else diag = j;
                    }
                    Sol[i] = (1. - alpha) * Sol[i] + alpha * sum / A->Values[i][diag];
                }
        }
    free(color);
}

void refh_ssor_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha,
                        int fix, int iter, int deg_freedom) {
    int n = A->size, k, ii, i, j, ind, diag;
    double sum;
    int backward = sym_next_phase(A);
    for (k = 0; k < iter; k++)
        for (ii = 0; ii < n; ii++) {
            i = backward ? n - 1 - ii : ii;
            if (Label[i] == fix) continue;
            sum = F[i];
            diag = -1;
            for (j = 0; j < A->Len[i]; j++) {
                ind = A->Indices[i][j];
                if (ind != i) sum -= A->Values[i][j] * Sol[ind];
                else diag = j;
            }
            Sol[i] = (1. - alpha) * Sol[i] + alpha * sum / A->Values[i][diag];
        }
}

/* Level-dependent smoother: multicolour SSOR on the finest level, l1-Jacobi
 * with its own sweep count on every coarser level (symmetric V-cycle). */
static sparse_matrix* hyb_finest = NULL;
static int hyb_coarse_sweeps = 2;
void refh_hybrid_smoother(sparse_matrix* A, double* F, double* Sol, int* Label, double alpha, int fix, int iter,
                          int deg_freedom) {
    if (A == hyb_finest) refh_mc_ssor_smoother(A, F, Sol, Label, alpha, fix, iter, deg_freedom);
    else refh_l1_jacobi_smoother(A, F, Sol, Label, alpha, fix, hyb_coarse_sweeps, deg_freedom);
}
void refh_set_hybrid(int coarse_sweeps) { hyb_coarse_sweeps = coarse_sweeps; }

int refh_colour_count(sparse_matrix* A, int* color_out) {
    int ncol;
    int* c = greedy_colouring(A, &ncol);
    if (color_out) memcpy(color_out, c, A->size * sizeof(int));
    free(c);
    return ncol;
}

/* smoother ids shared with the product's lina_hip.h */
static void set_smoother(int id) {
    switch (id) {
        case 1: AMG_set_smoother(&AMG_parallel_Jacobi_smoother); break;
        case 2: AMG_set_smoother(&refh_mc_sor_smoother); break;
        case 3: AMG_set_smoother(&refh_l1_jacobi_smoother); break;
        case 4: AMG_set_smoother(&refh_mc_ssor_smoother); break;
        case 5: AMG_set_smoother(&refh_ssor_smoother); break;
        case 6: AMG_set_smoother(&refh_hybrid_smoother); break;
        default: AMG_set_smoother(&AMG_SOR_smoother); break;
    }
}

/* ------------------------------------------------------------- AMG setup */

/* cycle_structure of the next refh_amg_setup (STAND_ALONE or FULL_AMG) */
static int h_cycle = STAND_ALONE;
void refh_set_cycle(int cycle) { h_cycle = cycle; }

/* AMG_setup with the reference's own arguments (Multigrid.c:1613-1754) */
amg_system_info* refh_amg_setup(sparse_matrix* A, int dof, int depth, int smiter, int aggr_smiter,
                                int coarsening, int matrix_tol_set, double matrix_tol) {
    AMG_print_info(NO);
    if (matrix_tol_set) AMG_set_matrix_tolerance(matrix_tol);
    return AMG_setup(A, dof, depth, smiter, aggr_smiter, coarsening, h_cycle, DIRECT, NULL, NULL);
}

void refh_amg_free(amg_system_info* I) {
    if (I == NULL) return;
    AMG_free_data(I);
    free(I);
}

void refh_amg_sizes(amg_system_info* I, int* fine, int* coarse) {
    int l;
    for (l = 0; l < I->depth; l++) {
        fine[l] = I->Fine_sizes[l];
        coarse[l] = I->Coarse_sizes[l];
    }
}

sparse_matrix* refh_amg_A(amg_system_info* I, int l) { return I->Coarse_A[l]; }
sparse_matrix* refh_amg_P(amg_system_info* I, int l) { return I->Prolongation[l]; }
sparse_matrix* refh_amg_R(amg_system_info* I, int l) { return I->Restriction[l]; }
sparse_matrix* refh_amg_L(amg_system_info* I, int l) { return I->L[l]; }
sparse_matrix* refh_amg_U(amg_system_info* I, int l) { return I->U[l]; }
sparse_matrix* refh_amg_signs(amg_system_info* I) { return I->Signs; }
void refh_amg_labels(amg_system_info* I, int l, int* out) {
    memcpy(out, I->Coarsening[l], I->Fine_sizes[l] * sizeof(int));
}

void refh_amg_recompute(amg_system_info* I, sparse_matrix* M) { AMG_recompute_coarse_matrices(I, M); }

/* AMG_solve with explicit iteration controls (Multigrid.c:1863-1931) */
int refh_amg_solve(amg_system_info* I, double* F, double* X, int smoother, double sor_alpha,
                   double eps, int max_iter) {
    set_smoother(smoother);
    hyb_finest = I->Coarse_A[I->depth - 1];
    AMG_set_SOR_relaxation_coeff(sor_alpha);
    AMG_set_eps(eps);
    AMG_set_max_iter(max_iter);
    AMG_print_info(NO);
    refh_sym_reset();
    return AMG_solve(I, F, X, NULL);
}

/* ------------------------------------------------------------------- pcg */

static amg_system_info* h_amg = NULL;
static int* h_calls = NULL;        /* V-cycles returned by each precon call */
static int h_ncalls = 0, h_cap = 0;
static int h_zero_init = 0;        /* V-cycle variant: Z := 0 before every call */

static int h_amg_precon(int n, double* R, double* Z) {
    int k;
    if (h_zero_init) memset(Z, 0, n * sizeof(double));
    refh_sym_reset();
    k = AMG_solve(h_amg, R, Z, NULL);
    if (h_calls && h_ncalls < h_cap) h_calls[h_ncalls] = k;
    h_ncalls++;
    return k;
}

/* pcg (PCGS.c:163-233) with PCG_default's matvec and an AMG preconditioner
 * built exactly as PCG_set_AMG_preconditioner does (PCGS.c:118-129), except
 * that smoother, relaxation, eps, max V-cycles, smoothing sweeps and
 * coarsening are arguments, and zero_init=1 selects the linear V-cycle
 * preconditioner variant (Z zeroed before each AMG_solve call; with
 * amg_max_iter=1 this is one V-cycle per pcg iteration).  X and B are in/out as in the reference.
 * Returns the pcg iteration count; *res = get_pcg_residuum(). */
int refh_pcg_amg(sparse_matrix* A, int dof, int depth, int smiter, int aggr_smiter, int coarsening,
                 int smoother, double sor_alpha, double amg_eps, int amg_max_iter, int zero_init,
                 double* X, double* B, double tol, int max_iter, double* res,
                 int* calls, int calls_cap, int* ncalls, double* t_setup, double* t_solve,
                 amg_system_info** keep) {
    struct timespec t0, t1, t2;
    int it;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    PCG_default(A, NULL, NULL);
    AMG_print_info(NO);
    h_amg = AMG_setup(A, dof, depth, smiter, aggr_smiter, coarsening, h_cycle, DIRECT, NULL, NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    set_smoother(smoother);
    hyb_finest = h_amg->Coarse_A[h_amg->depth - 1];
    AMG_set_SOR_relaxation_coeff(sor_alpha);
    AMG_set_eps(amg_eps);
    AMG_set_max_iter(amg_max_iter);
    PCG_set_precon(&h_amg_precon);
    h_calls = calls; h_cap = calls_cap; h_ncalls = 0; h_zero_init = zero_init;
    it = pcg(X, B, A->size, tol, max_iter);
    h_zero_init = 0;
    clock_gettime(CLOCK_MONOTONIC, &t2);
    if (res) *res = get_pcg_residuum();
    if (ncalls) *ncalls = h_ncalls;
    if (t_setup) *t_setup = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    if (t_solve) *t_solve = (t2.tv_sec - t1.tv_sec) + 1e-9 * (t2.tv_nsec - t1.tv_nsec);
    h_calls = NULL; h_cap = 0;
    if (keep) *keep = h_amg;
    else { AMG_free_data(h_amg); free(h_amg); }
    h_amg = NULL;
    return it;
}

/* ------------------------------------------ an assembled hierarchy (timing)
 *
 * The CPU baseline of the 2.1M-DOF benchmark times the reference's pcg and
 * V-cycle (PCGS.c:163-233, Multigrid.c:1536-1580, the smoothers above) on
 * the full system.  Its AMG_setup takes ~14 minutes there on one core, so
 * the hierarchy is assembled from level matrices built elsewhere (the
 * benchmark hands over liblina_hip's host setup, which the tests pin bit for
 * bit to AMG_setup), and the coarsest level is solved with a precomputed
 * dense inverse of its dense rows (the same multiply-adds as the triangular
 * solves of its LU factors) instead of sparse_LU_factorization's O(n^3)
 * factorisation.  Timing only: iterates are not compared from this path. */
static double* hd_inv = NULL;
static int* hd_idx = NULL;
static double* hd_tdiag = NULL;
static int hd_nd = 0;

static void refh_dense_coarse(sparse_matrix* L, sparse_matrix* U, double* F, double* Sol, int size) {
    int i, g, h;
    (void)L;
    (void)U;
    for (i = 0; i < size; i++)
        if (hd_tdiag[i] != 0) Sol[i] = F[i] / hd_tdiag[i];
    for (g = 0; g < hd_nd; g++) {
        double s = 0;
        const double* row = hd_inv + (size_t)g * hd_nd;
        for (h = 0; h < hd_nd; h++) s += row[h] * F[hd_idx[h]];
        Sol[hd_idx[g]] = s;
    }
}

/* A[l], P[l], R[l] (l >= 1), labels[l] (l >= 1): level l as Multigrid.c
 * indexes it (depth-1 finest); signs: the finest row signs; the coarsest
 * level's dense rows didx[0..nd) with the row-major inverse inv, tdiag[i]
 * the diagonal of each other row (0 for dense rows).  The finest level's
 * ILU(0) factors (AMG_solve's residual diagnostics) are the reference's. */
amg_system_info* refh_amg_assemble(int depth, int dof, int smiter, int aggr_smiter, int coarsening,
                                   sparse_matrix** A, sparse_matrix** P, sparse_matrix** R, int** labels,
                                   const double* signs, int nd, const int* didx, const double* inv,
                                   const double* tdiag) {
    amg_system_info* I = (amg_system_info*)calloc(1, sizeof(amg_system_info));
    int l, i, n0 = A[0]->size, n = A[depth - 1]->size;
    I->depth = depth;
    I->Coarse_A = (sparse_matrix**)calloc(depth, sizeof(sparse_matrix*));
    I->Prolongation = (sparse_matrix**)calloc(depth, sizeof(sparse_matrix*));
    I->Restriction = (sparse_matrix**)calloc(depth, sizeof(sparse_matrix*));
    I->L = (sparse_matrix**)calloc(depth, sizeof(sparse_matrix*));
    I->U = (sparse_matrix**)calloc(depth, sizeof(sparse_matrix*));
    I->Coarsening = (int**)calloc(depth, sizeof(int*));
    I->Fine_sizes = (int*)calloc(depth, sizeof(int));
    I->Coarse_sizes = (int*)calloc(depth, sizeof(int));
    for (l = 0; l < depth; l++) {
        I->Coarse_A[l] = A[l];
        I->Fine_sizes[l] = A[l]->size;
        I->L[l] = sparse_zero(A[l]->size);
        I->U[l] = sparse_zero(A[l]->size);
        if (l == 0) continue;
        I->Prolongation[l] = P[l];
        I->Restriction[l] = R[l];
        I->Coarsening[l] = labels[l];
        I->Coarse_sizes[l] = R[l]->size;
    }
    incomplete_LU_factorization(I->Coarse_A[depth - 1], I->L[depth - 1], I->U[depth - 1]);
    I->Signs = sparse_identity(n);
    for (i = 0; i < n; i++) I->Signs->Values[i][0] = signs[i];
    I->AMG_cycle_structure = STAND_ALONE;
    I->AMG_prolongation_method = DIRECT;
    I->AMG_smoothing_iterations = smiter;
    I->AMG_aggr_smoothing_iterations = aggr_smiter;
    I->AMG_coarsening = coarsening;
    I->AMG_info = NO;
    I->AMG_deg_of_freedom = dof;
    I->AMG_coarsest_solver = &refh_dense_coarse;
    I->Get_iter = &AMG_default_iterations;
    free(hd_inv);
    free(hd_idx);
    free(hd_tdiag);
    hd_nd = nd;
    hd_inv = (double*)malloc(((size_t)nd * nd + 1) * sizeof(double));
    hd_idx = (int*)malloc((nd + 1) * sizeof(int));
    hd_tdiag = (double*)malloc((n0 + 1) * sizeof(double));
    memcpy(hd_inv, inv, (size_t)nd * nd * sizeof(double));
    memcpy(hd_idx, didx, nd * sizeof(int));
    memcpy(hd_tdiag, tdiag, n0 * sizeof(double));
    return I;
}

void refh_set_threads(int n) { omp_set_num_threads(n); }
int refh_get_threads(void) { return omp_get_max_threads(); }

/* pcg preconditioned by AMG_solve on an assembled hierarchy (the settings of
 * refh_pcg_amg); *t_solve: wall time of pcg */
int refh_pcg_assembled(amg_system_info* I, sparse_matrix* A, int smoother, double sor_alpha, double amg_eps,
                       int amg_max_iter, int zero_init, double* X, double* B, double tol, int max_iter,
                       double* t_solve) {
    struct timespec t1, t2;
    int it;
    PCG_default(A, NULL, NULL);
    AMG_print_info(NO);
    h_amg = I;
    set_smoother(smoother);
    hyb_finest = I->Coarse_A[I->depth - 1];
    AMG_set_SOR_relaxation_coeff(sor_alpha);
    AMG_set_eps(amg_eps);
    AMG_set_max_iter(amg_max_iter);
    PCG_set_precon(&h_amg_precon);
    h_calls = NULL; h_cap = 0; h_ncalls = 0; h_zero_init = zero_init;
    clock_gettime(CLOCK_MONOTONIC, &t1);
    it = pcg(X, B, A->size, tol, max_iter);
    clock_gettime(CLOCK_MONOTONIC, &t2);
    h_zero_init = 0;
    h_amg = NULL;
    if (t_solve) *t_solve = (t2.tv_sec - t1.tv_sec) + 1e-9 * (t2.tv_nsec - t1.tv_nsec);
    return it;
}

/* the reference's own PCG_set_AMG_preconditioner + pcg, untouched */
int refh_pcg_amg_default(sparse_matrix* A, int dof, int depth, double* X, double* B, double tol,
                         int max_iter, double* res, double* t_setup, double* t_solve) {
    struct timespec t0, t1, t2;
    int it;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    PCG_default(A, NULL, NULL);
    AMG_set_smoother(&AMG_SOR_smoother);
    PCG_set_AMG_preconditioner(A, dof, depth);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    it = pcg(X, B, A->size, tol, max_iter);
    clock_gettime(CLOCK_MONOTONIC, &t2);
    if (res) *res = get_pcg_residuum();
    if (t_setup) *t_setup = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    if (t_solve) *t_solve = (t2.tv_sec - t1.tv_sec) + 1e-9 * (t2.tv_nsec - t1.tv_nsec);
    return it;
}

/* pcg with the ILU(0) preconditioner of PCG_default (PCGS.c:29-34,61-108) */
int refh_pcg_ilu(sparse_matrix* A, double* X, double* B, double tol, int max_iter, double* res) {
    int n = A->size, it;
    sparse_matrix* L = sparse_zero(n);
    sparse_matrix* U = sparse_zero(n);
    incomplete_LU_factorization(A, L, U);
    PCG_default(A, L, U);
    it = pcg(X, B, n, tol, max_iter);
    if (res) *res = get_pcg_residuum();
    Free_sparse(L); free(L);
    Free_sparse(U); free(U);
    return it;
}

/* ILU(0) factors (linear_algebra.c:1952-1982); returns the reference's flag */
int refh_ilu(sparse_matrix* A, sparse_matrix** Lo, sparse_matrix** Uo) {
    int n = A->size;
    sparse_matrix* L = sparse_zero(n);
    sparse_matrix* U = sparse_zero(n);
    int ok = incomplete_LU_factorization(A, L, U);
    *Lo = L; *Uo = U;
    return ok;
}

void refh_L_invert(sparse_matrix* L, double* x) { L_triang_invert(L, x); }
void refh_U_invert(sparse_matrix* U, double* x) { U_triang_invert(U, x); }

void refh_sparse_LU(sparse_matrix* A, sparse_matrix** Lo, sparse_matrix** Uo) {
    int n = A->size;
    *Lo = sparse_zero(n);
    *Uo = sparse_zero(n);
    sparse_LU_factorization(A, *Lo, *Uo);
}

sparse_matrix* refh_product(sparse_matrix* A, sparse_matrix* B) { return sparse_product(A, B); }

/* ---------------------------------------------------------------- BiCGStab */

static sparse_matrix* h_bi_A = NULL;
static void h_bi_matmult(int n, double* y, double* z) {
    int i, j;
    double sum;
    for (i = 0; i < n; i++) {
        sum = 0;
        for (j = 0; j < h_bi_A->Len[i]; j++) sum += h_bi_A->Values[i][j] * y[h_bi_A->Indices[i][j]];
        z[i] = sum;
    }
}

/* BiCGStab_solve (BiCGStab.c:115-198) with the AMG preconditioner set up as
 * BI_set_AMG_preconditioner does (BiCGStab.c:98-113); bi_eps/bi_max_iter are
 * the BI_set_eps/BI_set_max_iter values. */
int refh_bicgstab_amg(sparse_matrix* A, int dof, int depth, int smoother, double sor_alpha,
                      double bi_eps, int bi_max_iter, double* X, double* b) {
    int it;
    h_bi_A = A;
    BI_print_info(NO);
    BI_set_stat("amg", "bicgstab");
    BI_set_mat_mult(&h_bi_matmult);
    BI_set_AMG_preconditioner(A, dof, depth);
    set_smoother(smoother);
    AMG_set_SOR_relaxation_coeff(sor_alpha);
    BI_set_eps(bi_eps);
    BI_set_max_iter(bi_max_iter);
    it = BiCGStab_solve(X, b, A->size);
    amg_system_info* I = BI_get_AMG_setup_data();
    AMG_free_data(I); free(I);
    BI_set_AMG_setup_data(NULL);
    return it;
}

/* -------------------------------------------------------- reflective Newton */

/* reflective_newton (reflective_newton.c:660-711) on the quadratic
 * q(x) = x.H.x/2 + x.b with the reference's default functional/gradient,
 * box constraints lower<=x<=upper on every index (NaN = unbounded), and the
 * inner solver `solver` (RN_BICGSTAB=1, RN_AMG=2, RN_PCG=3).  For the AMG
 * variants the hierarchy is built from H with (dof, depth) as
 * RN_init_BiCGStab does (reflective_newton.c:205-217). */
int refh_reflective_newton(sparse_matrix* H, double* b, double* lower, double* upper, double* X,
                           int solver, int dof, int depth, int max_iter, double eps, double delta,
                           double* q_out) {
    int n = H->size, it;
    int* idx = get_all_indices(n);
    RN_print_info(NO);
    AMG_print_info(NO);
    BI_print_info(NO);
    set_RN_Hesse(H);
    set_RN_linear(b);
    set_RN_functional(NULL);
    RN_set_gradient(NULL);
    RN_set_Hesse_function(NULL, n);
    set_RN_Hesse(H);
    set_RN_constraints(idx, lower, upper, n);
    RN_set_params(max_iter, eps, 0.1, 0.9, 0.1, delta, 1., 0.6, 1.1, 1e-3, 1000.);
    RN_set_local_solver(NULL);
    RN_set_solver(solver);
    amg_system_info* I = NULL;
    if (solver == RN_BICGSTAB) {
        BI_set_stat("amg", "bicgstab");
        BI_set_eps(1e-3);
        BI_set_max_iter(10);
        BI_set_mat_mult(&RN_default_mat_mult);
        BI_set_AMG_preconditioner(H, dof, depth);
        I = BI_get_AMG_setup_data();
        RN_set_AMG_data(I);
    } else if (solver == RN_AMG) {
        I = AMG_setup(H, dof, depth, 2, 3, AGGRESSIVE, STAND_ALONE, DIRECT, NULL, NULL);
        AMG_set_eps(1e-4);
        AMG_set_max_iter(20);
        RN_set_AMG_data(I);
    }
    AMG_set_smoother(&AMG_SOR_smoother);
    AMG_set_SOR_relaxation_coeff(1.0);
    it = reflective_newton(X);
    if (q_out) {
        double* Hx = sparse_mult(H, X);
        *q_out = scalar(X, Hx, n) / 2. + scalar(b, X, n);
        free(Hx);
    }
    if (I) { AMG_free_data(I); free(I); }
    BI_set_AMG_setup_data(NULL);
    free(idx);
    return it;
}

/* RN_get_M_D (reflective_newton.c:250-262) for H, gradient G and iterate X
 * with bounds on every index; D receives the scaling */
extern sparse_matrix* RN_get_M_D(double* Gradient, double* X, double* D);
sparse_matrix* refh_rn_get_M_D(sparse_matrix* H, double* G, double* X, double* lower, double* upper, double* D) {
    int n = H->size;
    int* idx = get_all_indices(n);
    RN_print_info(NO);
    set_RN_Hesse(H);
    set_RN_constraints(idx, lower, upper, n);
    sparse_matrix* M = RN_get_M_D(G, X, D);
    set_RN_constraints(NULL, NULL, NULL, 0);
    free(idx);
    return M;
}

/* ------------------------------------------------------------ GMRES(m) */
#include "itlin.h"
extern void gmres(int n, double* y, MATVEC* matvec, PRECON* preconr, PRECON* preconl, double* b,
                  struct ITLIN_OPT* opt, struct ITLIN_INFO* info);
extern void PCG_mult(int n, double* y, double* z);
extern int PCG_preconlr(int n, double* z, double* w);
static void h_gm_matvec(int n, double* y, double* z) { PCG_mult(n, y, z); }
static void h_gm_precon(int n, double* z, double* w) { PCG_preconlr(n, z, w); }

/* FEM.c's GMRES solver path (FEM.c:591-611, 730-745): ILU(0) factors of A,
 * gmres(n, H, matvec, preconlr, NULL, F) with CheckOnRestart on H = x, then
 * x = preconlr(H).  Returns the iteration count reported in info. */
int refh_gmres_ilu(sparse_matrix* A, double* x, double* b, int m, double tol, int max_iter, double* eta_out) {
    int n = A->size, i;
    struct ITLIN_OPT opt;
    struct ITLIN_INFO info;
    sparse_matrix* L = sparse_zero(n);
    sparse_matrix* U = sparse_zero(n);
    double* H = (double*)malloc(n * sizeof(double));
    incomplete_LU_factorization(A, L, U);
    PCG_default(A, L, U);
    memset(&opt, 0, sizeof(opt));
    memset(&info, 0, sizeof(info));
    opt.tol = tol;
    opt.i_max = m;
    opt.maxiter = max_iter;
    opt.termcheck = CheckOnRestart;
    opt.errorlevel = None;
    opt.monitorlevel = None;
    opt.datalevel = None;
    for (i = 0; i < n; i++) H[i] = x[i];
    gmres(n, H, &h_gm_matvec, &h_gm_precon, NULL, b, &opt, &info);
    PCG_preconlr(n, H, x);
    if (eta_out) *eta_out = info.precision;
    free(H);
    Free_sparse(L); free(L);
    Free_sparse(U); free(U);
    return info.iter;
}
