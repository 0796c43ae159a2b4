/* Stationary sweeps, residual helpers, extended-precision norms and the
 * thread controls of the LinA drop-in (host C on caller-owned host data):
 * SOR (linear_algebra.c:2464-2512), gauss_seidel (:2514-2544), residuum
 * (:1999-2006), hp_scalar / hp_euklid_norm (:265-282), set_thread_num and
 * set_omp_chunk_size (:23-36).  The GPU solvers' own smoothers are device
 * kernels; these are the standalone routines LinA's callers use directly. */
#include <math.h>
#include <omp.h>

#include "dropin.h"

int lg_thread_num = 1;   /* linear_algebra.c:6-7 */
int omp_chunk_size = 1000;

/* the host threads of the library's setup work (AMG coarsening, Galerkin
 * products); AUTOMATIC (-1): every processor */
void set_thread_num(int n) {
    lg_thread_num = n == AUTOMATIC ? omp_get_num_procs() : n;
    if (lg_thread_num < 1) lg_thread_num = 1;
    omp_set_num_threads(lg_thread_num);
    printf("liblinalg_hip: %d host threads\n", lg_thread_num);
}

void set_omp_chunk_size(int chunk) { omp_chunk_size = chunk; }

long double hp_scalar(double* x1, double* x2, int n) {
    long double s = 0;
    int i;
    for (i = 0; i < n; i++) s += (long double)x1[i] * x2[i];
    return s;
}

long double hp_euklid_norm(double* x1, int n) { return sqrtl(hp_scalar(x1, x1, n)); }

/* one forward SOR pass: x_i = (1 - alpha) x_i + alpha (f_i - sum_{j != i}
 * a_ij x_j) / a_ii over the rows in order, each row summed in its stored
 * order; -1 when a row has no diagonal entry */
static int sor_pass(sparse_matrix* A, double* F, double* X, double alpha) {
    int i, j;
    for (i = 0; i < A->size; i++) {
        double s = F[i], d = 0;
        int has = 0;
        for (j = 0; j < A->Len[i]; j++) {
            const int c = A->Indices[i][j];
            if (c == i) {
                d = A->Values[i][j];
                has = 1;
            } else {
                s -= A->Values[i][j] * X[c];
            }
        }
        if (!has) {
            fprintf(stderr, "SOR: row %d has no diagonal entry\n", i);
            return -1;
        }
        X[i] = (1. - alpha) * X[i] + alpha * s / d;
    }
    return 0;
}

/* iter > 0: that many passes; iter <= 0: passes until |F - A Sol| <= 1e-10,
 * at most -iter of them, the residual printed after each (as LinA does) */
void SOR(sparse_matrix* A, double* F, double* Sol, double alpha, int iter) {
    int k;
    if (iter > 0) {
        for (k = 0; k < iter; k++)
            if (sor_pass(A, F, Sol, alpha) < 0) return;
        return;
    }
    k = 0;
    do { /* (iter = 0 still runs one pass) */
        double r;
        if (sor_pass(A, F, Sol, alpha) < 0) return;
        r = matrix_residuum(A, Sol, F);
        printf("residuum: %e\n", r);
        k++;
        if (!(fabs(r) > 1e-10)) break;
    } while (k < -iter);
    if (k == -iter) printf("Warning: SOR did not converge after %d iterations\n", -iter);
}

double residuum(void (*Matvec)(int n, double* y, double* z), double* Sol, double* b, int n) {
    double* R = zero_vector(n);
    double r;
    Matvec(n, Sol, R);
    vector_add(R, b, -1, n);
    r = euklid_norm(R, n);
    free(R);
    return r;
}

/* the caller's sweep Inv_NB_matvec applied `iterations` times, or (iterations
 * <= 0) until the residual of NB_matvec falls by tol, at most -iterations
 * sweeps; a sweep returning < 0 stops the loop.  Returns the sweeps run. */
int gauss_seidel(int (*Inv_NB_matvec)(int n, double* x, double* b), void (*NB_matvec)(int n, double* y, double* z),
                 double* Solution, double* b, int size, int iterations, double tol) {
    int k = 0;
    double r0, r;
    if (iterations > 0) {
        for (k = 0; k < iterations; k++)
            if (Inv_NB_matvec(size, Solution, b) < 0) {
                printf("gauss_seidel: sweep failed, stopping\n");
                break;
            }
        return iterations;
    }
    r0 = residuum(NB_matvec, Solution, b, size);
    do {
        if (Inv_NB_matvec(size, Solution, b) < 0) {
            printf("gauss_seidel: sweep failed, stopping\n");
            break;
        }
        r = residuum(NB_matvec, Solution, b, size);
        k++;
    } while (r / r0 > tol && k < -iterations);
    return k;
}
