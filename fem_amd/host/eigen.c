/* LinA's eigenvalue searches (linear_algebra.c:3441-3450, 4802-4865): the
 * power iterations run on the GPU (lina_matrix_power) on a device copy of
 * the row-list matrix; the start vectors are drawn as in the reference
 * (srand(time(NULL)), rand()/RAND_MAX, normalised). */
#include <math.h>
#include <stdlib.h>
#include <time.h>

#include "dropin.h"

/* matrix_norm (linear_algebra.c:3441-3450): sum of |a_ij| over the rows / n */
double matrix_norm(sparse_matrix* A) {
    int i, j;
    double s = 0;
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) s += fabs(A->Values[i][j]);
    return s / A->size;
}

/* set_random_normal_vector (linear_algebra.c:4824-4829) */
void set_random_normal_vector(double* X, int n) {
    int i;
    double s;
    srand((unsigned)time(NULL));
    for (i = 0; i < n; i++) X[i] = (double)rand() / RAND_MAX;
    s = 1. / euklid_norm(X, n);
    for (i = 0; i < n; i++) X[i] *= s;
}

/* power_method (linear_algebra.c:4802-4822): Max_vec holds the start vector
 * (normalised) and ends as the last iterate; *max_eig enters as the estimate
 * the first convergence test compares with.  Original and eps_neg take no
 * part (the reference's use of them is commented out). */
int power_method(sparse_matrix* A, double* max_eig, double* Max_vec, int max_iter, double eps,
                 sparse_matrix* Original, double eps_neg) {
    int it;
    lina_matrix* m = dropin_device_matrix(A);
    (void)Original;
    (void)eps_neg;
    it = lina_matrix_power(m, Max_vec, max_iter, eps, max_eig);
    lina_matrix_destroy(m);
    dropin_check(it, "power_method");
    return it;
}

/* the two-stage search of get_most_negative_eigenvector and
 * get_neg_curvature_direction: the largest eigenvalue s of A from pre_iter
 * iterations, then the dominant eigenvalue of A - s I plus s.  The first
 * stage's estimate starts at 0 (the reference reads an uninitialised local
 * there); the second starts from the caller's *value, as in the reference. */
static int shifted_search(sparse_matrix* A, double* value, double* V, int pre_iter, int max_iter, double eps,
                          sparse_matrix* Original, double neg_eps) {
    int i, it, n = A->size;
    double shift = 0;
    sparse_matrix* B;
    set_random_normal_vector(V, n);
    power_method(A, &shift, V, pre_iter, eps, NULL, 0);
    B = clone(A);
    for (i = 0; i < n; i++) insert_sparse(B, -shift, i, i);
    set_random_normal_vector(V, n);
    it = power_method(B, value, V, max_iter, eps, Original, neg_eps);
    *value += shift;
    free_sparse(B);
    return it;
}

/* get_most_negative_eigenvector (linear_algebra.c:4831-4847) */
int get_most_negative_eigenvector(sparse_matrix* A, double* eigenvalue, double* Eigenvector, int max_iter,
                                  double eps) {
    return shifted_search(A, eigenvalue, Eigenvector, 100, max_iter, eps, NULL, 0);
}

/* get_neg_curvature_direction (linear_algebra.c:4849-4865) */
int get_neg_curvature_direction(sparse_matrix* A, double* curvature, double* Direction, int max_iter, double eps,
                                double neg_eps) {
    return shifted_search(A, curvature, Direction, 1000, max_iter, eps, A, neg_eps);
}
