/* A LinA caller in plain C, unchanged between the reference library and the
 * drop-in: a 2-D Poisson system assembled with sparse_zero/insert_sparse,
 * solved by pcg with PCG_set_AMG_preconditioner.  Prints the iteration
 * count, get_pcg_residuum and the true relative residual. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "PCGS.h"
#include "linear_algebra.h"

int main(int argc, char** argv) {
    int m = argc > 1 ? atoi(argv[1]) : 65;
    int n = m * m, i, r, c, it;
    sparse_matrix* A = sparse_zero(n);
    double *X = zero_vector(n), *B = zero_vector(n), *B0, *AX;
    double res = 0, nb = 0;
    for (r = 0; r < m; r++)
        for (c = 0; c < m; c++) {
            i = r * m + c;
            if (r == 0 || c == 0 || r == m - 1 || c == m - 1) {
                insert_sparse(A, 1.0, i, i);
                continue;
            }
            insert_sparse(A, 4.0, i, i);
            if (r > 1) insert_sparse(A, -1.0, i, i - m);
            if (r < m - 2) insert_sparse(A, -1.0, i, i + m);
            if (c > 1) insert_sparse(A, -1.0, i, i - 1);
            if (c < m - 2) insert_sparse(A, -1.0, i, i + 1);
            B[i] = sin(0.37 * i) + 0.5;
        }
    B0 = clone_vector(B, n);
    PCG_default(A, NULL, NULL);
    PCG_set_AMG_preconditioner(A, 1, 4);
    it = pcg(X, B, n, 1e-10, 500);
    AX = sparse_mult(A, X);
    for (i = 0; i < n; i++) {
        res += (B0[i] - AX[i]) * (B0[i] - AX[i]);
        nb += B0[i] * B0[i];
    }
    printf("iterations %d residuum %.17g relres %.17g x0 %.17g\n", it, get_pcg_residuum(), sqrt(res / nb), X[n / 2]);
    free(X);
    free(B);
    free(B0);
    free(AX);
    free_sparse(A);
    return 0;
}
