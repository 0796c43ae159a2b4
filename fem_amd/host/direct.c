/* Dense direct factorisations of the LinA drop-in (host C):
 * sparse_LU_factorization (linear_algebra.c:3040-3062, LinA's coarsest-
 * level factors), complete_LU_factorization (:3122-3139), the dense/row-list
 * converters (:3064-3090) and diagonal_preconditioner (:1984-1997).
 *
 * The factorisation is the bordering form of Doolittle LU without pivoting
 * (L unit lower, U upper): with the leading s x s factors known, row s of L
 * solves l U = a_s,0:s (a forward substitution with U^T), column s of U
 * solves L u = a_0:s,s, and u_ss = a_ss - l.u.  Each substitution
 * accumulates in index order and each pivot is subtracted in that order, so
 * the factors are those of LinA's LU_factorzation (:2909-2942) bit for bit
 * -- the property the drop-in's callers compare (L/U of a coarse level).
 * These are setup-time routines on caller-owned host matrices: the GPU
 * solvers factor their own coarsest level on the device. */
#include <string.h>

#include "dropin.h"

/* L and U (n x n, row-major) of the dense matrix M by bordering */
static void border_lu(const double* M, int n, double* L, double* U) {
    int s, i, j;
    double* u = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double* l = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    memset(L, 0, (size_t)n * n * sizeof(double));
    memset(U, 0, (size_t)n * n * sizeof(double));
    if (n > 0) {
        L[0] = 1.0;
        U[0] = M[0];
    }
    for (s = 1; s < n; s++) {
        double d;
        for (i = 0; i < s; i++) { /* L u = M[0:s, s] */
            double t = M[(size_t)i * n + s];
            for (j = 0; j < i; j++) t -= L[(size_t)i * n + j] * u[j];
            u[i] = t / L[(size_t)i * n + i];
        }
        for (i = 0; i < s; i++) { /* U^T l = M[s, 0:s] */
            double t = M[(size_t)s * n + i];
            for (j = 0; j < i; j++) t -= U[(size_t)j * n + i] * l[j];
            l[i] = t / U[(size_t)i * n + i];
        }
        d = M[(size_t)s * n + s];
        for (i = 0; i < s; i++) {
            U[(size_t)i * n + s] = u[i];
            L[(size_t)s * n + i] = l[i];
            d -= u[i] * l[i];
        }
        U[(size_t)s * n + s] = d;
        L[(size_t)s * n + s] = 1.0;
    }
    free(u);
    free(l);
}

/* a row-list matrix as a dense row-major array with `cols` columns */
static double* to_dense(sparse_matrix* A, int cols) {
    int i, j;
    double* D = (double*)calloc((size_t)A->size * (cols > 0 ? cols : 1), sizeof(double));
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) {
            const int c = A->Indices[i][j];
            if (c < 0 || c >= cols) {
                fprintf(stderr, "liblinalg_hip: row-list to dense: column %d outside [0, %d)\n", c, cols);
                continue;
            }
            D[(size_t)i * cols + c] = A->Values[i][j];
        }
    return D;
}

/* linear_algebra.c:3082-3090: the nonzero entries of the dense rows x cols
 * array M_A added into A (a sparse_zero matrix) */
void convert_array_to_sparse(double** M_A, sparse_matrix* A, int col_size, int row_size) {
    int i, j;
    for (i = 0; i < col_size; i++)
        for (j = 0; j < row_size; j++)
            if (M_A[i][j] != 0) insert_sparse(A, M_A[i][j], i, j);
}

/* linear_algebra.c:3040-3062: rows holding more than one nonzero value form
 * a dense system (in their original order) that is factored; their entries
 * of L and U enter the caller's sparse_zero L and U at the original indices,
 * every other row gets L = 1 and U = its diagonal */
void sparse_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U) {
    const int n = A->size;
    int* map = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int* rows = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    int i, j, N = 0;
    double *M, *Lf, *Uf;
    for (i = 0; i < n; i++) {
        int nz = 0;
        for (j = 0; j < A->Len[i]; j++) nz += A->Values[i][j] != 0;
        map[i] = nz > 1 ? N : -1;
        if (nz > 1) rows[N++] = i;
    }
    M = (double*)calloc((size_t)N * N + 1, sizeof(double));
    Lf = (double*)malloc(((size_t)N * N + 1) * sizeof(double));
    Uf = (double*)malloc(((size_t)N * N + 1) * sizeof(double));
    for (i = 0; i < N; i++) {
        const int r = rows[i];
        for (j = 0; j < A->Len[r]; j++) {
            const int c = map[A->Indices[r][j]];
            if (c >= 0) M[(size_t)i * N + c] = A->Values[r][j];
        }
    }
    border_lu(M, N, Lf, Uf);
    for (i = 0; i < N; i++)
        for (j = 0; j < N; j++) {
            if (Lf[(size_t)i * N + j] != 0) insert_sparse(L, Lf[(size_t)i * N + j], rows[i], rows[j]);
            if (Uf[(size_t)i * N + j] != 0) insert_sparse(U, Uf[(size_t)i * N + j], rows[i], rows[j]);
        }
    for (i = 0; i < n; i++)
        if (map[i] < 0) {
            insert_sparse(L, 1.0, i, i);
            insert_sparse(U, get_diag(A, i), i, i);
        }
    free(M);
    free(Lf);
    free(Uf);
    free(map);
    free(rows);
}

/* linear_algebra.c:3122-3139: the dense LU of the whole matrix into the
 * caller's sparse_zero L and U (nonzero entries only) */
void complete_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U) {
    const int n = A->size;
    int i, j;
    double* M = to_dense(A, n);
    double* Lf = (double*)malloc(((size_t)n * n + 1) * sizeof(double));
    double* Uf = (double*)malloc(((size_t)n * n + 1) * sizeof(double));
    border_lu(M, n, Lf, Uf);
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) {
            if (Lf[(size_t)i * n + j] != 0) insert_sparse(L, Lf[(size_t)i * n + j], i, j);
            if (Uf[(size_t)i * n + j] != 0) insert_sparse(U, Uf[(size_t)i * n + j], i, j);
        }
    free(M);
    free(Lf);
    free(Uf);
}

/* linear_algebra.c:1984-1997: L = diag(A), U = I, added into the caller's
 * sparse_zero L and U, stopping at the first zero diagonal */
void diagonal_preconditioner(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U) {
    int i;
    for (i = 0; i < A->size; i++) {
        const double a = get_matrix_element(A, i, i);
        if (a == 0) {
            fprintf(stderr, "diagonal_preconditioner: zero diagonal in row %d\n", i);
            break;
        }
        insert_sparse(L, a, i, i);
        insert_sparse(U, 1, i, i);
    }
}
