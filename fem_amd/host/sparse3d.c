/* Third-order sparse tensors and nodal rank-4 tensors of the LinA drop-in
 * (host C): the containers FEM2D's assembly uses for the strain-gradient
 * tensor B_abc and the nodal moduli (LinA/linear_algebra.c:91-133,
 * 510-760, 1343-1380).  Values and orders of the results are those of the
 * cited functions: row a of a sparse_matrix3D holds its (b, c, value)
 * triples in insertion order; contractions sum in that order. */
#include <string.h>

#include "dropin.h"

sparse_matrix3D* sparse_zero3D(int n) {                             /* :510-525: rows empty */
    sparse_matrix3D* B = (sparse_matrix3D*)malloc(sizeof(sparse_matrix3D));
    const size_t m = (size_t)(n > 0 ? n : 1);
    B->size = n;
    B->Values = (double**)calloc(m, sizeof(double*));
    B->Indices1 = (int**)calloc(m, sizeof(int*));
    B->Indices2 = (int**)calloc(m, sizeof(int*));
    B->Len = (int*)calloc(m, sizeof(int));
    return B;
}

/* a triple appended to row a */
static void push3(sparse_matrix3D* B, int a, int b, int c, double v) {
    const int m = ++B->Len[a];
    B->Values[a] = (double*)realloc(B->Values[a], (size_t)m * sizeof(double));
    B->Indices1[a] = (int*)realloc(B->Indices1[a], (size_t)m * sizeof(int));
    B->Indices2[a] = (int*)realloc(B->Indices2[a], (size_t)m * sizeof(int));
    B->Values[a][m - 1] = v;
    B->Indices1[a][m - 1] = b;
    B->Indices2[a][m - 1] = c;
}

/* rows left without triples get the single (a, a, 0.0) */
static void seed_empty3(sparse_matrix3D* B) {
    int a;
    for (a = 0; a < B->size; a++)
        if (B->Len[a] == 0) push3(B, a, a, a, 0.0);
}

sparse_matrix* sparse3D_vB(sparse_matrix3D* B, double* X, int new_size) { /* :527-535: X_a B_abc */
    int a, j;
    sparse_matrix* R = sparse_zero(new_size);
    for (a = 0; a < B->size; a++)
        for (j = 0; j < B->Len[a]; j++) insert_sparse(R, X[a] * B->Values[a][j], B->Indices1[a][j], B->Indices2[a][j]);
    return R;
}

sparse_matrix* sparse3D_Bv(sparse_matrix3D* B, double* X) {         /* :551-561: B_abc X_c */
    int a, j;
    sparse_matrix* R = sparse_zero(B->size);
    for (a = 0; a < B->size; a++)
        for (j = 0; j < B->Len[a]; j++) insert_sparse(R, B->Values[a][j] * X[B->Indices2[a][j]], a, B->Indices1[a][j]);
    return R;
}

/* :537-549: (B_abc X_c + B_abc X_b) / 2, both halves inserted triple by triple */
sparse_matrix* sparse3D_Bv_symmetric(sparse_matrix3D* B, double* X) {
    int a, j;
    sparse_matrix* R = sparse_zero(B->size);
    for (a = 0; a < B->size; a++)
        for (j = 0; j < B->Len[a]; j++) {
            insert_sparse(R, B->Values[a][j] * X[B->Indices2[a][j]] / 2., a, B->Indices1[a][j]);
            insert_sparse(R, B->Values[a][j] * X[B->Indices1[a][j]] / 2., a, B->Indices2[a][j]);
        }
    return R;
}

double* sparse3D_Bvv(sparse_matrix3D* B, double* X1, double* X2) {  /* :573-588: B_abc X1_b X2_c */
    int a, j;
    double* r = zero_vector(B->size);
    for (a = 0; a < B->size; a++) {
        double s = 0;
        for (j = 0; j < B->Len[a]; j++) s += B->Values[a][j] * X1[B->Indices1[a][j]] * X2[B->Indices2[a][j]];
        r[a] = s;
    }
    return r;
}

double* sparse3D_vvB(sparse_matrix3D* B, double* X1, double* X2, int size) { /* :590-601: X1_a X2_b B_abc */
    int a, j;
    double* r = zero_vector(size);
    for (a = 0; a < B->size; a++)
        for (j = 0; j < B->Len[a]; j++) r[B->Indices2[a][j]] += B->Values[a][j] * X1[a] * X2[B->Indices1[a][j]];
    return r;
}

double* sparse3D_bilinear(sparse_matrix3D* B, double* X1, double* X2) { /* :679-686 */
    int a, j;
    double* r = zero_vector(B->size);
    for (a = 0; a < B->size; a++)
        for (j = 0; j < B->Len[a]; j++) r[a] += B->Values[a][j] * X1[B->Indices1[a][j]] * X2[B->Indices2[a][j]];
    return r;
}

void sparse3D_scalar_mult(sparse_matrix3D* B, double a) {           /* :640-647 */
    int i, j;
    for (i = 0; i < B->size; i++)
        for (j = 0; j < B->Len[i]; j++) B->Values[i][j] *= a;
}

/* :649-677: A += B triple by triple: added to A's first triple with the same
 * (b, c), else appended */
void sparse3D_add(sparse_matrix3D* A, sparse_matrix3D* B) {
    int a, k, j;
    for (a = 0; a < A->size; a++)
        for (k = 0; k < B->Len[a]; k++) {
            const int b = B->Indices1[a][k], c = B->Indices2[a][k];
            for (j = 0; j < A->Len[a]; j++)
                if (A->Indices1[a][j] == b && A->Indices2[a][j] == c) break;
            if (j < A->Len[a]) A->Values[a][j] += B->Values[a][k];
            else push3(A, a, b, c, B->Values[a][k]);
        }
}

/* :703-741: the triples of each row with equal (b, c) summed into the first
 * of them, rows keeping first-occurrence order */
void sparse3D_merge(sparse_matrix3D* B) {
    int a, j, k;
    for (a = 0; a < B->size; a++) {
        int w = 0;
        if (B->Len[a] <= 1) continue;
        for (j = 0; j < B->Len[a]; j++) {
            const int b = B->Indices1[a][j], c = B->Indices2[a][j];
            for (k = 0; k < w; k++)
                if (B->Indices1[a][k] == b && B->Indices2[a][k] == c) break;
            if (k < w) {
                B->Values[a][k] += B->Values[a][j];
            } else {
                B->Indices1[a][w] = b;
                B->Indices2[a][w] = c;
                B->Values[a][w++] = B->Values[a][j];
            }
        }
        B->Len[a] = w;
    }
}

/* :603-638: C_cba = B_abc over the nonzero triples (row c of the result
 * lists them in B's order), then the empty rows seeded */
sparse_matrix3D* sparse3D_transpose13(sparse_matrix3D* B, int new_size) {
    int a, l;
    sparse_matrix3D* R = sparse_zero3D(new_size);
    for (a = 0; a < B->size; a++)
        for (l = 0; l < B->Len[a]; l++) {
            const int c = B->Indices2[a][l];
            if (B->Values[a][l] == 0) continue;
            if (c >= new_size) {
                fprintf(stderr, "sparse3D_transpose13: index %d beyond %d\n", c, new_size);
                continue;
            }
            push3(R, c, B->Indices1[a][l], a, B->Values[a][l]);
        }
    seed_empty3(R);
    return R;
}

/* :1343-1380: the same over every triple, equal triples merged, then the
 * empty rows seeded */
sparse_matrix3D* get_transpose3D_13(sparse_matrix3D* B, int new_size) {
    int a, l;
    sparse_matrix3D* R = sparse_zero3D(new_size);
    for (a = 0; a < B->size; a++)
        for (l = 0; l < B->Len[a]; l++) {
            const int c = B->Indices2[a][l];
            if (c >= new_size) {
                fprintf(stderr, "get_transpose3D_13: index %d beyond %d\n", c, new_size);
                continue;
            }
            push3(R, c, B->Indices1[a][l], a, B->Values[a][l]);
        }
    sparse3D_merge(R);
    seed_empty3(R);
    return R;
}

/* ------------------------------------------------ nodal rank-4 tensors
 *
 * A field of dim^4 tensors over mesh_size nodes, component (a,b,c,d) of node
 * m at mesh_size * (((a dim + b) dim + c) dim + d) + m. */

static size_t t4(int a, int b, int c, int d, int dim, int mesh_size, int m) {
    return (size_t)mesh_size * (size_t)(((a * dim + b) * dim + c) * dim + d) + (size_t)m;
}

/* :91-110: (A B)_abcd = sum_ef A_abef B_efcd per node */
double* rank4_product(double* A, double* B, int dim, int mesh_size) {
    int m, a, b, c, d, e, f;
    double* R = zero_vector(dim * dim * dim * dim * mesh_size);
    for (m = 0; m < mesh_size; m++)
        for (a = 0; a < dim; a++)
            for (b = 0; b < dim; b++)
                for (c = 0; c < dim; c++)
                    for (d = 0; d < dim; d++)
                        for (e = 0; e < dim; e++)
                            for (f = 0; f < dim; f++)
                                R[t4(a, b, c, d, dim, mesh_size, m)] +=
                                    A[t4(a, b, e, f, dim, mesh_size, m)] * B[t4(e, f, c, d, dim, mesh_size, m)];
    return R;
}

/* :112-126: (A x B x C x D)_ijkl = A_i B_j C_k D_l per node (vectors of dim
 * components, component i of node m at mesh_size * i + m) */
double* rank4_tensor_product(double* A, double* B, double* C, double* D, int dim, int mesh_size) {
    int m, i, j, k, l;
    double* R = zero_vector(dim * dim * dim * dim * mesh_size);
    for (m = 0; m < mesh_size; m++)
        for (i = 0; i < dim; i++)
            for (j = 0; j < dim; j++)
                for (k = 0; k < dim; k++)
                    for (l = 0; l < dim; l++)
                        R[t4(i, j, k, l, dim, mesh_size, m)] =
                            A[mesh_size * i + m] * B[mesh_size * j + m] * C[mesh_size * k + m] * D[mesh_size * l + m];
    return R;
}

void add_rank4_tensor_product(double* X, double* A, double* B, double* C, double* D, int dim, int mesh_size) {
    double* Y = rank4_tensor_product(A, B, C, D, dim, mesh_size); /* :128-132 */
    vector_add(X, Y, 1., dim * dim * dim * dim * mesh_size);
    free(Y);
}
