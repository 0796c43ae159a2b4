/* Row-list matrix algebra of the LinA drop-in (host C, caller-owned
 * matrices): products, sums, scalings, row/column sums, restrictions and
 * single-entry edits, with the value and pattern semantics of the
 * functions of LinA/linear_algebra.c cited at each definition. */
#include <string.h>

#include "dropin.h"

/* ------------------------------------------------------------- products */

/* A dense-indexed accumulator for one output row: slot[c] >= 0 while column
 * c holds a value in the row being formed. */
typedef struct {
    int* slot;
    int* cols;
    double* vals;
    int len, width;
} row_acc;

static void acc_init(row_acc* r, int width) {
    int c;
    r->width = width;
    r->slot = (int*)malloc((size_t)(width > 0 ? width : 1) * sizeof(int));
    for (c = 0; c < width; c++) r->slot[c] = -1;
    r->cols = (int*)malloc((size_t)(width > 0 ? width : 1) * sizeof(int));
    r->vals = (double*)malloc((size_t)(width > 0 ? width : 1) * sizeof(double));
    r->len = 0;
}

static void acc_free(row_acc* r) {
    free(r->slot);
    free(r->cols);
    free(r->vals);
}

/* insert_sparse semantics: the first value at a column is stored as given,
 * later ones are added to it */
static void acc_add(row_acc* r, int c, double v) {
    if (r->slot[c] < 0) {
        r->slot[c] = r->len;
        r->cols[r->len] = c;
        r->vals[r->len++] = v;
    } else {
        r->vals[r->slot[c]] += v;
    }
}

static int cmp_int(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* the row's entries in column order into row i of P (replacing its
 * arrays), then delete_zeros' rule (linear_algebra.c:793-805): a row of
 * several entries loses its exact zeros and keeps (i, 0.0) if nothing is
 * left */
static void acc_emit(row_acc* r, sparse_matrix* P, int i) {
    int k, w = 0, n = r->len;
    int* order = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    memcpy(order, r->cols, (size_t)n * sizeof(int));
    qsort(order, (size_t)n, sizeof(int), cmp_int);
    free(P->Indices[i]);
    free(P->Values[i]);
    P->Indices[i] = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    P->Values[i] = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (k = 0; k < n; k++) {
        const int c = order[k];
        const double v = r->vals[r->slot[c]];
        r->slot[c] = -1;
        if (n > 1 && v == 0) continue;
        P->Indices[i][w] = c;
        P->Values[i][w++] = v;
    }
    if (w == 0) {
        P->Indices[i][0] = i;
        P->Values[i][0] = 0.0;
        w = 1;
    }
    P->Len[i] = w;
    r->len = 0;
    free(order);
}

/* linear_algebra.c:186-222: P = A B.  Row i of P starts as sparse_zero's
 * (i, 0.0); for each nonzero A(i,k) in row order and each nonzero B(k,j) in
 * row order A(i,k) B(k,j) is added at (i, j); then delete_zeros. */
sparse_matrix* sparse_product(sparse_matrix* A, sparse_matrix* B) {
    const int n = A->size;
    int i, p, q, width = n;
    row_acc r;
    sparse_matrix* P = sparse_zero(n);
    for (i = 0; i < B->size; i++)
        for (q = 0; q < B->Len[i]; q++)
            if (B->Indices[i][q] + 1 > width) width = B->Indices[i][q] + 1;
    acc_init(&r, width);
    for (i = 0; i < n; i++) {
        acc_add(&r, i, 0.0);
        for (p = 0; p < A->Len[i]; p++) {
            const double a = A->Values[i][p];
            const int k = A->Indices[i][p];
            if (a == 0) continue;
            for (q = 0; q < B->Len[k]; q++) {
                const double b = B->Values[k][q];
                if (b != 0) acc_add(&r, B->Indices[k][q], a * b);
            }
        }
        acc_emit(&r, P, i);
    }
    acc_free(&r);
    return P;
}
