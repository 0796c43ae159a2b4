/* Row-list matrix algebra of the LinA drop-in (host C, caller-owned
 * matrices): products, sums, scalings, row/column sums, restrictions and
 * single-entry edits, with the value and pattern semantics of the
 * functions of LinA/linear_algebra.c cited at each definition. */
#include <math.h>
#include <stdarg.h>
#include <string.h>

#include "dropin.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

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

/* ------------------------------------------------- sums and scalings */

void scalar_sparse_mult(double a, sparse_matrix* A) {               /* :318-323 */
    int i, j;
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) A->Values[i][j] *= a;
}

/* :325-334: A += factor B (every entry of B's first A->size rows added
 * through insert_sparse) */
void sparse_add(sparse_matrix* A, sparse_matrix* B, double factor) {
    int i, j;
    if (!B) return;
    for (i = 0; i < A->size; i++)
        for (j = 0; j < B->Len[i]; j++) insert_sparse(A, factor * B->Values[i][j], i, B->Indices[i][j]);
}

/* :1209-1219: the nonzero entries of B added into A at offset (i_o, j_o) */
void add_subsparse_to_sparse(sparse_matrix* A, sparse_matrix* B, int i_o, int j_o) {
    int i, k;
    for (i = 0; i < B->size; i++)
        for (k = 0; k < B->Len[i]; k++)
            if (B->Values[i][k] != 0) insert_sparse(A, B->Values[i][k], i + i_o, B->Indices[i][k] + j_o);
}

double* sparse_row_sum(sparse_matrix* A) {                          /* :3271-3278 */
    int i, j;
    double* r = zero_vector(A->size);
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) r[i] += A->Values[i][j];
    return r;
}

double* get_col_sums(sparse_matrix* A, int col_num) {               /* :3464-3471 */
    int i, j;
    double* r = zero_vector(col_num);
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) r[A->Indices[i][j]] += A->Values[i][j];
    return r;
}

/* :455-464: each row with a nonzero sum scaled to sum Row_Sums[i] (1 when
 * Row_Sums is NULL) */
void normalize_row_sum(sparse_matrix* A, double* Row_Sums) {
    int i, j;
    for (i = 0; i < A->size; i++) {
        double s = 0, d = Row_Sums ? Row_Sums[i] : 1.;
        for (j = 0; j < A->Len[i]; j++) s += A->Values[i][j];
        if (s != 0)
            for (j = 0; j < A->Len[i]; j++) A->Values[i][j] *= d / s;
    }
}

/* -------------------------------------------------------- builders */

sparse_matrix* sparse_diagonal(double* d, int n) {                  /* :484-499 */
    int i;
    sparse_matrix* R = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    R->size = n;
    R->Len = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    R->Indices = (int**)malloc((size_t)(n > 0 ? n : 1) * sizeof(int*));
    R->Values = (double**)malloc((size_t)(n > 0 ? n : 1) * sizeof(double*));
    for (i = 0; i < n; i++) {
        R->Len[i] = 1;
        R->Indices[i] = (int*)malloc(sizeof(int));
        R->Values[i] = (double*)malloc(sizeof(double));
        R->Indices[i][0] = i;
        R->Values[i][0] = d[i];
    }
    return R;
}

sparse_matrix* sparse_identity(int length) {                        /* :774-781 */
    sparse_matrix* R = sparse_zero(length);
    int i;
    for (i = 0; i < length; i++) R->Values[i][0] = 1;
    return R;
}

/* :4192-4207: the second-difference matrix of u'' = f on n points with
 * u = 0 at both ends (identity boundary rows, the boundary coupling left
 * out), F scaled by -h^2 inside and zeroed at the ends */
sparse_matrix* get_1D_Laplace(double* F, int n) {
    const double h = 1. / (n - 1);
    int i;
    sparse_matrix* A = sparse_zero(n);
    for (i = 1; i < n - 1; i++) {
        insert_sparse(A, 2., i, i);
        if (i != 1) insert_sparse(A, -1., i, i - 1);
        if (i != n - 2) insert_sparse(A, -1., i, i + 1);
        F[i] *= -h * h;
    }
    insert_sparse(A, 1., 0, 0);
    insert_sparse(A, 1., n - 1, n - 1);
    F[0] = 0;
    F[n - 1] = 0;
    return A;
}

/* ---------------------------------------------------- single entries */

/* :1552-1560: position of column col in row `row`; -1 (with a message)
 * when absent (LinA exits) */
int find_position_in_row(sparse_matrix* A, int row, int col) {
    int flag;
    const int pos = get_position(A->Indices[row], A->Len[row], col, &flag);
    if (flag != 1) {
        fprintf(stderr, "find_position_in_row: (%d,%d) is not in the pattern\n", row, col);
        return -1;
    }
    return pos;
}

/* :1161-1179: the entry at pos leaves row i; a row that would become empty
 * keeps the single entry (i, 0.0) */
void remove_element_at(sparse_matrix* A, int i, int pos) {
    const int len = A->Len[i];
    if (len > 1) {
        memmove(A->Values[i] + pos, A->Values[i] + pos + 1, (size_t)(len - 1 - pos) * sizeof(double));
        memmove(A->Indices[i] + pos, A->Indices[i] + pos + 1, (size_t)(len - 1 - pos) * sizeof(int));
        A->Len[i] = len - 1;
    } else {
        A->Values[i][0] = 0;
        A->Indices[i][0] = i;
    }
}

/* :1905-1915: the last zero entry of a row (one per row and call) leaves it */
void remove_zero(sparse_matrix* A, int i) {
    int j, k = -1;
    for (j = 0; j < A->Len[i]; j++)
        if (A->Values[i][j] == 0) k = j;
    if (k >= 0) remove_element_at(A, i, k);
}

void remove_zeros(sparse_matrix* A) {
    int i;
    for (i = 0; i < A->size; i++) remove_zero(A, i);
}

void reset_row(sparse_matrix* A, int i) {                           /* :1245-1253: row i = (i, 0.0) */
    free(A->Values[i]);
    free(A->Indices[i]);
    A->Values[i] = (double*)malloc(sizeof(double));
    A->Indices[i] = (int*)malloc(sizeof(int));
    A->Values[i][0] = 0;
    A->Indices[i][0] = i;
    A->Len[i] = 1;
}

/* -------------------------------------------------------- restriction */

/* :3242-3269: the rows not listed in List (ascending) as a new matrix;
 * index_map[i] = new row of row i, -1 for listed rows */
sparse_matrix* restrict_matrix_rows(sparse_matrix* A, int* List, int list_size, int* index_map) {
    sparse_matrix* R = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    const int m = A->size - list_size;
    int i, next = 0, k = 0;
    R->size = m;
    R->Len = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
    R->Indices = (int**)malloc((size_t)(m > 0 ? m : 1) * sizeof(int*));
    R->Values = (double**)malloc((size_t)(m > 0 ? m : 1) * sizeof(double*));
    for (i = 0; i < A->size; i++) {
        const int len = A->Len[i];
        if (next < list_size && List[next] == i) {
            index_map[i] = -1;
            next++;
            continue;
        }
        index_map[i] = k;
        R->Len[k] = len;
        R->Indices[k] = (int*)malloc((size_t)(len > 0 ? len : 1) * sizeof(int));
        R->Values[k] = (double*)malloc((size_t)(len > 0 ? len : 1) * sizeof(double));
        memcpy(R->Indices[k], A->Indices[i], (size_t)len * sizeof(int));
        memcpy(R->Values[k], A->Values[i], (size_t)len * sizeof(double));
        k++;
    }
    return R;
}

/* :3280-3299: columns renumbered by index_map in place; entries in
 * unmapped columns leave the row and move to the right-hand side, Res_i =
 * -sum a_ij Sol_j over them (in row order).  A row reduced to its (i, 0.0)
 * seed in an unmapped column i is left there (LinA loops forever on it). */
double* restrict_matrix_cols(sparse_matrix* A, double* Sol, int* index_map) {
    int i, j;
    double* r = zero_vector(A->size);
    for (i = 0; i < A->size; i++) {
        j = 0;
        while (j < A->Len[i]) {
            const int c = A->Indices[i][j];
            if (index_map[c] >= 0) {
                A->Indices[i][j] = index_map[c];
                j++;
            } else {
                r[i] -= A->Values[i][j] * Sol[c];
                if (A->Len[i] == 1 && c == i && A->Values[i][j] == 0) break;
                remove_element_at(A, i, j);
            }
        }
    }
    return r;
}

/* -------------------------------------------------- formats and maps */

/* :3182-3212: compressed columns (UMFPACK's layout) of the nonzero entries
 * of A's first col_num columns, rows ascending in each column */
void convert_sparse_to_UMFPACK(sparse_matrix* A, int col_num, int** Col_start, int** Indices, double** Values,
                               int* Size) {
    int i, j, k = 0, nnz = 0;
    int* cnt = (int*)calloc((size_t)col_num + 1, sizeof(int));
    for (i = 0; i < A->size; i++)
        for (j = 0; j < A->Len[i]; j++) {
            const int c = A->Indices[i][j];
            if (c < col_num && A->Values[i][j] != 0) cnt[c + 1]++;
        }
    for (i = 0; i < col_num; i++) cnt[i + 1] += cnt[i];
    nnz = cnt[col_num];
    *Col_start = (int*)malloc((size_t)(col_num + 1) * sizeof(int));
    *Indices = (int*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    *Values = (double*)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    memcpy(*Col_start, cnt, (size_t)(col_num + 1) * sizeof(int));
    for (i = 0; i < A->size; i++) /* rows in order: ascending inside each column */
        for (j = 0; j < A->Len[i]; j++) {
            const int c = A->Indices[i][j];
            if (c < col_num && A->Values[i][j] != 0) {
                k = cnt[c]++;
                (*Indices)[k] = i;
                (*Values)[k] = A->Values[i][j];
            }
        }
    *Size = nnz;
    free(cnt);
}

void free_UMFPACK_matrix_info(umfpack_matrix_info* A) {             /* :5515-5519 */
    free(A->Ind);
    free(A->Col);
    free(A->Val);
}

void free_sparse_index_map(map_sparse** Map) {                      /* :696-701 */
    free((*Map)->Indices);
    Free_sparse((*Map)->Pattern);
    free((*Map)->Pattern);
    free(*Map);
    *Map = NULL;
}

sparse_general* new_sparse_general(int rank, int* Dim) {            /* :5224-5231 */
    sparse_general* R = (sparse_general*)malloc(sizeof(sparse_general));
    R->Elements = NULL;
    R->size = 0;
    R->rank = rank;
    R->Dimension = clone_list(Dim, rank);
    return R;
}

/* :5241-5256: an element of `rank` indices (the trailing int arguments) */
void sparse_general_append(sparse_general* A, double val, int rank, ...) {
    va_list ap;
    int i;
    matrix_element* e;
    if (A->rank != rank) {
        fprintf(stderr, "sparse_general_append: element of rank %d in a rank-%d matrix\n", rank, A->rank);
        return;
    }
    A->Elements = (matrix_element*)realloc(A->Elements, (size_t)(A->size + 1) * sizeof(matrix_element));
    e = &A->Elements[A->size++];
    e->val = val;
    e->Ind = zero_int_list(rank);
    va_start(ap, rank);
    for (i = 0; i < rank; i++) e->Ind[i] = va_arg(ap, int);
    va_end(ap);
}

/* ---------------------------------------------- fixed-pattern products */

/* set_product_pattern (:5078-5116) keeps, for each entry (i, J) of A B
 * (sparse_product's pattern), the addresses of the value pairs A(i,k),
 * B(k,J) = BT(J,k) it sums, found by walking A's row i and BT's row J
 * together (get_row_pairs, :5053-5076; the walk stops when BT's row runs
 * out); fast_sparse_product (:5134-5172) then refills P's values from the
 * current values at those addresses, pair by pair.  BT is the transpose of
 * B, or B itself when B_transposed is NULL (B symmetric). */
static double** row_pairs(sparse_matrix* A, sparse_matrix* BT, int i, int J, int* count) {
    double** out = (double**)malloc((size_t)(2 * A->Len[i] > 0 ? 2 * A->Len[i] : 1) * sizeof(double*));
    int l, k = 0, c = 0;
    const int bl = BT->Len[J];
    for (l = 0; l < A->Len[i]; l++) {
        const int r = A->Indices[i][l];
        while (k < bl && BT->Indices[J][k] < r) k++;
        if (k == bl) break;
        if (BT->Indices[J][k] == r) {
            out[c++] = &A->Values[i][l];
            out[c++] = &BT->Values[J][k];
        }
    }
    *count = c;
    return (double**)realloc(out, (size_t)(c > 0 ? c : 1) * sizeof(double*));
}

product_pattern set_product_pattern(sparse_matrix* A, sparse_matrix* B, sparse_matrix* B_transposed,
                                    sparse_matrix** AB) {
    product_pattern pp;
    sparse_matrix *F, *BT = B_transposed ? B_transposed : B;
    int i, j;
    if (B == NULL && B_transposed != NULL) {
        int maxc = 0;
        sparse_matrix* H;
        for (i = 0; i < B_transposed->size; i++)
            for (j = 0; j < B_transposed->Len[i]; j++)
                if (B_transposed->Indices[i][j] > maxc) maxc = B_transposed->Indices[i][j];
        H = get_transpose(B_transposed, maxc + 1);
        F = sparse_product(A, H);
        Free_sparse(H);
        free(H);
    } else {
        F = sparse_product(A, B);
    }
    pp.Size_0 = F->size;
    pp.Pairs = (double****)malloc((size_t)(F->size > 0 ? F->size : 1) * sizeof(double***));
    pp.Indices = (int**)malloc((size_t)(F->size > 0 ? F->size : 1) * sizeof(int*));
    pp.Size_2 = (int**)malloc((size_t)(F->size > 0 ? F->size : 1) * sizeof(int*));
    pp.Size_1 = (int*)malloc((size_t)(F->size > 0 ? F->size : 1) * sizeof(int));
    for (i = 0; i < F->size; i++) {
        const int m = F->Len[i];
        pp.Size_1[i] = m;
        pp.Pairs[i] = (double***)malloc((size_t)(m > 0 ? m : 1) * sizeof(double**));
        pp.Indices[i] = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
        pp.Size_2[i] = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
        for (j = 0; j < m; j++) {
            pp.Indices[i][j] = F->Indices[i][j];
            pp.Pairs[i][j] = row_pairs(A, BT, i, F->Indices[i][j], &pp.Size_2[i][j]);
        }
    }
    if (AB) *AB = F;
    else {
        Free_sparse(F);
        free(F);
    }
    return pp;
}

void fast_sparse_product(sparse_matrix* P, sparse_matrix* A, sparse_matrix* BT, product_pattern* Pattern) {
    int i, j, k;
    (void)A;
    (void)BT;
    for (i = 0; i < Pattern->Size_0; i++)
        for (j = 0; j < Pattern->Size_1[i]; j++) {
            double s = 0;
            for (k = 0; k < Pattern->Size_2[i][j]; k += 2) s += *Pattern->Pairs[i][j][k] * *Pattern->Pairs[i][j][k + 1];
            P->Indices[i][j] = Pattern->Indices[i][j];
            P->Values[i][j] = s;
        }
}

/* ------------------------------------------------------------ test matrix */

/* :244-254 with copy_sparse_content :1473-1495: A B written over B, or
 * over A */
static void replace_rows(sparse_matrix* dst, sparse_matrix* src) {
    int i;
    for (i = 0; i < dst->size; i++) {
        free(dst->Values[i]);
        free(dst->Indices[i]);
    }
    free(dst->Len);
    free(dst->Values);
    free(dst->Indices);
    *dst = *src;
    free(src);
}

/* :3885-3911: diag((i+1)/n) with the middle eigenvalue negated, rotated by
 * random_rot Givens rotations G (i, j, angle from rand() after srand(0),
 * G = I with (i,i), (j,j) += cos a - 1, (i,j) -= sin a, (j,i) += sin a):
 * M = G M G^T each; the eigenvalues go to the file "vector" as in LinA */
sparse_matrix* get_test_matrix(int n, int random_rot) {
    int k;
    double* eig = zero_vector(n);
    sparse_matrix* M;
    for (k = 0; k < n; k++) eig[k] = (double)(k + 1) / n;
    eig[n / 2] = -eig[n / 2];
    print_vector(eig, n);
    srand(0);
    M = sparse_diagonal(eig, n);
    for (k = 0; k < random_rot; k++) {
        const int i = rand() % n, j = rand() % n;
        if (i != j) {
            const double a = (double)2. * M_PI * rand() / RAND_MAX;
            sparse_matrix* G = sparse_identity(n);
            sparse_matrix *GT, *T;
            insert_sparse(G, cos(a) - 1., i, i);
            insert_sparse(G, cos(a) - 1., j, j);
            insert_sparse(G, -sin(a), i, j);
            insert_sparse(G, sin(a), j, i);
            GT = get_transpose(G, n);
            T = sparse_product(M, GT);
            replace_rows(M, T);
            T = sparse_product(G, M);
            replace_rows(M, T);
            Free_sparse(G);
            free(G);
            Free_sparse(GT);
            free(GT);
        }
    }
    free(eig);
    return M;
}

/* :918-935 writes to a fixed directory of the author's machine; here the
 * file "sparse_matrix" in the working directory, one row per line of
 * "(j|v)" pairs */
void print_sparse(sparse_matrix* A) {
    int i, j;
    FILE* f;
    if (!A) {
        printf("Matrix = NULL");
        return;
    }
    f = fopen("sparse_matrix", "w");
    if (!f) return;
    for (i = 0; i < A->size; i++) {
        const int n = A->Len[i] - 1;
        for (j = 0; j < n; j++) fprintf(f, "(%d|%.15e)  ", A->Indices[i][j], A->Values[i][j]);
        fprintf(f, "(%d|%.15e)\n", A->Indices[i][n], A->Values[i][n]);
    }
    fclose(f);
}

/* :2836-2855: row i becomes the old row Map[i] (the row arrays are moved,
 * not copied) */
void sparse_row_permutation(sparse_matrix* A, int* Map) {
    const int n = A->size;
    int i;
    double** v = (double**)malloc((size_t)(n > 0 ? n : 1) * sizeof(double*));
    int** c = (int**)malloc((size_t)(n > 0 ? n : 1) * sizeof(int*));
    int* l = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (i = 0; i < n; i++) {
        v[i] = A->Values[i];
        c[i] = A->Indices[i];
        l[i] = A->Len[i];
    }
    for (i = 0; i < n; i++) {
        A->Values[i] = v[Map[i]];
        A->Indices[i] = c[Map[i]];
        A->Len[i] = l[Map[i]];
    }
    free(v);
    free(c);
    free(l);
}

/* :2413-2429: max over rows of sum_{j != i} |a_ij| / |a_ii|; a row without a
 * diagonal entry divides by the last diagonal seen, as the reference does */
double get_diag_dominance(sparse_matrix* A) {
    int i, j;
    double worst = 0, diag = 0;
    for (i = 0; i < A->size; i++) {
        double off = 0;
        for (j = 0; j < A->Len[i]; j++) {
            if (A->Indices[i][j] != i) off += fabs(A->Values[i][j]);
            else diag = fabs(A->Values[i][j]);
        }
        off /= diag;
        if (off > worst) worst = off;
    }
    return worst;
}

/* :4312-4339: the five-point operator 2(a+b), -a (rows i +- 1), -b (columns
 * j +- 1) on the interior of a dom_size^2 grid, then 1 added on every
 * boundary node once per boundary side it lies on (corners get 2) */
sparse_matrix* Lap_2D_Dirichlet(double a, double b, int dom_size) {
    const int m = dom_size;
    int i, j;
    sparse_matrix* A = sparse_zero(m * m);
    for (i = 1; i < m - 1; i++)
        for (j = 1; j < m - 1; j++) {
            const int k = i * m + j;
            insert_sparse(A, 2. * (a + b), k, k);
            insert_sparse(A, -a, k, k + m);
            insert_sparse(A, -a, k, k - m);
            insert_sparse(A, -b, k, k + 1);
            insert_sparse(A, -b, k, k - 1);
        }
    for (i = 0; i < m; i++) {
        insert_sparse(A, 1., i * m, i * m);
        insert_sparse(A, 1., i * m + m - 1, i * m + m - 1);
        insert_sparse(A, 1., i, i);
        insert_sparse(A, 1., (m - 1) * m + i, (m - 1) * m + i);
    }
    return A;
}
