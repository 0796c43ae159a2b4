/* Bordered systems and the primal-dual active-set method of the LinA drop-in.
 *
 * Behaviour kept from LinA (linear_algebra.c:4109-4156, :4867-5042): the
 * bordered matrices hold the same entries in the same rows, with the values
 * that LinA's insert_sparse sequence produces, and active_set_solver hands
 * the caller's solver the same sequence of systems and right-hand sides.
 *
 * The construction here is this library's own: a bordered matrix is built
 * in one pass from the base rows and a list of added (row, column, value)
 * triplets (stable-sorted by row and column, merged into each row; a
 * triplet on an existing entry is added to it, later triplets on the same
 * position are added in the order they were listed), instead of cloning,
 * enlarging and inserting entry by entry; a block shift of a matrix is a
 * renumbering of its row table. */
#include <string.h>

#include "dropin.h"

/* ------------------------------------------------------------ utilities */

int* zero_int_list(int n) { return (int*)calloc(n > 0 ? n : 1, sizeof(int)); }

void copy_vector_content(double* x, double* y, int x_pos, int y_pos, int size) {
    if (size > 0) memmove(y + y_pos, x + x_pos, (size_t)size * sizeof(double));
}

static void seed_row(sparse_matrix* A, int r) {
    A->Len[r] = 1;
    A->Indices[r] = (int*)malloc(sizeof(int));
    A->Values[r] = (double*)malloc(sizeof(double));
    A->Indices[r][0] = r;
    A->Values[r][0] = 0.0;
}

/* Row r of the result is row r - row_shift of A (columns + col_shift) where
 * that row exists, else a seeded row (r, 0.0).  Rows of A that fall outside
 * the result are released. */
static void shift_rows(sparse_matrix* A, int new_size, int row_shift, int col_shift) {
    const int old = A->size;
    const int cap = new_size > 0 ? new_size : 1;
    int* len = (int*)malloc(cap * sizeof(int));
    int** ind = (int**)malloc(cap * sizeof(int*));
    double** val = (double**)malloc(cap * sizeof(double*));
    int r, k;
    for (r = 0; r < old; r++) {
        const int to = r + row_shift;
        if (to < 0 || to >= new_size) {
            free(A->Indices[r]);
            free(A->Values[r]);
            continue;
        }
        len[to] = A->Len[r];
        ind[to] = A->Indices[r];
        val[to] = A->Values[r];
        if (col_shift)
            for (k = 0; k < len[to]; k++) ind[to][k] += col_shift;
    }
    free(A->Len);
    free(A->Indices);
    free(A->Values);
    A->Len = len;
    A->Indices = ind;
    A->Values = val;
    A->size = new_size;
    for (r = 0; r < new_size; r++)
        if (r - row_shift < 0 || r - row_shift >= old) seed_row(A, r);
}

/* linear_algebra.c:4109-4137: A's rows move down by equ_index blocks of
 * mesh_size rows and its columns right by var_index blocks; a block keeps
 * its internal numbering, so both moves are plain shifts.  The matrix gets
 * new_size rows; rows that receive nothing are seeded (i, 0.0). */
void enlarge_matrix(sparse_matrix* A, int mesh_size, int new_size, int equ_index, int var_index) {
    shift_rows(A, new_size, equ_index * mesh_size, var_index * mesh_size);
}

/* linear_algebra.c:4139-4156: add_size seeded rows at the end */
void simple_enlarge_matrix(sparse_matrix* A, int add_size) { shift_rows(A, A->size + add_size, 0, 0); }

/* ------------------------------------------------------------- triplets */

typedef struct {
    int row, col, seq;
    double val;
} triplet;

typedef struct {
    triplet* t;
    int n, cap;
} triplets;

static void tr_push(triplets* L, int row, int col, double val) {
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 64;
        L->t = (triplet*)realloc(L->t, (size_t)L->cap * sizeof(triplet));
    }
    L->t[L->n].row = row;
    L->t[L->n].col = col;
    L->t[L->n].val = val;
    L->t[L->n].seq = L->n;
    L->n++;
}

static int tr_order(const void* a, const void* b) {
    const triplet* x = (const triplet*)a;
    const triplet* y = (const triplet*)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* A new matrix of `size` rows: row r starts as row r of A (r < A->size) or
 * as the seed (r, 0.0), then receives the triplets of row r. */
static sparse_matrix* build_bordered(sparse_matrix* A, int size, triplets* L) {
    sparse_matrix* R = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    const int cap = size > 0 ? size : 1;
    int r, p = 0;
    R->size = size;
    R->Len = (int*)malloc(cap * sizeof(int));
    R->Indices = (int**)malloc(cap * sizeof(int*));
    R->Values = (double**)malloc(cap * sizeof(double*));
    qsort(L->t, (size_t)L->n, sizeof(triplet), tr_order);
    while (p < L->n && L->t[p].row < 0) p++;
    for (r = 0; r < size; r++) {
        static const double zero = 0.0;
        const int seed_col = r;
        const int* bi = r < A->size ? A->Indices[r] : &seed_col;
        const double* bv = r < A->size ? A->Values[r] : &zero;
        const int bl = r < A->size ? A->Len[r] : 1;
        int q = p, a = 0, w = 0;
        while (q < L->n && L->t[q].row == r) q++;
        R->Indices[r] = (int*)malloc((size_t)(bl + q - p > 0 ? bl + q - p : 1) * sizeof(int));
        R->Values[r] = (double*)malloc((size_t)(bl + q - p > 0 ? bl + q - p : 1) * sizeof(double));
        while (a < bl || p < q) {
            int col;
            double v;
            if (p >= q || (a < bl && bi[a] <= L->t[p].col)) {
                col = bi[a];
                v = bv[a++];
            } else {
                col = L->t[p].col;
                v = L->t[p++].val;
            }
            while (p < q && L->t[p].col == col) v += L->t[p++].val;
            R->Indices[r][w] = col;
            R->Values[r][w] = v;
            w++;
        }
        R->Len[r] = w;
    }
    return R;
}

/* --------------------------------------------------- bordered systems */

/* linear_algebra.c:4867-4884: [A c_k; c_k^T 0] over the active constraints
 * k, where c_k is column act[k] of C as LinA reads it through
 * get_transpose(C, C->size): that transpose starts from seeded rows, so
 * c_k always holds the entry (act[k], 0.0 + C[act[k]][act[k]]), and entries
 * of C in columns >= C->size do not take part. */
sparse_matrix* add_constraints_default(sparse_matrix* A, sparse_matrix* C, int* act, int nact) {
    const int N = A->size, m = C->size;
    int* cnt = (int*)calloc((size_t)m + 1, sizeof(int));
    int* rows;
    double* vals;
    triplets L = {0};
    int i, j, k;
    sparse_matrix* R;
    /* the columns of C, rows ascending, each with its seed merged in */
    for (i = 0; i < m; i++) {
        int seeded = 0;
        for (j = 0; j < C->Len[i]; j++)
            if (C->Indices[i][j] < m) cnt[C->Indices[i][j] + 1]++;
        for (j = 0; j < C->Len[i]; j++) seeded |= C->Indices[i][j] == i;
        if (!seeded) cnt[i + 1]++;
    }
    for (i = 0; i < m; i++) cnt[i + 1] += cnt[i];
    rows = (int*)malloc((size_t)(cnt[m] > 0 ? cnt[m] : 1) * sizeof(int));
    vals = (double*)malloc((size_t)(cnt[m] > 0 ? cnt[m] : 1) * sizeof(double));
    {
        int* fill = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
        memcpy(fill, cnt, (size_t)m * sizeof(int));
        for (i = 0; i < m; i++) {
            int seeded = 0;
            for (j = 0; j < C->Len[i]; j++) seeded |= C->Indices[i][j] == i;
            if (!seeded) { /* column i gets the seed at row i, in row order */
                rows[fill[i]] = i;
                vals[fill[i]++] = 0.0;
            }
            for (j = 0; j < C->Len[i]; j++) {
                const int c = C->Indices[i][j];
                if (c >= m) continue;
                rows[fill[c]] = i;
                vals[fill[c]++] = c == i ? 0.0 + C->Values[i][j] : C->Values[i][j];
            }
        }
        free(fill);
    }
    for (k = 0; k < nact; k++) {
        const int c = act[k];
        int p;
        for (p = cnt[c]; p < cnt[c + 1]; p++) {
            tr_push(&L, rows[p], N + k, vals[p]);
            tr_push(&L, N + k, rows[p], vals[p]);
        }
    }
    R = build_bordered(A, N + nact, &L);
    free(L.t);
    free(rows);
    free(vals);
    free(cnt);
    return R;
}

/* linear_algebra.c:4886-4918 (diagonal constraints on the second field):
 * constraint k is row act[k] of C; its entries border rows C->size + j of
 * A in column N + k and fill the new row N + k. */
sparse_matrix* add_constraints_flux(sparse_matrix* A, sparse_matrix* C, int* act, int nact) {
    const int N = A->size, off = C->size;
    triplets L = {0};
    int j, k;
    sparse_matrix* R;
    for (k = 0; k < nact; k++) {
        const int c = act[k];
        for (j = 0; j < C->Len[c]; j++) {
            if (off + C->Indices[c][j] < N + nact) tr_push(&L, off + C->Indices[c][j], N + k, C->Values[c][j]);
            tr_push(&L, N + k, C->Indices[c][j], C->Values[c][j]);
        }
    }
    R = build_bordered(A, N + nact, &L);
    free(L.t);
    return R;
}

void free_ass_info(ass_info** info) {
    if (info == NULL || *info == NULL) return;
    free((*info)->Active);
    free((*info)->Signs);
    free((*info)->Constraint);
    free((*info)->Iter_history);
    free(*info);
    *info = NULL;
}

/* ----------------------------------------------------- active-set method */

static ass_info* ass_state(ass_info** info, int m) {
    ass_info* s = *info;
    if (s == NULL) s = *info = (ass_info*)calloc(1, sizeof(ass_info));
    if (s->Active == NULL) s->Active = zero_int_list(m);
    if (s->Signs == NULL) s->Signs = zero_vector(m);
    if (s->Constraint == NULL) s->Constraint = zero_vector(m);
    return s;
}

/* One pass: solve the bordered system of the current active set, then
 * (1) release constraints whose multiplier has the sign of `definiteness`
 * (+1 minimum, -1 maximum) and (2) activate every constraint B x violates
 * by more than 1e-10, at the bound it violates.  Step (1) runs over the
 * leading entries of the set while their index stays below the set's live
 * size, which shrinks with every release -- LinA's loop bound
 * (linear_algebra.c:5001-5007), kept because it decides which multipliers
 * are looked at and so the systems of the following passes.  Returns 1
 * when nothing changed. */
static int ass_pass(ass_info* s, sparse_matrix* A, sparse_matrix* B, double* X, double* rhs, int* act,
                    double* sgn, double* L, double* U, int (*Solver)(sparse_matrix*, double*, double*),
                    double definiteness, int mode) {
    const double tol = 1e-10;
    const int n = A->size, m = B->size;
    int i, k, nact = 0, released = 0, settled = 1;
    sparse_matrix* M;
    double* Y;
    for (i = 0; i < m; i++)
        if (s->Active[i]) {
            act[nact] = i;
            sgn[nact] = s->Signs[i];
            rhs[n + nact] = s->Constraint[i];
            nact++;
        }
    s->active_size = nact;
    M = mode == ASS_FLUX ? add_constraints_flux(A, B, act, nact) : add_constraints_default(A, B, act, nact);
    Solver(M, X, rhs);
    free_sparse(M);
    Y = sparse_mult(B, X);
    for (k = 0; k < nact - released; k++)
        if (sgn[k] * X[n + k] * definiteness > 0) {
            s->Active[act[k]] = 0;
            released++;
            settled = 0;
        }
    s->active_size = nact - released;
    for (i = 0; i < m; i++) {
        const int low = Y[i] < L[i] - tol, high = Y[i] > U[i] + tol;
        if (!low && !high) continue;
        if (!s->Active[i]) s->active_size++;
        s->Active[i] = 1;
        if (low) {
            s->Signs[i] = 1.0;
            s->Constraint[i] = L[i];
        }
        if (high) {
            s->Signs[i] = -1.0;
            s->Constraint[i] = U[i];
        }
        settled = 0;
    }
    free(Y);
    return settled;
}

/* linear_algebra.c:4933-5042.  X (the solver's iterate, multipliers after
 * the n unknowns) persists across passes; the right-hand side is b followed
 * by the active bounds.  Sol receives the first n entries. */
int active_set_solver(sparse_matrix* A, double* Sol, double* b, sparse_matrix* B, double* L, double* U,
                      int (*Solver)(sparse_matrix* A, double* X, double* b), int max_iter, double definiteness,
                      int mode, ass_info** solver_info) {
    const int n = A->size, m = B->size;
    ass_info* s = ass_state(solver_info, m);
    double* X = zero_vector(n + m);
    double* rhs = zero_vector(n + m);
    int* act = zero_int_list(m);
    double* sgn = zero_vector(m);
    int passes = 0, settled;
    memcpy(X, Sol, (size_t)n * sizeof(double));
    memcpy(rhs, b, (size_t)n * sizeof(double));
    do {
        settled = ass_pass(s, A, B, X, rhs, act, sgn, L, U, Solver, definiteness, mode);
        passes++;
    } while (!settled && passes < max_iter);
    memcpy(Sol, X, (size_t)n * sizeof(double));
    if (passes >= max_iter) fprintf(stderr, "active_set_solver: active set still changing after %d passes\n", passes);
    s->Iter_history = (int*)realloc(s->Iter_history, (size_t)(s->calls + 1) * sizeof(int));
    s->Iter_history[s->calls++] = passes;
    free(X);
    free(rhs);
    free(act);
    free(sgn);
    return passes;
}
