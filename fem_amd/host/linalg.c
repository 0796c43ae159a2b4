/* Container, vector and device glue of the LinA drop-in.
 *
 * The row-list sparse_matrix keeps the reference's semantics where they
 * change values or patterns: sparse_zero seeds every row with (i, 0.0)
 * (linear_algebra.c:501-508), insert_sparse adds to an existing entry
 * (:1124-1133), delete_zeros removes zero entries and leaves (i, 0.0) in a row
 * that loses all of them (:793-805, remove_element_at :1161-1180).  The
 * products accumulate each row in column order like the reference loops
 * (:134-184).  ILU(0) and the triangular solves run through liblina_hip. */
#include <string.h>

#include "dropin.h"

/* ------------------------------------------------------------- errors */

void dropin_check(int rc, const char* what) {
    if (rc < 0) fprintf(stderr, "liblinalg_hip: %s failed: %s\n", what, lina_last_error());
}

/* --------------------------------------------------------- containers */

sparse_matrix* sparse_zero(int length) {
    int i;
    sparse_matrix* A = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    A->size = length;
    A->Len = (int*)malloc((length > 0 ? length : 1) * sizeof(int));
    A->Values = (double**)malloc((length > 0 ? length : 1) * sizeof(double*));
    A->Indices = (int**)malloc((length > 0 ? length : 1) * sizeof(int*));
    for (i = 0; i < length; i++) {
        A->Len[i] = 1;
        A->Values[i] = (double*)malloc(sizeof(double));
        A->Indices[i] = (int*)malloc(sizeof(int));
        A->Values[i][0] = 0.0;
        A->Indices[i][0] = i;
    }
    return A;
}

/* binary search in a sorted list: position of val (Flag 1) or of its
 * insertion point (Flag -1) */
int get_position(int* list, int size, int val, int* Flag) {
    int lo = 0, hi = size;
    *Flag = -1;
    if (size <= 0) return -1;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (list[mid] < val) lo = mid + 1;
        else hi = mid;
    }
    if (lo < size && list[lo] == val) *Flag = 1;
    return lo;
}

static void insert_at(sparse_matrix* A, double a, int i, int j, int pos) {
    int k, len = A->Len[i];
    A->Values[i] = (double*)realloc(A->Values[i], (len + 1) * sizeof(double));
    A->Indices[i] = (int*)realloc(A->Indices[i], (len + 1) * sizeof(int));
    for (k = len - 1; k >= pos; k--) {
        A->Values[i][k + 1] = A->Values[i][k];
        A->Indices[i][k + 1] = A->Indices[i][k];
    }
    A->Values[i][pos] = a;
    A->Indices[i][pos] = j;
    A->Len[i] = len + 1;
}

int insert_sparse(sparse_matrix* A, double a, int i, int j) {
    int flag;
    int pos = get_position(A->Indices[i], A->Len[i], j, &flag);
    if (pos < 0) pos = 0;
    if (flag == 1) A->Values[i][pos] += a;
    else insert_at(A, a, i, j, pos);
    return pos;
}

void set_sparse(sparse_matrix* A, double a, int i, int j) {
    int flag;
    int pos = get_position(A->Indices[i], A->Len[i], j, &flag);
    if (pos < 0) pos = 0;
    if (flag == 1) A->Values[i][pos] = a;
    else insert_at(A, a, i, j, pos);
}

double get_matrix_element(sparse_matrix* A, int i, int j) {
    int flag;
    int pos = get_position(A->Indices[i], A->Len[i], j, &flag);
    return flag == 1 ? A->Values[i][pos] : 0.0;
}

double get_diag(sparse_matrix* A, int i) { return get_matrix_element(A, i, i); }

void delete_zeros(sparse_matrix* A) {
    int i, j;
    for (i = 0; i < A->size; i++) {
        int len = A->Len[i], w = 0;
        if (len <= 1) continue;
        for (j = 0; j < len; j++)
            if (A->Values[i][j] != 0) {
                A->Values[i][w] = A->Values[i][j];
                A->Indices[i][w] = A->Indices[i][j];
                w++;
            }
        if (w == 0) {
            A->Values[i][0] = 0.0;
            A->Indices[i][0] = i;
            w = 1;
        }
        A->Len[i] = w;
    }
}

sparse_matrix* clone(sparse_matrix* A) {
    int i, n = A->size;
    sparse_matrix* R = (sparse_matrix*)malloc(sizeof(sparse_matrix));
    R->size = n;
    R->Len = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    R->Values = (double**)malloc((n > 0 ? n : 1) * sizeof(double*));
    R->Indices = (int**)malloc((n > 0 ? n : 1) * sizeof(int*));
    for (i = 0; i < n; i++) {
        int m = A->Len[i];
        R->Len[i] = m;
        R->Values[i] = (double*)malloc((m > 0 ? m : 1) * sizeof(double));
        R->Indices[i] = (int*)malloc((m > 0 ? m : 1) * sizeof(int));
        memcpy(R->Values[i], A->Values[i], m * sizeof(double));
        memcpy(R->Indices[i], A->Indices[i], m * sizeof(int));
    }
    return R;
}

void Free_sparse(sparse_matrix* A) {
    int i;
    for (i = 0; i < A->size; i++) {
        free(A->Values[i]);
        free(A->Indices[i]);
    }
    free(A->Len);
    free(A->Values);
    free(A->Indices);
}

/* ------------------------------------------------------ row products */

static double row_dot(sparse_matrix* A, int i, const double* x) {
    int j;
    double sum = 0;
    for (j = 0; j < A->Len[i]; j++) sum += A->Values[i][j] * x[A->Indices[i][j]];
    return sum;
}

double* sparse_mult(sparse_matrix* A, double* x) {
    int i;
    double* y = (double*)malloc((A->size > 0 ? A->size : 1) * sizeof(double));
    for (i = 0; i < A->size; i++) y[i] = row_dot(A, i, x);
    return y;
}

void linear_map(double* x, double fac, sparse_matrix* A, double* y) {
    int i;
    for (i = 0; i < A->size; i++) {
        double sum = row_dot(A, i, y);
        sum *= fac;
        x[i] += sum;
    }
}

void sparse_multiplication(sparse_matrix* A, double* x) {
    int i, n = A->size;
    double* y = zero_vector(n);
    for (i = 0; i < n; i++) {
        int j;
        for (j = 0; j < A->Len[i]; j++) y[i] += A->Values[i][j] * x[A->Indices[i][j]];
    }
    copy_vector_to(y, x, n);
    free(y);
}

double sparse_bilinear(double* y, sparse_matrix* A, double* x) {
    double* Ax = sparse_mult(A, x);
    double r = scalar(y, Ax, A->size);
    free(Ax);
    return r;
}

double matrix_residuum(sparse_matrix* A, double* Sol, double* b) {
    int n = A->size;
    double* R = clone_vector(b, n);
    double r;
    linear_map(R, -1., A, Sol);
    r = euklid_norm(R, n);
    free(R);
    return r;
}

/* ------------------------------------------------------------ vectors */

double scalar(double* x1, double* x2, int n) {
    int i;
    double res = 0;
    for (i = 0; i < n; i++) res = res + x1[i] * x2[i];
    return res;
}

double euklid_norm(double* x1, int n) { return sqrt(scalar(x1, x1, n)); }

void scalar_mult(double a, double* x, int n) {
    int i;
    for (i = 0; i < n; i++) x[i] *= a;
}

void vector_add(double* x, double* y, double factor, int n) {
    int i;
    for (i = 0; i < n; i++) x[i] += factor * y[i];
}

double* zero_vector(int n) {
    double* x = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    return x;
}

double* clone_vector(double* x, int n) {
    double* y = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    memcpy(y, x, n * sizeof(double));
    return y;
}

void copy_vector_to(double* x, double* y, int n) { memcpy(y, x, n * sizeof(double)); }

/* --------------------------------------------------------- CSR glue */

/* CSR copy; a row without entries (read_sparse can produce one) becomes the
 * explicit zero row (i, 0.0) of sparse_zero's convention, which the device
 * matrices require and which multiplies like an empty row */
dropin_csr dropin_to_csr(sparse_matrix* A) {
    dropin_csr c;
    int i, k = 0, nnz = 0;
    c.n = A->size;
    for (i = 0; i < A->size; i++) nnz += A->Len[i] > 0 ? A->Len[i] : 1;
    c.rp = (int*)malloc((c.n + 1) * sizeof(int));
    c.ci = (int*)malloc((nnz > 0 ? nnz : 1) * sizeof(int));
    c.v = (double*)malloc((nnz > 0 ? nnz : 1) * sizeof(double));
    c.rp[0] = 0;
    for (i = 0; i < A->size; i++) {
        if (A->Len[i] > 0) {
            memcpy(c.ci + k, A->Indices[i], A->Len[i] * sizeof(int));
            memcpy(c.v + k, A->Values[i], A->Len[i] * sizeof(double));
            k += A->Len[i];
        } else {
            c.ci[k] = i;
            c.v[k++] = 0.0;
        }
        c.rp[i + 1] = k;
    }
    return c;
}

void dropin_free_csr(dropin_csr* c) {
    free(c->rp);
    free(c->ci);
    free(c->v);
    c->rp = c->ci = NULL;
    c->v = NULL;
}

lina_matrix* dropin_device_matrix(sparse_matrix* A) {
    dropin_csr c = dropin_to_csr(A);
    lina_matrix* m = lina_matrix_create(c.n, c.rp, c.ci, c.v);
    if (!m) dropin_check(-1, "device matrix");
    dropin_free_csr(&c);
    return m;
}

/* ------------------------------------------------------ device glue */

double* dropin_dalloc(int n) {
    double* d = lina_dev_alloc(n > 0 ? n : 1);
    if (!d) {
        dropin_check(-1, "device allocation");
        return NULL;
    }
    dropin_check(lina_dev_fill(d, n > 0 ? n : 1, 0.0), "fill");
    return d;
}
void dropin_dfree(double* d) {
    if (d) lina_dev_free(d);
}
void dropin_h2d(double* d, const double* h, int n) { dropin_check(lina_dev_copy(d, h, n, 0), "copy"); }
void dropin_d2h(double* h, const double* d, int n) { dropin_check(lina_dev_copy(h, d, n, 1), "copy"); }
void dropin_d2d(double* dst, const double* src, int n) { dropin_check(lina_dev_copy(dst, src, n, 2), "copy"); }
double dropin_dot(const double* a, const double* b, int n) { return lina_dev_dot(a, b, n); }

void dropin_stage_init(dropin_stage* s, int n) {
    s->n = n;
    s->hx = NULL;
    s->hy = NULL;
}
void dropin_stage_free(dropin_stage* s) {
    free(s->hx);
    free(s->hy);
}
static void stage_ready(dropin_stage* s) {
    if (!s->hx) {
        s->hx = (double*)calloc(s->n > 0 ? s->n : 1, sizeof(double));
        s->hy = (double*)calloc(s->n > 0 ? s->n : 1, sizeof(double));
    }
}

void dropin_op_apply(const dropin_op* o, dropin_stage* s, const double* dx, double* dy) {
    if (o->dev) {
        dropin_check(lina_matrix_spmv_device(o->dev, dx, dy), "SpMV");
        return;
    }
    stage_ready(s);
    dropin_d2h(s->hx, dx, s->n);
    o->host(s->n, s->hx, s->hy);
    dropin_h2d(dy, s->hy, s->n);
}

int dropin_prec_apply(const dropin_prec* p, dropin_stage* s, const double* dr, double* dz) {
    int k;
    if (p->amg) {
        k = lina_amg_solve_device(p->amg, dr, dz);
        dropin_check(k, "AMG_solve");
        return k;
    }
    if (p->ilu) {
        k = lina_ilu_apply_device(p->ilu, dr, dz);
        dropin_check(k, "ILU solve");
        return 1;
    }
    stage_ready(s);
    dropin_d2h(s->hx, dr, s->n);
    dropin_d2h(s->hy, dz, s->n);
    k = p->host(s->n, s->hx, s->hy);
    dropin_h2d(dz, s->hy, s->n);
    return k;
}

/* ------------------------------------------------------------ ILU(0) */

static void fill_rows(sparse_matrix* T, int n, const int* rp, const int* ci, const double* v) {
    int i;
    for (i = 0; i < n; i++) {
        int m = rp[i + 1] - rp[i];
        free(T->Values[i]);
        free(T->Indices[i]);
        T->Len[i] = m;
        T->Values[i] = (double*)malloc((m > 0 ? m : 1) * sizeof(double));
        T->Indices[i] = (int*)malloc((m > 0 ? m : 1) * sizeof(int));
        memcpy(T->Values[i], v + rp[i], m * sizeof(double));
        memcpy(T->Indices[i], ci + rp[i], m * sizeof(int));
    }
}

int incomplete_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U) {
    int n = A->size, ok, nl, nu;
    dropin_csr c = dropin_to_csr(A);
    lina_ilu* h = lina_ilu_create(n, c.rp, c.ci, c.v, 0);
    dropin_free_csr(&c);
    if (!h) {
        dropin_check(-1, "incomplete_LU_factorization");
        return 0;
    }
    ok = lina_ilu_ok(h);
    nl = lina_ilu_factor_nnz(h, 'L');
    nu = lina_ilu_factor_nnz(h, 'U');
    if (ok && nl >= 0 && nu >= 0) {
        int* rp = (int*)malloc((n + 1) * sizeof(int));
        int* ci = (int*)malloc(((nl > nu ? nl : nu) + 1) * sizeof(int));
        double* v = (double*)malloc(((nl > nu ? nl : nu) + 1) * sizeof(double));
        lina_ilu_factor_copy(h, 'L', rp, ci, v);
        fill_rows(L, n, rp, ci, v);
        lina_ilu_factor_copy(h, 'U', rp, ci, v);
        fill_rows(U, n, rp, ci, v);
        free(rp);
        free(ci);
        free(v);
    }
    lina_ilu_destroy(h);
    return ok;
}

/* L x = b in place (diagonal last in each row of L) or U x = b (diagonal
 * first in each row of U): a level-scheduled triangular solve on the GPU */
static void triang_invert(sparse_matrix* T, double* x, int lower) {
    int i, n = T->size;
    dropin_csr t = dropin_to_csr(T);
    int* irp = (int*)malloc((n + 1) * sizeof(int));
    int* ici = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    double* iv = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    lina_ilu* h;
    double* w = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    for (i = 0; i <= n; i++) irp[i] = i;
    for (i = 0; i < n; i++) {
        ici[i] = i;
        iv[i] = 1.0;
    }
    h = lower ? lina_ilu_create_factors(n, t.rp, t.ci, t.v, irp, ici, iv)
              : lina_ilu_create_factors(n, irp, ici, iv, t.rp, t.ci, t.v);
    if (!h) {
        dropin_check(-1, lower ? "L_triang_invert" : "U_triang_invert");
    } else {
        dropin_check(lina_ilu_apply(h, x, w), "triangular solve");
        memcpy(x, w, n * sizeof(double));
        lina_ilu_destroy(h);
    }
    free(w);
    free(irp);
    free(ici);
    free(iv);
    dropin_free_csr(&t);
}

void L_triang_invert(sparse_matrix* L, double* x) { triang_invert(L, x, 1); }
void U_triang_invert(sparse_matrix* U, double* x) { triang_invert(U, x, 0); }

/* ------------------------------------------------------- transpose */

/* get_transpose (linear_algebra.c:1290-1300): rows of a new sparse_zero
 * matrix of new_size rows receive A's entries through insert_sparse, A's rows
 * in order (entries whose column has no row there are skipped: the reference
 * writes past its row arrays for them) */
sparse_matrix* get_transpose(sparse_matrix* A, int new_size) {
    sparse_matrix* T = sparse_zero(new_size);
    int i, r;
    for (i = 0; i < A->size; i++)
        for (r = 0; r < A->Len[i]; r++)
            if (A->Indices[i][r] < new_size) insert_sparse(T, A->Values[i][r], A->Indices[i][r], i);
    return T;
}
