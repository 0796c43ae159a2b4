/*
 * lina_oracle.c -- CPU restatement of the LinA solver path, the parity
 * oracle for liblina_hip.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py through oracle/bind.py; never linked into the product.
 *
 * Each routine restates the reference routine cited above it, operation by
 * operation (compiled with -ffp-contract=off: no fused multiply-adds), on
 * CSR arrays instead of the reference's row-lists.  The reference's growth of
 * rows by insert_sparse (linear_algebra.c:1124-1159) is replaced by a dense
 * row accumulator that applies the same additions in the same order.  The
 * red-black tree that orders coarsening candidates (red_black_tree.c) is
 * replaced by a binary max-heap on (importance, insertion number): equal keys
 * go to the right of existing ones in the tree, so the rightmost node taken
 * by RBLargest is the most recently inserted among the largest keys.
 *
 * Pinning: tests/test_oracle_cpu.py checks every routine here against the
 * compiled reference (oracle/_ref) and against the committed golden vectors
 * in tests/golden/, plus closed-form checks (dense solves, CG in exact
 * arithmetic on tiny systems, invariants of the AMG hierarchy).
 */
#include "lina_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { UNDECIDED = 0, COARSE = 1, FINE = 2, ISOLATED = 3 };
enum { AGGRESSIVE = 1, STANDARD = 2 };

/* ------------------------------------------------------------ containers */

typedef struct {
    int* a;
    int n, cap;
} ivec;
typedef struct {
    double* a;
    int n, cap;
} dvec;

static void ipush(ivec* v, int x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->a = (int*)realloc(v->a, v->cap * sizeof(int));
    }
    v->a[v->n++] = x;
}
static void dpush(dvec* v, double x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 16;
        v->a = (double*)realloc(v->a, v->cap * sizeof(double));
    }
    v->a[v->n++] = x;
}

typedef struct {
    int n, ncols;
    int* rp;
    ivec ci;
    dvec v;
} Mat;

static void mat_init(Mat* M, int n, int ncols) {
    memset(M, 0, sizeof(*M));
    M->n = n;
    M->ncols = ncols;
    M->rp = (int*)calloc(n + 1, sizeof(int));
}
static void mat_free(Mat* M) {
    free(M->rp);
    free(M->ci.a);
    free(M->v.a);
    memset(M, 0, sizeof(*M));
}
static int mat_nnz(const Mat* M) { return M->rp[M->n]; }
static void mat_from(Mat* M, int n, const int* rp, const int* ci, const double* v) {
    int k;
    mat_init(M, n, n);
    memcpy(M->rp, rp, (n + 1) * sizeof(int));
    for (k = 0; k < rp[n]; k++) {
        ipush(&M->ci, ci[k]);
        dpush(&M->v, v[k]);
    }
}

/* dense row accumulator: the value/presence of one row being built */
typedef struct {
    double* val;
    char* has;
    ivec touched;
    int width;
} Acc;

static void acc_init(Acc* A, int width) {
    A->val = (double*)calloc(width > 0 ? width : 1, sizeof(double));
    A->has = (char*)calloc(width > 0 ? width : 1, 1);
    A->touched.a = NULL;
    A->touched.n = A->touched.cap = 0;
    A->width = width;
}
static void acc_free(Acc* A) {
    free(A->val);
    free(A->has);
    free(A->touched.a);
}
/* insert_sparse: add to an existing entry, create a new one otherwise */
static void acc_add(Acc* A, int j, double a) {
    if (A->has[j]) A->val[j] += a;
    else {
        A->has[j] = 1;
        A->val[j] = a;
        ipush(&A->touched, j);
    }
}
/* sparse_zero's seed (j, 0.0) */
static void acc_seed(Acc* A, int j) {
    if (!A->has[j]) {
        A->has[j] = 1;
        A->val[j] = 0.0;
        ipush(&A->touched, j);
    }
}
/* remove_element (linear_algebra.c:1181-1185) */
static void acc_remove(Acc* A, int j) {
    if (A->has[j]) {
        A->has[j] = 0;
        A->val[j] = 0.0;
    }
}
static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}
/* move the live entries, sorted by column, into (c, v); reset */
static void acc_take(Acc* A, ivec* c, dvec* v) {
    int t, last = -1;
    c->n = 0;
    v->n = 0;
    qsort(A->touched.a, A->touched.n, sizeof(int), cmp_int);
    for (t = 0; t < A->touched.n; t++) {
        int j = A->touched.a[t];
        if (j == last) continue;
        last = j;
        if (A->has[j]) {
            ipush(c, j);
            dpush(v, A->val[j]);
        }
        A->has[j] = 0;
        A->val[j] = 0.0;
    }
    A->touched.n = 0;
}

/* delete_zeros (linear_algebra.c:793-805) on one row seeded by sparse_zero:
 * rows with more than one entry lose their exact zeros; an all-zero row ends
 * as the single entry (row, 0.0) */
static void row_delete_zeros(int row, ivec* c, dvec* v) {
    int k, w = 0;
    if (c->n <= 1) return;
    for (k = 0; k < c->n; k++)
        if (v->a[k] != 0.0) {
            c->a[w] = c->a[k];
            v->a[w] = v->a[k];
            w++;
        }
    if (w == 0) {
        c->a[0] = row;
        v->a[0] = 0.0;
        w = 1;
    }
    c->n = w;
    v->n = w;
}

static void mat_append_row(Mat* M, int i, const ivec* c, const dvec* v) {
    int k;
    for (k = 0; k < c->n; k++) {
        ipush(&M->ci, c->a[k]);
        dpush(&M->v, v->a[k]);
    }
    M->rp[i + 1] = M->ci.n;
}

/* get_diag (linear_algebra.c:3434-3439) */
static double diag_of(const Mat* A, int i) {
    int p;
    for (p = A->rp[i]; p < A->rp[i + 1]; p++)
        if (A->ci.a[p] == i) return A->v.a[p];
    return 0.0;
}

/* sparse_product (linear_algebra.c:186-222): C = A*B, row i seeded with
 * (i,0.0), products added in A-row then B-row order, then delete_zeros */
static void sparse_product(const Mat* A, const Mat* B, int b_ncols, Mat* C) {
    int i, p, q, width = b_ncols > A->n ? b_ncols : A->n;
    Acc acc;
    ivec c = {0};
    dvec v = {0};
    acc_init(&acc, width);
    mat_init(C, A->n, b_ncols);
    for (i = 0; i < A->n; i++) {
        acc_seed(&acc, i);
        for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
            double a = A->v.a[p];
            int k = A->ci.a[p];
            if (a == 0) continue;
            for (q = B->rp[k]; q < B->rp[k + 1]; q++) {
                double b = B->v.a[q];
                if (b != 0) acc_add(&acc, B->ci.a[q], a * b);
            }
        }
        acc_take(&acc, &c, &v);
        row_delete_zeros(i, &c, &v);
        mat_append_row(C, i, &c, &v);
    }
    acc_free(&acc);
    free(c.a);
    free(v.a);
}

/* get_restricted_matrix (Multigrid.c:475-484) = R*(A*P), square_matrix_clean
 * (linear_algebra.c:3100-3120) */
static int galerkin(const Mat* A, const Mat* R, const Mat* P, Mat* out) {
    Mat AP, RAP;
    int i, p, n;
    sparse_product(A, P, P->ncols, &AP);
    sparse_product(R, &AP, P->ncols, &RAP);
    mat_free(&AP);
    n = RAP.n;
    mat_init(out, n, n);
    for (i = 0; i < n; i++) {
        for (p = RAP.rp[i]; p < RAP.rp[i + 1]; p++) {
            int j = RAP.ci.a[p];
            double x = RAP.v.a[p];
            if (j >= n) {
                if (x == 0) continue;
                mat_free(&RAP);
                return 0;
            }
            if (x == 0 && j != i) continue;
            ipush(&out->ci, j);
            dpush(&out->v, x);
        }
        out->rp[i + 1] = out->ci.n;
    }
    mat_free(&RAP);
    return 1;
}

/* ----------------------------------------------------- coarsening heap */

typedef struct {
    double key;
    long seq;
    int node;
} hent;
typedef struct {
    hent* h;
    int n, cap;
    int* pos; /* heap index of a node, -1 when absent (Node_list[I]==NULL) */
    long seq;
} Heap;

static int h_above(const hent* a, const hent* b) { return a->key > b->key || (a->key == b->key && a->seq > b->seq); }
static void h_swap(Heap* H, int i, int j) {
    hent t = H->h[i];
    H->h[i] = H->h[j];
    H->h[j] = t;
    H->pos[H->h[i].node] = i;
    H->pos[H->h[j].node] = j;
}
static void h_up(Heap* H, int i) {
    while (i > 0) {
        int p = (i - 1) / 2;
        if (!h_above(&H->h[i], &H->h[p])) break;
        h_swap(H, i, p);
        i = p;
    }
}
static void h_down(Heap* H, int i) {
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < H->n && h_above(&H->h[l], &H->h[m])) m = l;
        if (r < H->n && h_above(&H->h[r], &H->h[m])) m = r;
        if (m == i) break;
        h_swap(H, i, m);
        i = m;
    }
}
static void h_init(Heap* H, int nodes) {
    int i;
    H->n = 0;
    H->cap = 64;
    H->h = (hent*)malloc(H->cap * sizeof(hent));
    H->pos = (int*)malloc((nodes > 0 ? nodes : 1) * sizeof(int));
    for (i = 0; i < nodes; i++) H->pos[i] = -1;
    H->seq = 0;
}
static void h_free(Heap* H) {
    free(H->h);
    free(H->pos);
}
static void h_insert(Heap* H, int node, double key) {
    if (H->n == H->cap) {
        H->cap *= 2;
        H->h = (hent*)realloc(H->h, H->cap * sizeof(hent));
    }
    H->h[H->n].key = key;
    H->h[H->n].seq = H->seq++;
    H->h[H->n].node = node;
    H->pos[node] = H->n;
    H->n++;
    h_up(H, H->n - 1);
}
static void h_remove(Heap* H, int node) {
    int i = H->pos[node];
    if (i < 0) return;
    H->n--;
    if (i != H->n) {
        H->h[i] = H->h[H->n];
        H->pos[H->h[i].node] = i;
        h_down(H, i);
        h_up(H, i);
    }
    H->pos[node] = -1;
}

/* ------------------------------------------------------------ coarsening */

typedef struct {
    int* ptr;
    int* idx;
} Lists;
static void lists_free(Lists* L) {
    free(L->ptr);
    free(L->idx);
}

/* get_strong_connection_sets (Multigrid.c:240-258) with get_strong_connected
 * (:156-184) and get_max_coupled (:86-97) */
static void strong_sets(const Mat* A, const int* levels, double eps_minus, double eps_plus, Lists* S, Lists* T) {
    int n = A->n, i, p, k;
    ivec idx = {0};
    int* cnt;
    S->ptr = (int*)calloc(n + 1, sizeof(int));
    for (i = 0; i < n; i++) {
        double mn = 0;
        int mi = -1;
        for (p = A->rp[i]; p < A->rp[i + 1]; p++)
            if (A->ci.a[p] != i && A->v.a[p] < mn) {
                mn = A->v.a[p];
                mi = p;
            }
        if (mi >= 0) {
            double lo = -eps_minus * fabs(A->v.a[mi]);
            double hi = eps_plus * fabs(A->v.a[mi]);
            for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
                int J = A->ci.a[p];
                double a;
                if (J == i) continue;
                a = levels[i] == levels[J] ? A->v.a[p] : 0;
                if (a < lo || a > hi) ipush(&idx, J);
            }
        }
        S->ptr[i + 1] = idx.n;
    }
    S->idx = idx.a;
    T->ptr = (int*)calloc(n + 1, sizeof(int));
    T->idx = (int*)malloc((idx.n > 0 ? idx.n : 1) * sizeof(int));
    for (k = 0; k < idx.n; k++) T->ptr[idx.a[k] + 1]++;
    for (i = 0; i < n; i++) T->ptr[i + 1] += T->ptr[i];
    cnt = (int*)malloc((n + 1) * sizeof(int));
    memcpy(cnt, T->ptr, (n + 1) * sizeof(int));
    for (i = 0; i < n; i++)
        for (p = S->ptr[i]; p < S->ptr[i + 1]; p++) T->idx[cnt[S->idx[p]]++] = i;
    free(cnt);
}

/* AMG_get_lambda (Multigrid.c:364-375) */
static double lambda_of(const int* label, const Lists* T, int i, int size) {
    int p, undec = 0, fine = 0;
    for (p = T->ptr[i]; p < T->ptr[i + 1]; p++) {
        int J = T->idx[p];
        if (label[J] == UNDECIDED) undec++;
        if (label[J] == FINE) fine++;
    }
    return (double)i / size + undec + 2 * fine;
}

/* AMG_coarsener (Multigrid.c:377-473); `levels` is indexed with this
 * matrix's own row numbers as in the reference */
static void coarsener(const Mat* A, int* label, double* imp, const Lists* S, const Lists* T, Heap* H,
                      const int* levels, int size, int seed) {
    int level = levels[seed], next = seed, p, q, i;
    do {
        label[next] = COARSE;
        for (p = A->rp[next]; p < A->rp[next + 1]; p++) {
            int I = A->ci.a[p];
            if (label[I] == UNDECIDED && A->v.a[p] != 0. && H->pos[I] < 0) h_insert(H, I, imp[I]);
        }
        for (q = T->ptr[next]; q < T->ptr[next + 1]; q++) {
            int I = T->idx[q];
            if (H->pos[I] >= 0) {
                label[I] = FINE;
                h_remove(H, I);
            }
        }
        for (p = A->rp[next]; p < A->rp[next + 1]; p++) {
            int I = A->ci.a[p];
            if (A->v.a[p] != 0. && H->pos[I] >= 0) {
                double lam = lambda_of(label, T, I, size);
                if (lam != imp[I]) {
                    h_remove(H, I);
                    imp[I] = lam;
                    h_insert(H, I, imp[I]);
                }
            }
        }
        for (p = A->rp[next]; p < A->rp[next + 1]; p++) {
            int I = A->ci.a[p];
            if (I == next) continue;
            for (q = A->rp[I]; q < A->rp[I + 1]; q++) {
                int J = A->ci.a[q];
                if (label[J] == UNDECIDED && A->v.a[q] != 0. && H->pos[J] < 0) h_insert(H, J, imp[J]);
            }
            for (q = S->ptr[I]; q < S->ptr[I + 1]; q++) {
                int J = S->idx[q];
                if (label[J] == UNDECIDED) {
                    double lam = lambda_of(label, T, J, size);
                    if (lam != imp[J]) {
                        h_remove(H, J);
                        imp[J] = lam;
                        h_insert(H, J, imp[J]);
                    }
                }
            }
        }
        if (H->n > 0) {
            next = H->h[0].node;
            h_remove(H, next);
        } else {
            next = -1;
            for (i = 0; i < size; i++)
                if (label[i] == UNDECIDED && levels[i] == level) {
                    next = i;
                    break;
                }
        }
    } while (next >= 0);
}

/* AMG_is_isolated (Multigrid.c:498-503) */
static int isolated_row(const Mat* A, int i) {
    int p, nz = 0;
    for (p = A->rp[i]; p < A->rp[i + 1]; p++)
        if (A->v.a[p] != 0) nz++;
    return nz == 1;
}

/* standard_coarsening (Multigrid.c:506-548); premode: isolated not counted */
static int standard_coarsening(const Mat* A, int* label, Lists* S, Lists* T, const int* levels, int dof,
                               int premode, double em, double ep) {
    int n = A->n, i, j, d, coarse = 0;
    double* imp = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    Heap H;
    strong_sets(A, levels, em, ep, S, T);
    for (i = 0; i < n; i++) {
        if (S->ptr[i + 1] == S->ptr[i]) label[i] = isolated_row(A, i) ? ISOLATED : COARSE;
        else label[i] = UNDECIDED;
    }
    h_init(&H, n);
    for (d = 0; d < dof; d++) {
        int seed = -1;
        for (j = 0; j < n; j++)
            if (label[j] == UNDECIDED && levels[j] == d) {
                seed = j;
                break;
            }
        if (seed >= 0) coarsener(A, label, imp, S, T, &H, levels, n, seed);
    }
    h_free(&H);
    for (i = 0; i < n; i++) {
        if (label[i] == UNDECIDED) label[i] = FINE;
        if (label[i] == COARSE) coarse++;
    }
    if (!premode)
        for (i = 0; i < n; i++)
            if (label[i] == ISOLATED) coarse++;
    free(imp);
    return coarse;
}

/* aggressive_coarsening (Multigrid.c:550-607) with get_2nd_max_coupled
 * (:99-119), Get_2nd_strong_connected (:187-217) and
 * get_2nd_strong_connection_sets (:260-292).  The uninitialised
 * Importance_2nd of the reference is taken as zero. */
static int aggressive_coarsening(const Mat* A, int* label, Lists* S, Lists* T, const int* levels, int dof,
                                 double em, double ep) {
    int n = A->n, i, p, q, I, d, k, coarse = 0;
    int* pre = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    int pre_size = standard_coarsening(A, pre, S, T, levels, dof, 1, em, ep);
    int* map = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    int* inv = (int*)malloc((pre_size > 0 ? pre_size : 1) * sizeof(int));
    Mat PA;
    Lists M2, T2;
    ivec m2 = {0}, c = {0};
    dvec v = {0};
    Acc acc;
    int *label2, *cnt;
    double* imp2;
    Heap H;
    for (i = 0; i < n; i++) label[i] = pre[i];
    k = 0;
    for (i = 0; i < n; i++) {
        map[i] = -1;
        if (pre[i] == COARSE) {
            map[i] = k;
            inv[k] = i;
            k++;
        }
    }
    mat_init(&PA, pre_size, pre_size);
    M2.ptr = (int*)calloc(pre_size + 1, sizeof(int));
    acc_init(&acc, pre_size);
    for (i = 0; i < n; i++) {
        double mx = 0;
        if (pre[i] != COARSE) continue;
        I = map[i];
        /* get_2nd_max_coupled */
        for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
            double val = A->v.a[p];
            int J = A->ci.a[p];
            double dJ = diag_of(A, J);
            if (J != i && val != 0)
                for (q = A->rp[J]; q < A->rp[J + 1]; q++) {
                    double val2 = A->v.a[q];
                    double pp = fabs(val * val2 / dJ);
                    int K = A->ci.a[q];
                    if (K != J && K != i && val2 != 0 && pp > mx) mx = pp;
                }
        }
        if (mx > 0) {
            double lo = -em * mx, hi = ep * mx;
            for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
                int J = A->ci.a[p];
                double a = A->v.a[p], dJ = diag_of(A, J);
                if (J != i && a != 0 && levels[i] == levels[J])
                    for (q = A->rp[J]; q < A->rp[J + 1]; q++) {
                        int K = A->ci.a[q];
                        double b = A->v.a[q];
                        if (pre[K] == COARSE && K != J && K != i && a * b / dJ > 0 && levels[K] == levels[i])
                            if (-a * b / dJ < lo || a * b / dJ > hi) ipush(&m2, map[K]);
                    }
            }
        }
        M2.ptr[I + 1] = m2.n;
        acc_seed(&acc, I);
        for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
            int J = A->ci.a[p];
            if (J != i && A->v.a[p] != 0)
                for (q = A->rp[J]; q < A->rp[J + 1]; q++) {
                    int K = A->ci.a[q];
                    if (pre[K] == COARSE && K != J && K != i && A->v.a[q] != 0) acc_add(&acc, map[K], 1.);
                }
        }
        acc_take(&acc, &c, &v);
        mat_append_row(&PA, I, &c, &v);
    }
    M2.idx = m2.a;
    T2.ptr = (int*)calloc(pre_size + 1, sizeof(int));
    T2.idx = (int*)malloc((m2.n > 0 ? m2.n : 1) * sizeof(int));
    for (k = 0; k < m2.n; k++) T2.ptr[m2.a[k] + 1]++;
    for (I = 0; I < pre_size; I++) T2.ptr[I + 1] += T2.ptr[I];
    cnt = (int*)malloc((pre_size + 1) * sizeof(int));
    memcpy(cnt, T2.ptr, (pre_size + 1) * sizeof(int));
    for (I = 0; I < pre_size; I++)
        for (p = M2.ptr[I]; p < M2.ptr[I + 1]; p++) T2.idx[cnt[M2.idx[p]]++] = I;
    free(cnt);
    label2 = (int*)calloc(pre_size > 0 ? pre_size : 1, sizeof(int));
    imp2 = (double*)calloc(pre_size > 0 ? pre_size : 1, sizeof(double));
    h_init(&H, pre_size);
    for (d = 0; d < dof; d++) {
        int seed = -1;
        for (I = 0; I < pre_size; I++)
            if (levels[inv[I]] == d) {
                seed = I;
                break;
            }
        if (seed >= 0) coarsener(&PA, label2, imp2, &M2, &T2, &H, levels, pre_size, seed);
    }
    h_free(&H);
    for (i = 0; i < n; i++) {
        if (label[i] == COARSE) {
            label[i] = label2[map[i]];
            if (label[i] == COARSE) coarse++;
        } else if (label[i] == ISOLATED)
            coarse++;
    }
    acc_free(&acc);
    mat_free(&PA);
    lists_free(&M2);
    lists_free(&T2);
    free(label2);
    free(imp2);
    free(pre);
    free(map);
    free(inv);
    free(c.a);
    free(v.a);
    return coarse;
}

/* AMG_simple_standard_interpolation (Multigrid.c:989-1079), then
 * sparse_approximate (linear_algebra.c:1418-1442) and normalize_row_sum
 * (:455-464) as in set_prolongation_map (Multigrid.c:1265-1286) */
static int interpolation(const Mat* A, const int* label, const int* levels, const int* map, int coarse,
                         int passes, double tol, Mat* P) {
    int n = A->n, i, p, k, t, stop = 0;
    double* D = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    Acc acc;
    ivec snc = {0}, rc = {0};
    dvec snv = {0}, rv = {0};
    for (i = 0; i < n; i++) {
        D[i] = diag_of(A, i);
        if (D[i] == 0) {
            free(D);
            return 0;
        }
    }
    acc_init(&acc, n);
    mat_init(P, n, coarse);
    for (i = 0; i < n; i++) {
        if (stop) {
            acc_seed(&acc, i);
        } else if (label[i] == FINE) {
            double s_p = 0, c_p = 0, s_m = 0, c_m = 0, alpha = 0, beta = 0, Di = D[i];
            int w = 0;
            for (p = A->rp[i]; p < A->rp[i + 1]; p++) acc_add(&acc, A->ci.a[p], A->v.a[p] / D[i]);
            for (k = 0; k < passes; k++) {
                acc_take(&acc, &snc, &snv);
                for (t = 0; t < snc.n; t++) acc_add(&acc, snc.a[t], snv.a[t]);
                for (t = 0; t < snc.n; t++) {
                    int J = snc.a[t];
                    double q = snv.a[t];
                    if (J != i && label[J] == FINE && q != 0) {
                        acc_remove(&acc, J);
                        for (p = A->rp[J]; p < A->rp[J + 1]; p++) {
                            int L = A->ci.a[p];
                            double a = A->v.a[p] / D[J];
                            if (L != J && a != 0) acc_add(&acc, L, -a * q);
                        }
                    }
                }
            }
            acc_take(&acc, &rc, &rv);
            for (t = 0; t < rc.n; t++)
                if (levels[rc.a[t]] == levels[i]) {
                    rc.a[w] = rc.a[t];
                    rv.a[w] = rv.a[t];
                    w++;
                }
            rc.n = rv.n = w;
            for (t = 0; t < rc.n; t++) {
                int J = rc.a[t];
                double q = rv.a[t];
                if (J != i) {
                    if (q < 0) {
                        s_m += q;
                        if (label[J] == COARSE) c_m += q;
                    } else {
                        s_p += q;
                        if (label[J] == COARSE) c_p += q;
                    }
                } else
                    Di = q;
            }
            if (c_m != 0) alpha = -s_m / c_m;
            else Di += s_m;
            if (c_p != 0) beta = -s_p / c_p;
            else Di += s_p;
            acc_seed(&acc, i);
            if (alpha == 0 && beta == 0) {
                stop = 1; /* the reference breaks out of its row loop here */
            } else {
                alpha /= Di;
                beta /= Di;
                for (t = 0; t < rc.n; t++) {
                    int J = rc.a[t];
                    double q = rv.a[t];
                    if (J != i && label[J] == COARSE) {
                        if (q < 0 && alpha != 0) acc_add(&acc, map[J], alpha * q);
                        if (q > 0 && beta != 0) acc_add(&acc, map[J], beta * q);
                    }
                }
            }
        } else {
            acc_seed(&acc, i);
            acc_add(&acc, map[i], 1.);
        }
        acc_take(&acc, &rc, &rv);
        row_delete_zeros(i, &rc, &rv);
        {
            double mx = 0, sum = 0;
            int w = 0;
            for (t = 0; t < rc.n; t++)
                if (rc.a[t] != i && fabs(rv.a[t]) > mx) mx = fabs(rv.a[t]);
            for (t = 0; t < rc.n; t++) {
                if (rc.a[t] != i && fabs(rv.a[t]) < mx * tol) continue;
                rc.a[w] = rc.a[t];
                rv.a[w] = rv.a[t];
                w++;
            }
            rc.n = rv.n = w;
            for (t = 0; t < rc.n; t++) sum += rv.a[t];
            if (sum != 0)
                for (t = 0; t < rc.n; t++) rv.a[t] *= 1. / sum;
        }
        mat_append_row(P, i, &rc, &rv);
    }
    acc_free(&acc);
    free(D);
    free(snc.a);
    free(snv.a);
    free(rc.a);
    free(rv.a);
    return 1;
}

/* get_AMG_transposed (Multigrid.c:308-321): R = P^T, zeros of P skipped,
 * each row k seeded with (k,0.0) by sparse_zero */
static void restriction(const Mat* P, int coarse, Mat* R) {
    int i, p, k;
    ivec c = {0};
    dvec v = {0};
    /* counting sort of P's nonzeros by column, then each row k merged with
     * its (k,0.0) seed */
    int* cnt = (int*)calloc(coarse + 1, sizeof(int));
    int* fill;
    int* ci;
    double* cv;
    for (i = 0; i < P->n; i++)
        for (p = P->rp[i]; p < P->rp[i + 1]; p++)
            if (P->v.a[p] != 0) cnt[P->ci.a[p] + 1]++;
    for (k = 0; k < coarse; k++) cnt[k + 1] += cnt[k];
    ci = (int*)malloc((cnt[coarse] > 0 ? cnt[coarse] : 1) * sizeof(int));
    cv = (double*)malloc((cnt[coarse] > 0 ? cnt[coarse] : 1) * sizeof(double));
    fill = (int*)malloc((coarse + 1) * sizeof(int));
    memcpy(fill, cnt, (coarse + 1) * sizeof(int));
    for (i = 0; i < P->n; i++)
        for (p = P->rp[i]; p < P->rp[i + 1]; p++)
            if (P->v.a[p] != 0) {
                k = P->ci.a[p];
                ci[fill[k]] = i;
                cv[fill[k]] = P->v.a[p];
                fill[k]++;
            }
    mat_init(R, coarse, P->n);
    for (k = 0; k < coarse; k++) {
        int seeded = 0;
        c.n = v.n = 0;
        for (p = cnt[k]; p < cnt[k + 1]; p++) {
            if (!seeded && ci[p] >= k) {
                if (ci[p] > k) {
                    ipush(&c, k);
                    dpush(&v, 0.0);
                }
                seeded = 1;
            }
            ipush(&c, ci[p]);
            dpush(&v, cv[p]);
        }
        if (!seeded) {
            ipush(&c, k);
            dpush(&v, 0.0);
        }
        mat_append_row(R, k, &c, &v);
    }
    free(cnt);
    free(fill);
    free(ci);
    free(cv);
    free(c.a);
    free(v.a);
}

/* ------------------------------------------------------------- hierarchy */

struct orc_amg {
    int depth, dof, smiter, aggr_smiter, coarsening;
    double em, ep, tol;
    Mat* A;      /* Coarse_A[l] */
    Mat* P;      /* Prolongation[l] */
    Mat* R;      /* Restriction[l] */
    int** label; /* Coarsening[l] */
    int* coarse;
    double* signs;
    /* coarsest: sparse_LU_factorization (linear_algebra.c:3040-3062) */
    int nd;
    int* didx;
    int* dmap;
    double *L, *U; /* dense factors over the non-trivial rows, row-major */
    /* colourings per level for the multicolour smoothers */
    int** colour;
    int* ncolour;
    /* smoother of the levels below the finest (-1: the same as the finest) */
    int coarse_sm, coarse_sweeps;
    /* cycle structure of the preconditioner calls (0 STAND_ALONE, 2 FULL_AMG) */
    int cycle;
};

/* LU_factorzation (linear_algebra.c:2909-2942): unit-lower L, U by bordering,
 * with Lower_triangular_invert (:2865-2878) and
 * Upper_triangular_transpose_invert (:2894-2907) */
static void dense_lu(const double* M, int N, double* L, double* U) {
    int s, i, j;
    double* u = (double*)malloc((N > 0 ? N : 1) * sizeof(double));
    double* l = (double*)malloc((N > 0 ? N : 1) * sizeof(double));
    memset(L, 0, (size_t)N * N * sizeof(double));
    memset(U, 0, (size_t)N * N * sizeof(double));
    if (N == 0) {
        free(u);
        free(l);
        return;
    }
    L[0] = 1.;
    U[0] = M[0];
    for (s = 1; s < N; s++) {
        for (i = 0; i < s; i++) u[i] = M[(size_t)i * N + s];
        for (i = 0; i < s; i++) {
            for (j = 0; j < i; j++) u[i] -= L[(size_t)i * N + j] * u[j];
            u[i] /= L[(size_t)i * N + i];
        }
        for (i = 0; i < s; i++) l[i] = M[(size_t)s * N + i];
        for (i = 0; i < s; i++) {
            for (j = 0; j < i; j++) l[i] -= U[(size_t)j * N + i] * l[j];
            l[i] /= U[(size_t)i * N + i];
        }
        for (i = 0; i < s; i++) {
            U[(size_t)i * N + s] = u[i];
            L[(size_t)s * N + i] = l[i];
        }
        {
            double dg = M[(size_t)s * N + s];
            for (i = 0; i < s; i++) dg -= u[i] * l[i];
            U[(size_t)s * N + s] = dg;
        }
        L[(size_t)s * N + s] = 1.;
    }
    free(u);
    free(l);
}

static void colour_level(orc_amg* h, int l) {
    const Mat* A = &h->A[l];
    h->colour[l] = (int*)malloc((A->n > 0 ? A->n : 1) * sizeof(int));
    h->ncolour[l] = orc_colouring(A->n, A->rp, A->ci.a, A->v.a, h->colour[l]);
}

static void coarse_factor(orc_amg* h) {
    const Mat* A0 = &h->A[0];
    int n = A0->n, i, p, k = 0;
    double* M;
    h->dmap = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    h->didx = (int*)malloc((n > 0 ? n : 1) * sizeof(int));
    for (i = 0; i < n; i++) {
        int nz = 0;
        for (p = A0->rp[i]; p < A0->rp[i + 1]; p++)
            if (A0->v.a[p] != 0) nz++;
        h->dmap[i] = -1;
        if (nz > 1) {
            h->dmap[i] = k;
            h->didx[k++] = i;
        }
    }
    h->nd = k;
    M = (double*)calloc((size_t)k * k + 1, sizeof(double));
    for (i = 0; i < k; i++) {
        int r = h->didx[i];
        for (p = A0->rp[r]; p < A0->rp[r + 1]; p++) {
            int J = h->dmap[A0->ci.a[p]];
            if (J >= 0) M[(size_t)i * k + J] = A0->v.a[p];
        }
    }
    h->L = (double*)malloc(((size_t)k * k + 1) * sizeof(double));
    h->U = (double*)malloc(((size_t)k * k + 1) * sizeof(double));
    dense_lu(M, k, h->L, h->U);
    free(M);
}

static void coarse_free(orc_amg* h) {
    free(h->dmap);
    free(h->didx);
    free(h->L);
    free(h->U);
    h->dmap = h->didx = NULL;
    h->L = h->U = NULL;
}

static int build_levels(orc_amg* h, const int* finest_levels) {
    int depth = h->depth, i, r;
    int* levels = (int*)malloc((h->A[depth - 1].n > 0 ? h->A[depth - 1].n : 1) * sizeof(int));
    memcpy(levels, finest_levels, h->A[depth - 1].n * sizeof(int));
    for (i = depth - 1; i > 0; i--) {
        Mat* A = &h->A[i];
        Lists S, T;
        int coarse, k = 0, *map, *nl, ok;
        h->label[i] = (int*)calloc(A->n > 0 ? A->n : 1, sizeof(int));
        if (h->coarsening == AGGRESSIVE && i == depth - 1)
            coarse = aggressive_coarsening(A, h->label[i], &S, &T, levels, h->dof, h->em, h->ep);
        else
            coarse = standard_coarsening(A, h->label[i], &S, &T, levels, h->dof, 0, h->em, h->ep);
        lists_free(&S);
        lists_free(&T);
        h->coarse[i] = coarse;
        map = (int*)malloc((A->n > 0 ? A->n : 1) * sizeof(int));
        for (r = 0; r < A->n; r++) map[r] = h->label[i][r] != FINE ? k++ : -1;
        if (k != coarse) {
            free(map);
            free(levels);
            return 0;
        }
        nl = (int*)malloc((coarse > 0 ? coarse : 1) * sizeof(int));
        for (r = 0; r < A->n; r++)
            if (map[r] >= 0) nl[map[r]] = levels[r];
        ok = interpolation(A, h->label[i], levels, map, coarse,
                           (h->coarsening == AGGRESSIVE && i == depth - 1) ? 2 : 1, h->tol, &h->P[i]);
        free(map);
        if (!ok) {
            free(nl);
            free(levels);
            return 0;
        }
        restriction(&h->P[i], coarse, &h->R[i]);
        if (!galerkin(A, &h->R[i], &h->P[i], &h->A[i - 1])) {
            free(nl);
            free(levels);
            return 0;
        }
        free(levels);
        levels = nl;
    }
    free(levels);
    coarse_factor(h);
    for (i = 1; i < depth; i++) colour_level(h, i);
    return 1;
}

orc_amg* orc_amg_setup(int n, const int* rp, const int* ci, const double* v, int dof, int depth, int smiter,
                       int aggr_smiter, int coarsening, double eps_minus, double eps_plus, double matrix_tol) {
    orc_amg* h = (orc_amg*)calloc(1, sizeof(orc_amg));
    int i, p, mesh;
    int* levels;
    h->depth = depth;
    h->dof = dof;
    h->smiter = smiter;
    h->aggr_smiter = aggr_smiter;
    h->coarsening = coarsening;
    h->em = eps_minus;
    h->ep = eps_plus;
    h->tol = matrix_tol;
    h->A = (Mat*)calloc(depth, sizeof(Mat));
    h->P = (Mat*)calloc(depth, sizeof(Mat));
    h->R = (Mat*)calloc(depth, sizeof(Mat));
    h->label = (int**)calloc(depth, sizeof(int*));
    h->coarse = (int*)calloc(depth, sizeof(int));
    h->colour = (int**)calloc(depth, sizeof(int*));
    h->ncolour = (int*)calloc(depth, sizeof(int));
    h->coarse_sm = -1;
    h->coarse_sweeps = 1;
    /* clone + make_positive_diag (linear_algebra.c:2612-2620) */
    mat_from(&h->A[depth - 1], n, rp, ci, v);
    h->signs = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    for (i = 0; i < n; i++) {
        h->signs[i] = 1.;
        if (diag_of(&h->A[depth - 1], i) < 0) {
            for (p = rp[i]; p < rp[i + 1]; p++) h->A[depth - 1].v.a[p] *= -1.;
            h->signs[i] = -1.;
        }
    }
    /* AMG_init_finest_tupels (Multigrid.c:785-804) */
    levels = (int*)calloc(n > 0 ? n : 1, sizeof(int));
    mesh = n / dof;
    for (i = 0; i < mesh; i++)
        for (p = 1; p <= dof; p++) levels[(p - 1) * mesh + i] = p - 1;
    if (!build_levels(h, levels)) {
        free(levels);
        orc_amg_free(h);
        return NULL;
    }
    free(levels);
    return h;
}

void orc_amg_free(orc_amg* h) {
    int l;
    if (!h) return;
    for (l = 0; l < h->depth; l++) {
        mat_free(&h->A[l]);
        mat_free(&h->P[l]);
        mat_free(&h->R[l]);
        free(h->label[l]);
        free(h->colour[l]);
    }
    free(h->A);
    free(h->P);
    free(h->R);
    free(h->label);
    free(h->coarse);
    free(h->colour);
    free(h->ncolour);
    free(h->signs);
    coarse_free(h);
    free(h);
}

int orc_amg_level(const orc_amg* h, int l, int* fine, int* coarse, int* nnz_a, int* nnz_p, int* nnz_r) {
    if (l < 0 || l >= h->depth) return -1;
    *fine = h->A[l].n;
    *coarse = h->coarse[l];
    *nnz_a = mat_nnz(&h->A[l]);
    *nnz_p = l > 0 ? mat_nnz(&h->P[l]) : 0;
    *nnz_r = l > 0 ? mat_nnz(&h->R[l]) : 0;
    return 0;
}

int orc_amg_copy(const orc_amg* h, int l, char which, int* rp, int* ci, double* v) {
    const Mat* M;
    if (l < 0 || l >= h->depth) return -1;
    M = which == 'A' ? &h->A[l] : which == 'P' ? &h->P[l] : &h->R[l];
    if (l == 0 && which != 'A') return -1;
    memcpy(rp, M->rp, (M->n + 1) * sizeof(int));
    memcpy(ci, M->ci.a, mat_nnz(M) * sizeof(int));
    memcpy(v, M->v.a, mat_nnz(M) * sizeof(double));
    return 0;
}

int orc_amg_labels(const orc_amg* h, int l, int* labels) {
    if (l <= 0 || l >= h->depth) return -1;
    memcpy(labels, h->label[l], h->A[l].n * sizeof(int));
    return 0;
}

/* AMG_recompute_coarse_matrices (Multigrid.c:1756-1780): new finest values
 * (clone(New_matrix), Signs unchanged), Galerkin products with the kept P/R */
int orc_amg_recompute(orc_amg* h, const double* v) {
    int l, k, d = h->depth - 1;
    Mat* A = &h->A[d];
    for (k = 0; k < mat_nnz(A); k++) A->v.a[k] = v[k];
    for (l = d; l > 0; l--) {
        mat_free(&h->A[l - 1]);
        if (!galerkin(&h->A[l], &h->R[l], &h->P[l], &h->A[l - 1])) return -1;
    }
    coarse_free(h);
    coarse_factor(h);
    for (l = 1; l < h->depth; l++) {
        free(h->colour[l]);
        colour_level(h, l);
    }
    return 0;
}

/* ------------------------------------------------------------- kernels */

void orc_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y) {
    int i, p;
    for (i = 0; i < n; i++) {
        double s = 0;
        for (p = rp[i]; p < rp[i + 1]; p++) s += v[p] * x[ci[p]];
        y[i] = s;
    }
}

/* linear_map (linear_algebra.c:149-172): x += fac * A y */
static void linear_map(double* x, double fac, const Mat* A, const double* y) {
    int i, p;
    for (i = 0; i < A->n; i++) {
        double s = 0;
        for (p = A->rp[i]; p < A->rp[i + 1]; p++) s += A->v.a[p] * y[A->ci.a[p]];
        s *= fac;
        x[i] += s;
    }
}

static double dot(const double* a, const double* b, int n) {
    int i;
    double s = 0;
    for (i = 0; i < n; i++) s = s + a[i] * b[i];
    return s;
}

/* matrix_residuum (linear_algebra.c:2431-2438) */
static double residual_norm(const Mat* A, const double* X, const double* F) {
    int n = A->n;
    double* R = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    double r;
    memcpy(R, F, n * sizeof(double));
    linear_map(R, -1., A, X);
    r = sqrt(dot(R, R, n));
    free(R);
    return r;
}

int orc_colouring(int n, const int* rp, const int* ci, const double* v, int* colour) {
    int i, p, ncol = 0, cap = 64;
    int* tp = (int*)calloc(n + 1, sizeof(int));
    int *ta, *fill, *mark;
    for (i = 0; i < n; i++)
        for (p = rp[i]; p < rp[i + 1]; p++)
            if (ci[p] != i && v[p] != 0) tp[ci[p] + 1]++;
    for (i = 0; i < n; i++) tp[i + 1] += tp[i];
    ta = (int*)malloc((tp[n] > 0 ? tp[n] : 1) * sizeof(int));
    fill = (int*)malloc((n + 1) * sizeof(int));
    memcpy(fill, tp, (n + 1) * sizeof(int));
    for (i = 0; i < n; i++)
        for (p = rp[i]; p < rp[i + 1]; p++)
            if (ci[p] != i && v[p] != 0) ta[fill[ci[p]]++] = i;
    mark = (int*)malloc(cap * sizeof(int));
    for (p = 0; p < cap; p++) mark[p] = -1;
    for (i = 0; i < n; i++) {
        int c = 0;
        for (p = rp[i]; p < rp[i + 1]; p++)
            if (ci[p] < i && v[p] != 0) mark[colour[ci[p]]] = i;
        for (p = tp[i]; p < tp[i + 1]; p++)
            if (ta[p] < i) mark[colour[ta[p]]] = i;
        while (c < cap && mark[c] == i) c++;
        if (c == cap) {
            int q;
            mark = (int*)realloc(mark, 2 * cap * sizeof(int));
            for (q = cap; q < 2 * cap; q++) mark[q] = -1;
            cap *= 2;
        }
        colour[i] = c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    free(tp);
    free(ta);
    free(fill);
    free(mark);
    return ncol;
}

/* one SOR row update (AMG_SOR_smoother, Multigrid.c:1345-1365) */
static void sor_row(const Mat* A, const double* F, double* X, int i, double w) {
    int p, d = -1;
    double sum = F[i];
    for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
        int j = A->ci.a[p];
        if (j != i) sum -= A->v.a[p] * X[j];
        else d = p;
    }
    X[i] = (1. - w) * X[i] + w * sum / A->v.a[d];
}

/* planted mistakes (tests only; 0 = none): 1 = the multicolour sweep reads
 * every neighbour value from the start of the sweep (a stale gather),
 * 2 = the symmetric multicolour sweep keeps the forward colour order when
 * post-smoothing.  orc_plant(0) restores the restatement. */
static int planted = 0;
void orc_plant(int which) { planted = which; }

/* smoother ids: 0 SOR, 1 parallel Jacobi, 2 MC-SOR, 3 l1-Jacobi, 4 MC-SSOR, 5 SSOR.
 * SOR / SSOR: AMG_SOR_smoother's row update (Multigrid.c:1345-1365) over the
 * rows in order, the symmetric form's post-smoothing call in reverse order;
 * MC-SOR / MC-SSOR: the same row update colour by colour (rows of a colour
 * in natural order; the symmetric form's post call takes the colours in
 * reverse); parallel Jacobi: AMG_parallel_Jacobi_smoother (Multigrid.c:
 * 1367-1400, undamped, rows beyond (n/dof)*dof untouched); l1-Jacobi:
 * x += (f - A x) / sum_j |a_ij|. */
static void smooth_csr(const Mat* A, const int* colour, int nc, int dof, const double* F, double* X, int sweeps,
                       double w, int sm, int post) {
    int n = A->n, k, i, c, p;
    for (k = 0; k < sweeps; k++) {
        if (sm == 0 || sm == 5) {
            int back = sm == 5 && post;
            for (i = 0; i < n; i++) sor_row(A, F, X, back ? n - 1 - i : i, w);
        } else if (sm == 2 || sm == 4) {
            int back = sm == 4 && post && planted != 2;
            double* S = planted == 1 ? (double*)malloc((n > 0 ? n : 1) * sizeof(double)) : NULL;
            if (S) memcpy(S, X, n * sizeof(double));
            for (c = 0; c < nc; c++) {
                int cc = back ? nc - 1 - c : c;
                for (i = 0; i < n; i++)
                    if (colour[i] == cc) {
                        if (S) { /* planted: gathers from the sweep's start */
                            double sum = F[i];
                            int d = -1;
                            for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
                                if (A->ci.a[p] != i) sum -= A->v.a[p] * S[A->ci.a[p]];
                                else d = p;
                            }
                            X[i] = (1. - w) * X[i] + w * sum / A->v.a[d];
                        } else {
                            sor_row(A, F, X, i, w);
                        }
                    }
            }
            free(S);
        } else {
            double* O = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
            int lim = sm == 1 ? (n / dof) * dof : n;
            memcpy(O, X, n * sizeof(double));
            for (i = 0; i < lim; i++) {
                double sum = F[i], d = 0;
                for (p = A->rp[i]; p < A->rp[i + 1]; p++) {
                    int j = A->ci.a[p];
                    if (sm == 1) {
                        if (j != i) sum -= A->v.a[p] * O[j];
                        else d = A->v.a[p];
                    } else {
                        sum -= A->v.a[p] * O[j];
                        d += fabs(A->v.a[p]);
                    }
                }
                X[i] = sm == 1 ? sum / d : O[i] + sum / d;
            }
            free(O);
        }
    }
}

static void smooth(orc_amg* h, int l, const double* F, double* X, int sweeps, double w, int sm, int post) {
    smooth_csr(&h->A[l], h->colour[l], h->ncolour[l], h->dof, F, X, sweeps, w, sm, post);
}

void orc_smooth(int n, const int* rp, const int* ci, const double* v, const int* colour, int ncolour, int dof,
                const double* F, double* X, int sweeps, double w, int sm, int post) {
    Mat A;
    A.n = A.ncols = n;
    A.rp = (int*)rp;
    A.ci.a = (int*)ci;
    A.v.a = (double*)v;
    A.ci.n = A.ci.cap = A.v.n = A.v.cap = rp[n];
    smooth_csr(&A, colour, ncolour, dof, F, X, sweeps, w, sm, post);
}

/* AMG_Direct_solver (Multigrid.c:1450-1455): Sol = F, L_triang_invert,
 * U_triang_invert (linear_algebra.c:2008-2043) with the factors of
 * sparse_LU_factorization (trivial rows: L=1, U=diag) */
static void coarse_solve(orc_amg* h, const double* F, double* X) {
    const Mat* A0 = &h->A[0];
    int n = A0->n, N = h->nd, i, j;
    double* y = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    memcpy(X, F, n * sizeof(double));
    /* forward: rows in index order; dense rows use L, trivial rows L_ii = 1 */
    for (i = 0; i < n; i++) {
        int I = h->dmap[i];
        if (I < 0) continue;
        {
            double s = X[i];
            for (j = 0; j < I; j++) {
                double l = h->L[(size_t)I * N + j];
                if (l != 0) s -= l * X[h->didx[j]];
            }
            X[i] = s;
        }
    }
    memcpy(y, X, n * sizeof(double));
    for (i = n - 1; i >= 0; i--) {
        int I = h->dmap[i];
        if (I < 0) {
            X[i] = y[i] / diag_of(A0, i);
            continue;
        }
        {
            double s = y[i];
            for (j = I + 1; j < N; j++) {
                double u = h->U[(size_t)I * N + j];
                if (u != 0) s -= u * X[h->didx[j]];
            }
            X[i] = s / h->U[(size_t)I * N + I];
        }
    }
    free(y);
}

/* AMG_V_cycle (Multigrid.c:1536-1580), Iter[level] = 1 */
static void vcycle(orc_amg* h, int l, const double* F, double* X, int sm, double w) {
    const Mat* A = &h->A[l];
    int n = A->n, nc, sw;
    double *R, *Fc, *Xc;
    if (l == 0) {
        coarse_solve(h, F, X);
        return;
    }
    nc = h->A[l - 1].n;
    sw = (h->coarsening == AGGRESSIVE && l == h->depth - 1) ? h->aggr_smiter : h->smiter;
    if (l < h->depth - 1 && h->coarse_sm >= 0) {
        sm = h->coarse_sm;
        sw = h->coarse_sweeps;
    }
    smooth(h, l, F, X, sw, w, sm, 0);
    R = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    memcpy(R, F, n * sizeof(double));
    linear_map(R, -1., A, X);
    Fc = (double*)malloc((nc > 0 ? nc : 1) * sizeof(double));
    Xc = (double*)calloc(nc > 0 ? nc : 1, sizeof(double));
    {
        /* sparse_mult (linear_algebra.c:134-147) of the restriction */
        const Mat* Rm = &h->R[l];
        int i, p;
        for (i = 0; i < Rm->n; i++) {
            double s = 0;
            for (p = Rm->rp[i]; p < Rm->rp[i + 1]; p++) s += Rm->v.a[p] * R[Rm->ci.a[p]];
            Fc[i] = s;
        }
    }
    vcycle(h, l - 1, Fc, Xc, sm, w);
    {
        /* linear_map(X, 1, P, Xc); P's seed entries outside the coarse range are zero */
        const Mat* P = &h->P[l];
        int i, p;
        for (i = 0; i < P->n; i++) {
            double s = 0;
            for (p = P->rp[i]; p < P->rp[i + 1]; p++)
                if (P->ci.a[p] < nc) s += P->v.a[p] * Xc[P->ci.a[p]];
            s *= 1.;
            X[i] += s;
        }
    }
    smooth(h, l, F, X, sw, w, sm, 1);
    free(R);
    free(Fc);
    free(Xc);
}

/* AMG_solve (Multigrid.c:1863-1931), STAND_ALONE; the diagnostic ILU
 * residuals decide only `res0 != 0`, i.e. r0 != 0 */
int orc_amg_solve(orc_amg* h, const double* Fin, double* X, int sm, double w, double eps, int max_iter) {
    int d = h->depth - 1, n = h->A[d].n, i, p, it = 0;
    const Mat* A = &h->A[d];
    double *F = (double*)malloc((n > 0 ? n : 1) * sizeof(double)), r0, r;
    for (i = 0; i < n; i++) F[i] = 0 + h->signs[i] * Fin[i];
    /* AMG_fix_isolated (Multigrid.c:1301-1319) */
    for (i = 0; i < n; i++)
        if (h->label[d][i] == ISOLATED)
            for (p = A->rp[i]; p < A->rp[i + 1]; p++)
                if (A->v.a[p] != 0) {
                    X[i] = F[i] / A->v.a[p];
                    break;
                }
    r0 = residual_norm(A, X, F);
    r = r0;
    if (r0 != 0)
        while (fabs(r / r0) > eps && it < max_iter) {
            vcycle(h, d, F, X, sm, w);
            r = residual_norm(A, X, F);
            it++;
        }
    free(F);
    return it;
}

/* sparse_mult (linear_algebra.c:134-147) of a transfer matrix, y = M x with
 * x of length nx; P's seed entries past the coarse range are zero and skipped
 * (the reference reads past the coarse vector there and multiplies by 0) */
static double* transfer(const Mat* M, const double* x, int nx) {
    double* y = (double*)malloc((M->n > 0 ? M->n : 1) * sizeof(double));
    int i, p;
    for (i = 0; i < M->n; i++) {
        double s = 0;
        for (p = M->rp[i]; p < M->rp[i + 1]; p++)
            if (M->ci.a[p] < nx) s += M->v.a[p] * x[M->ci.a[p]];
        y[i] = s;
    }
    return y;
}

/* FMG_recursive (Multigrid.c:1582-1593): the iterate restricted, the level
 * below solved the same way, its result prolongated (replacing the level's
 * iterate) and one V-cycle on CF[l] */
static void fmg(orc_amg* h, double** CF, double* X, int l, int sm, double w) {
    double *Xc, *Xf;
    if (l == 0) {
        coarse_solve(h, CF[0], X);
        return;
    }
    Xc = transfer(&h->R[l], X, h->A[l].n);
    fmg(h, CF, Xc, l - 1, sm, w);
    Xf = transfer(&h->P[l], Xc, h->A[l - 1].n);
    vcycle(h, l, CF[l], Xf, sm, w);
    memcpy(X, Xf, h->A[l].n * sizeof(double));
    free(Xc);
    free(Xf);
}

/* AMG_solve (Multigrid.c:1863-1931), FULL_AMG: after the sign flip and
 * AMG_fix_isolated as in the STAND_ALONE case, one AMG_full_multigrid
 * (Multigrid.c:1595-1611) when r0 != 0: CF[depth-1] = F, CF[l-1] = R_l CF[l],
 * FMG_recursive from the finest level; returns the cycle count (1, or 0
 * when r0 == 0) */
/* the cycle structure AMG_solve applies when it preconditions pcg (AMG_setup's
 * cycle argument, Multigrid.c:1902-1915): 0 STAND_ALONE, 2 FULL_AMG */
void orc_set_cycle(orc_amg* h, int cycle) { h->cycle = cycle; }

int orc_amg_solve_fmg(orc_amg* h, const double* Fin, double* X, int sm, double w) {
    int d = h->depth - 1, n = h->A[d].n, i, p, l;
    const Mat* A = &h->A[d];
    double** CF;
    double r0;
    CF = (double**)malloc(h->depth * sizeof(double*));
    CF[d] = (double*)malloc((n > 0 ? n : 1) * sizeof(double));
    for (i = 0; i < n; i++) CF[d][i] = 0 + h->signs[i] * Fin[i];
    for (i = 0; i < n; i++)
        if (h->label[d][i] == ISOLATED)
            for (p = A->rp[i]; p < A->rp[i + 1]; p++)
                if (A->v.a[p] != 0) {
                    X[i] = CF[d][i] / A->v.a[p];
                    break;
                }
    r0 = residual_norm(A, X, CF[d]);
    for (l = d; l > 0; l--) CF[l - 1] = transfer(&h->R[l], CF[l], h->A[l].n);
    if (r0 != 0) fmg(h, CF, X, d, sm, w);
    for (l = 0; l <= d; l++) free(CF[l]);
    free(CF);
    return r0 != 0;
}

/* pcg (PCGS.c:163-233) with an AMG_solve preconditioner (PCG_AMG_precon) */
int orc_pcg_amg(orc_amg* h, double* X, double* B, double tol, int max_iter, int sm, double w, double amg_eps,
                int amg_max_iter, int zero_init, int stop, double* res, int* vcycles) {
    int n = h->A[h->depth - 1].n, i, it = 0;
    double hn, s = 0, s_old, a, b, r, r0;
    double *R, *P, *W, *Q, *Z, *Y;
    int vc = 0;
#define PRECON(RR, ZZ)                                                       \
    do {                                                                     \
        if (zero_init) memset((ZZ), 0, n * sizeof(double));                  \
        vc += h->cycle == 2 ? orc_amg_solve_fmg(h, (RR), (ZZ), sm, w)        \
                            : orc_amg_solve(h, (RR), (ZZ), sm, w, amg_eps, amg_max_iter); \
    } while (0)
    hn = sqrt(dot(B, B, n));
    if (hn == 0) hn = 1;
    for (i = 0; i < n; i++) B[i] *= (double)1 / hn;
    for (i = 0; i < n; i++) X[i] *= (double)1 / hn;
    W = (double*)calloc(n, sizeof(double));
    Q = (double*)calloc(n, sizeof(double));
    Z = (double*)calloc(n, sizeof(double));
    P = (double*)calloc(n, sizeof(double));
    R = (double*)malloc(n * sizeof(double));
    Y = (double*)malloc(n * sizeof(double));
    memcpy(R, B, n * sizeof(double));
    orc_spmv(n, h->A[h->depth - 1].rp, h->A[h->depth - 1].ci.a, h->A[h->depth - 1].v.a, X, Q);
    /* the finest Coarse_A has positive diagonal; PCG_mult uses the caller's
     * matrix, equal to it up to the row signs */
    for (i = 0; i < n; i++) Q[i] = h->signs[i] * Q[i];
    for (i = 0; i < n; i++) R[i] += -1. * Q[i];
    PRECON(R, W);
    r0 = sqrt(dot(W, W, n));
    if (r0 == 0) {
        *res = 0;
        *vcycles = vc;
        free(W); free(Q); free(Z); free(P); free(R); free(Y);
        return 0;
    }
    r = r0;
    do {
        PRECON(R, Z);
        s_old = s;
        s = dot(R, Z, n);
        if (it > 0) {
            b = s / s_old;
            for (i = 0; i < n; i++) P[i] *= b;
            for (i = 0; i < n; i++) P[i] += 1 * Z[i];
        } else
            memcpy(P, Z, n * sizeof(double));
        orc_spmv(n, h->A[h->depth - 1].rp, h->A[h->depth - 1].ci.a, h->A[h->depth - 1].v.a, P, Q);
        for (i = 0; i < n; i++) Q[i] = h->signs[i] * Q[i];
        a = s / dot(P, Q, n);
        for (i = 0; i < n; i++) X[i] += a * P[i];
        for (i = 0; i < n; i++) R[i] += -a * Q[i];
        r = sqrt(dot(R, R, n));
        it++;
    } while ((stop ? r : r / r0) > tol && it < max_iter);
    PRECON(R, W);
    r = sqrt(dot(W, W, n));
    for (i = 0; i < n; i++) B[i] *= hn;
    for (i = 0; i < n; i++) X[i] *= hn;
    *res = r / r0;
    *vcycles = vc;
    free(W); free(Q); free(Z); free(P); free(R); free(Y);
#undef PRECON
    return it;
}

/* smoother of the levels below the finest; sm < 0 restores the finest's */
void orc_set_coarse_smoother(orc_amg* h, int sm, int sweeps) {
    h->coarse_sm = sm;
    h->coarse_sweeps = sweeps;
}

/* ---------------------------------------------------------------- ILU(0) */

/* sum of l[k] * u[k] over the columns k <= lim held by both sorted lists, in
 * increasing k from 0 (row_times_row, linear_algebra.c:1795-1819) */
static double pair_dot(const int* lc, const double* lv, int ln, const int* uc, const double* uv, int un, int lim) {
    int a, b = 0;
    double s = 0;
    for (a = 0; a < ln && lc[a] <= lim; a++) {
        while (b < un && uc[b] < lc[a]) b++;
        if (b < un && uc[b] == lc[a]) s += lv[a] * uv[b];
    }
    return s;
}

/* incomplete_LU_factorization (linear_algebra.c:1958-1984).  The reference
 * seeds L and U with sparse_zero (linear_algebra.c:501-508): every row holds
 * its diagonal (0.0) before the sweep, and insert_sparse adds to an entry that
 * exists, so L's diagonal is 0.0 + (a_ii - sum).  U is built by columns (its
 * rows hold U^T) and transposed at the end.  A zero pivot returns 0 and leaves
 * the outputs undefined. */
int orc_ilu0(int n, const int* rp, const int* ci, const double* v, int* lrp, int* lci, double* lv, int* urp,
             int* uci, double* uv) {
    int i, j, p, nl = 0;
    int *cp = (int*)calloc(n + 1, sizeof(int)), *fill = (int*)malloc((n + 1) * sizeof(int));
    int *cr, *pos;
    double* cv;
    /* U^T rows: column j of U lists (i, u_ij), i < j, by increasing i */
    for (i = 0; i < n; i++)
        for (p = rp[i]; p < rp[i + 1]; p++)
            if (ci[p] > i) cp[ci[p] + 1]++;
    for (j = 0; j < n; j++) cp[j + 1] += cp[j];
    memcpy(fill, cp, (n + 1) * sizeof(int));
    cr = (int*)malloc((cp[n] > 0 ? cp[n] : 1) * sizeof(int));
    cv = (double*)malloc((cp[n] > 0 ? cp[n] : 1) * sizeof(double));
    for (i = 0; i < n; i++) {
        int l0 = nl;
        double dv = 0.0, s;
        lrp[i] = nl;
        for (p = rp[i]; p < rp[i + 1]; p++) {
            j = ci[p];
            if (j <= i) {
                s = v[p];
                s -= pair_dot(lci + l0, lv + l0, nl - l0, cr + cp[j], cv + cp[j], fill[j] - cp[j], j - 1);
                if (j < i) {
                    lci[nl] = j;
                    lv[nl] = s;
                    nl++;
                } else
                    dv = 0.0 + s;
            } else {
                s = v[p];
                s -= pair_dot(lci + l0, lv + l0, nl - l0, cr + cp[j], cv + cp[j], fill[j] - cp[j], i - 1);
                if (dv == 0) {
                    free(cp); free(fill); free(cr); free(cv);
                    return 0;
                }
                s /= dv;
                cr[fill[j]] = i;
                cv[fill[j]] = s;
                fill[j]++;
            }
        }
        lci[nl] = i;
        lv[nl] = dv;
        nl++;
    }
    lrp[n] = nl;
    /* transpose: U row i = (i, 1.0) then (j, u_ij) by increasing j */
    pos = (int*)malloc((n + 1) * sizeof(int));
    urp[0] = 0;
    for (i = 0; i < n; i++) {
        int up = 0;
        for (p = rp[i]; p < rp[i + 1]; p++) up += ci[p] > i;
        urp[i + 1] = urp[i] + 1 + up;
        uci[urp[i]] = i;
        uv[urp[i]] = 0.0 + 1.0;
        pos[i] = urp[i] + 1;
    }
    for (j = 0; j < n; j++)
        for (p = cp[j]; p < cp[j + 1]; p++) {
            uci[pos[cr[p]]] = j;
            uv[pos[cr[p]]] = cv[p];
            pos[cr[p]]++;
        }
    free(cp); free(fill); free(cr); free(cv); free(pos);
    return 1;
}

typedef struct {
    int n;
    const int *lrp, *lci, *urp, *uci;
    const double *lv, *uv;
    double* y;
} IluFactors;

/* PCG_preconlr (PCGS.c:102-108): w = U^-1 L^-1 z by PCG_preconl (forward,
 * L's diagonal last in each row, PCGS.c:62-78) and PCG_preconr (backward,
 * U's diagonal first, PCGS.c:80-100) */
static void ilu_apply(IluFactors* F, const double* z, double* w) {
    int i, p;
    double s;
    for (i = 0; i < F->n; i++) {
        int last = F->lrp[i + 1] - 1;
        s = z[i];
        for (p = F->lrp[i]; p < last; p++) s -= F->lv[p] * F->y[F->lci[p]];
        s /= F->lv[last];
        F->y[i] = s;
    }
    for (i = F->n - 1; i >= 0; i--) {
        int first = F->urp[i];
        s = F->y[i];
        for (p = first + 1; p < F->urp[i + 1]; p++) s -= F->uv[p] * w[F->uci[p]];
        s /= F->uv[first];
        w[i] = s;
    }
}

/* ----------------------------------------------------------- GMRES(m)+ILU */

/* the plane rotation that zeroes t2 below t1 (gmres_qrfact, gmres.c:521-527
 * and 539-545); returns the rotated diagonal c t1 - s t2 */
static double givens(double t1, double t2, double* c, double* s) {
    double t;
    if (t2 == 0.0) {
        *c = 1.0;
        *s = 0.0;
    } else if (fabs(t2) >= fabs(t1)) {
        t = t1 / t2;
        *s = -1.0 / sqrt(1.0 + t * t);
        *c = -*s * t;
    } else {
        t = t2 / t1;
        *c = 1.0 / sqrt(1.0 + t * t);
        *s = -*c * t;
    }
    return *c * t1 - *s * t2;
}

/* gmres_qrfact with ijob 2 (gmres.c:530-548), or ijob 1 for k == 1: column
 * k-1 of the Hessenberg matrix (column-major, leading dimension ld) is rotated
 * by the k-1 earlier rotations, then a new one zeroes its subdiagonal.
 * Returns nonzero when the new diagonal is zero (never for k == 1, as the
 * reference's info = k there is 0). */
static int qr_add_column(int k, int ld, double* a, double* q) {
    int r;
    double* col = a + (size_t)ld * (k - 1);
    double t1, t2;
    for (r = 0; r < k - 1; r++) {
        double c = q[2 * r], s = q[2 * r + 1];
        t1 = col[r];
        t2 = col[r + 1];
        col[r] = c * t1 - s * t2;
        col[r + 1] = s * t1 + c * t2;
    }
    col[k - 1] = givens(col[k - 1], col[k], &q[2 * (k - 1)], &q[2 * (k - 1) + 1]);
    return col[k - 1] == 0.0 && k > 1;
}

/* gmres_qrsolv (gmres.c:550-574): rotate the right-hand side, then solve the
 * triangle column by column from the last */
static void qr_solve(int k, int ld, const double* a, const double* q, double* z) {
    int r, i;
    for (r = 0; r < k; r++) {
        double c = q[2 * r], s = q[2 * r + 1], t1 = z[r], t2 = z[r + 1];
        z[r] = c * t1 - s * t2;
        z[r + 1] = s * t1 + c * t2;
    }
    for (r = k - 1; r >= 0; r--) {
        double t;
        z[r] /= a[r + (size_t)ld * r];
        t = -z[r];
        for (i = 0; i < r; i++) z[i] += t * a[i + (size_t)ld * r];
    }
}

static double norm2(const double* x, int n) {
    int i;
    double s = 0.0;
    for (i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* FEM.c's GMRES path (FEM.c:591-611, 730-745): ILU(0) factors of A, then the
 * ZIB gmres (gmres.c:308-502) on A M^-1 y = b with preconr = PCG_preconlr,
 * preconl = identity, matvec = PCG_mult, CheckOnRestart, from y = x; on return
 * x = M^-1 y.  Parameter checks as itlin_parcheck_and_print (utils.c:345-365):
 * tol <= 0 is an error (-1 returned), tol is clamped to [1e-15, 0.1];
 * max_iter <= 0 means 100.  Returns info.iter (Arnoldi steps); *eta is
 * info.precision; -2 for a zero pivot in ILU(0), -3 when gmres_qrfact fails
 * (x is then M^-1 of the iterate reached at the last restart). */
int orc_gmres_ilu(int n, const int* rp, const int* ci, const double* v, double* x, const double* b, int m,
                  double tol, int max_iter, double* eta_out) {
    int nnz = rp[n], i, j, l, iter = 0, fail = 0;
    IluFactors F;
    int *lrp, *lci, *urp, *uci;
    double *lv, *uv, *V, *h, *z, *Hi, *r, *y0, *w, *q, *y;
    double beta, beta0 = 0, eta = 0, s;
    const size_t N = (size_t)(n > 0 ? n : 1);
    const int ld = m + 1;
    if (tol <= 0 || n <= 0 || m <= 0) return -1;
    if (tol > 1.0e-1) tol = 1.0e-1;
    else if (tol < 1.0e-15) tol = 1.0e-15;
    if (max_iter <= 0) max_iter = 100;
    lrp = (int*)malloc((n + 1) * sizeof(int));
    urp = (int*)malloc((n + 1) * sizeof(int));
    lci = (int*)malloc((nnz + n) * sizeof(int));
    uci = (int*)malloc((nnz + n) * sizeof(int));
    lv = (double*)malloc((nnz + n) * sizeof(double));
    uv = (double*)malloc((nnz + n) * sizeof(double));
    if (!orc_ilu0(n, rp, ci, v, lrp, lci, lv, urp, uci, uv)) {
        free(lrp); free(urp); free(lci); free(uci); free(lv); free(uv);
        return -2;
    }
    F.n = n; F.lrp = lrp; F.lci = lci; F.lv = lv; F.urp = urp; F.uci = uci; F.uv = uv;
    F.y = (double*)malloc(N * sizeof(double));
    V = (double*)calloc((size_t)(m + 2) * N, sizeof(double));
    h = (double*)calloc(m + 1, sizeof(double));
    z = (double*)calloc(m + 1, sizeof(double));
    Hi = (double*)calloc((size_t)(m + 1) * (m + 1), sizeof(double));
    q = (double*)calloc(2 * (m + 1), sizeof(double));
    r = (double*)malloc(N * sizeof(double));
    y0 = (double*)malloc(N * sizeof(double));
    w = (double*)malloc(N * sizeof(double));
    y = (double*)malloc(N * sizeof(double));
    memcpy(y, x, n * sizeof(double));
    for (;;) {
        /* restart: r = b - A M^-1 y */
        memcpy(y0, y, n * sizeof(double));
        ilu_apply(&F, y, V);
        orc_spmv(n, rp, ci, v, V, w);
        for (j = 0; j < n; j++) r[j] = b[j] - w[j];
        beta = norm2(r, n);
        if (iter == 0) {
            beta0 = beta;
            eta = beta0 == 0.0 ? 0.0 : 1.0;
        } else
            eta = beta / beta0;
        for (j = 0; j < n; j++) V[j] = r[j] / beta;
        for (i = 1; i <= m && eta > tol; i++) {
            double *vi = V + (size_t)(i - 1) * N, *vn = V + (size_t)i * N, nv;
            /* Arnoldi step, classical Gram-Schmidt against r = A M^-1 v_{i-1} */
            ilu_apply(&F, vi, vn);
            orc_spmv(n, rp, ci, v, vn, w);
            memcpy(r, w, n * sizeof(double));
            for (l = 0; l < i; l++) {
                const double* vl = V + (size_t)l * N;
                s = 0.0;
                for (j = 0; j < n; j++) s += vl[j] * r[j];
                h[l] = s;
            }
            for (j = 0; j < n; j++) {
                s = 0.0;
                for (l = 0; l < i; l++) s += V[(size_t)l * N + j] * h[l];
                vn[j] = r[j] - s;
            }
            nv = norm2(vn, n);
            for (j = 0; j < n; j++) vn[j] /= nv;
            for (l = 0; l < i; l++) Hi[(size_t)ld * (i - 1) + l] = h[l];
            Hi[(size_t)ld * (i - 1) + i] = nv;
            if (qr_add_column(i, ld, Hi, q)) {
                fail = 1;
                break;
            }
            if (i == m) {
                /* least-squares update at the cycle's end */
                for (l = 0; l < i + 1; l++) z[l] = 0.0;
                z[0] = beta;
                qr_solve(i, ld, Hi, q, z);
                for (j = 0; j < n; j++) {
                    s = 0.0;
                    for (l = 0; l < i; l++) s += V[(size_t)l * N + j] * z[l];
                    y[j] = y0[j] + s;
                }
            }
            iter++;
        }
        if (fail || !(eta > tol && iter < max_iter)) break;
    }
    ilu_apply(&F, y, x);
    if (eta_out) *eta_out = eta;
    free(lrp); free(urp); free(lci); free(uci); free(lv); free(uv); free(F.y);
    free(V); free(h); free(z); free(Hi); free(q); free(r); free(y0); free(w); free(y);
    return fail ? -3 : iter;
}

/* -------------------------------------------------------- BiCGStab + AMG */

/* BiCGStab_solve (BiCGStab.c:115-198) with Precon = BI_AMG_precon
 * (BiCGStab.c:79-81: AMG_solve warm-started from the previous PP / PS) and
 * Mat_mult = the caller's matrix (rp, ci, v).  The hierarchy h is built as
 * BI_set_AMG_preconditioner does (BiCGStab.c:98-113: smiter 2, aggressive
 * coarsening, matrix tolerance 0.1) and AMG_solve stops at amg_eps / amg_max_iter
 * (1e-4 / 20 there).  Returns the iterations; *res = |R| / |R0| at the end
 * (0 when R0 = 0). */
int orc_bicgstab_amg(orc_amg* hh, int n, const int* rp, const int* ci, const double* v, double* X, const double* b,
                     int sm, double w, double amg_eps, int amg_max_iter, double eps, int max_iter, double* res) {
    const size_t N = (size_t)(n > 0 ? n : 1);
    double *S = (double*)calloc(N, sizeof(double)), *PS = (double*)calloc(N, sizeof(double));
    double *PP = (double*)calloc(N, sizeof(double)), *APP = (double*)calloc(N, sizeof(double));
    double *APS = (double*)calloc(N, sizeof(double)), *R = (double*)calloc(N, sizeof(double));
    double *R0 = (double*)malloc(N * sizeof(double)), *P = (double*)malloc(N * sizeof(double));
    double hdot, t, s, r = 0, r0, alpha, beta, omega;
    int i, iter = 0;
    /* R = -(A X - b) */
    orc_spmv(n, rp, ci, v, X, R);
    for (i = 0; i < n; i++) R[i] += -1. * b[i];
    for (i = 0; i < n; i++) R[i] *= -1;
    memcpy(R0, R, n * sizeof(double));
    memcpy(P, R, n * sizeof(double));
    r0 = sqrt(dot(R, R, n));
    if (r0 != 0) {
        do {
            orc_amg_solve(hh, P, PP, sm, w, amg_eps, amg_max_iter);
            orc_spmv(n, rp, ci, v, PP, APP);
            hdot = dot(R, R0, n);
            t = dot(APP, R0, n);
            if (t == 0) break;
            alpha = hdot / t;
            memcpy(S, R, n * sizeof(double));
            for (i = 0; i < n; i++) S[i] += -alpha * APP[i];
            orc_amg_solve(hh, S, PS, sm, w, amg_eps, amg_max_iter);
            orc_spmv(n, rp, ci, v, PS, APS);
            s = dot(APS, APS, n);
            if (s == 0) break;
            omega = dot(APS, S, n) / s;
            for (i = 0; i < n; i++) X[i] += alpha * PP[i];
            for (i = 0; i < n; i++) X[i] += omega * PS[i];
            memcpy(R, S, n * sizeof(double));
            for (i = 0; i < n; i++) R[i] += -omega * APS[i];
            beta = dot(R, R0, n) * alpha / (hdot * omega);
            for (i = 0; i < n; i++) P[i] += -omega * APP[i];
            for (i = 0; i < n; i++) P[i] *= beta;
            for (i = 0; i < n; i++) P[i] += 1. * R[i];
            r = sqrt(dot(R, R, n));
            iter++;
        } while (r / r0 > eps && iter <= max_iter);
    }
    *res = r0 != 0 ? r / r0 : 0;
    free(S); free(PS); free(PP); free(APP); free(APS); free(R); free(R0); free(P);
    return iter;
}

/* pcg (PCGS.c:163-233) as PCG_default sets it up (PCGS.c:29-34): operator
 * PCG_mult on A, preconditioner PCG_preconlr with the ILU(0) factors of A.
 * Returns iterations (-2 for a zero pivot); *res = get_pcg_residuum.  B and X
 * are scaled by 1/|B| and back in place, as the reference does (and left
 * scaled when the preconditioned initial residual is zero). */
int orc_pcg_ilu(int n, const int* rp, const int* ci, const double* v, double* X, double* B, double tol, int max_iter,
                double* res) {
    int nnz = rp[n], i, it = 0;
    IluFactors F;
    int *lrp, *lci, *urp, *uci;
    double *lv, *uv, *R, *P, *W, *Q, *Z;
    double hn, s = 0, s_old, a, b, r, r0;
    const size_t N = (size_t)(n > 0 ? n : 1);
    lrp = (int*)malloc((n + 1) * sizeof(int));
    urp = (int*)malloc((n + 1) * sizeof(int));
    lci = (int*)malloc((nnz + n) * sizeof(int));
    uci = (int*)malloc((nnz + n) * sizeof(int));
    lv = (double*)malloc((nnz + n) * sizeof(double));
    uv = (double*)malloc((nnz + n) * sizeof(double));
    if (!orc_ilu0(n, rp, ci, v, lrp, lci, lv, urp, uci, uv)) {
        free(lrp); free(urp); free(lci); free(uci); free(lv); free(uv);
        return -2;
    }
    F.n = n; F.lrp = lrp; F.lci = lci; F.lv = lv; F.urp = urp; F.uci = uci; F.uv = uv;
    F.y = (double*)malloc(N * sizeof(double));
    W = (double*)calloc(N, sizeof(double));
    Q = (double*)calloc(N, sizeof(double));
    Z = (double*)calloc(N, sizeof(double));
    P = (double*)calloc(N, sizeof(double));
    R = (double*)malloc(N * sizeof(double));
    hn = sqrt(dot(B, B, n));
    if (hn == 0) hn = 1;
    for (i = 0; i < n; i++) B[i] *= (double)1 / hn;
    for (i = 0; i < n; i++) X[i] *= (double)1 / hn;
    memcpy(R, B, n * sizeof(double));
    orc_spmv(n, rp, ci, v, X, Q);
    for (i = 0; i < n; i++) R[i] += -1. * Q[i];
    ilu_apply(&F, R, W);
    r0 = sqrt(dot(W, W, n));
    if (r0 != 0) {
        do {
            ilu_apply(&F, R, Z);
            s_old = s;
            s = dot(R, Z, n);
            if (it > 0) {
                b = s / s_old;
                for (i = 0; i < n; i++) P[i] *= b;
                for (i = 0; i < n; i++) P[i] += 1 * Z[i];
            } else
                memcpy(P, Z, n * sizeof(double));
            orc_spmv(n, rp, ci, v, P, Q);
            a = s / dot(P, Q, n);
            for (i = 0; i < n; i++) X[i] += a * P[i];
            for (i = 0; i < n; i++) R[i] += -a * Q[i];
            r = sqrt(dot(R, R, n));
            it++;
        } while (r / r0 > tol && it < max_iter);
        ilu_apply(&F, R, W);
        r = sqrt(dot(W, W, n));
        for (i = 0; i < n; i++) B[i] *= hn;
        for (i = 0; i < n; i++) X[i] *= hn;
        *res = r / r0;
    } else
        *res = 0;
    free(lrp); free(urp); free(lci); free(uci); free(lv); free(uv); free(F.y);
    free(W); free(Q); free(Z); free(P); free(R);
    return it;
}

/* PCG_preconlr (PCGS.c:102-108) with given factors: w = U^-1 L^-1 z */
void orc_ilu_apply(int n, const int* lrp, const int* lci, const double* lv, const int* urp, const int* uci,
                   const double* uv, const double* z, double* w) {
    IluFactors F;
    F.n = n; F.lrp = lrp; F.lci = lci; F.lv = lv; F.urp = urp; F.uci = uci; F.uv = uv;
    F.y = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    ilu_apply(&F, z, w);
    free(F.y);
}

/* ------------------------------------------------------ re-assembly (f1)
 *
 * FEM2D.c:8504-8532 (get_map) fixes for every node k, triple m of its row
 * (Indices1 = n b + j, Indices2 = n d + k') and (a, c) in the order (0,0),
 * (0,1), (1,1), (1,0) the target (n a + j, n c + k'); :8534-8600 (use_map)
 * starts from a zero matrix of that pattern and adds C[l] * p in the same
 * order, l = n (4b + d) + k, then l += 2n, l += 8n, l -= 2n.  Restated
 * with the index l written out as n (8a + 4b + 2c + d) + k. */
static int row_position(const int* rowptr, const int* col, int r, int c) {
    int lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (col[mid] == c) return mid;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

int orc_rank4_assembly(int n, const int* bptr, const int* bi1, const int* bi2, const double* bval, const double* C,
                       const int* rowptr, const int* col, double* val) {
    static const int ac[4][2] = {{0, 0}, {0, 1}, {1, 1}, {1, 0}};
    int k, q, s2, e;
    for (e = 0; e < rowptr[2 * n]; e++) val[e] = 0.0;
    for (k = 0; k < n; k++)
        for (q = bptr[k]; q < bptr[k + 1]; q++) {
            const double p = bval[q];
            const int b = bi1[q] < n ? 0 : 1, d = bi2[q] < n ? 0 : 1;
            const int j = bi1[q] % n, kk = bi2[q] % n;
            for (s2 = 0; s2 < 4; s2++) {
                const int a = ac[s2][0], c = ac[s2][1];
                const int l = planted == 3 ? n * (8 * a + 4 * d + 2 * c + b) + k : n * (8 * a + 4 * b + 2 * c + d) + k;
                const int pos = row_position(rowptr, col, n * a + j, n * c + kk);
                if (pos < 0) return -1;
                val[pos] += C[l] * p;
            }
        }
    return 0;
}

void orc_dirichlet_symmetric(int n, const int* rowptr, const int* col, const unsigned char* dir, const double* u,
                             double* val, double* b) {
    int r, p;
    for (r = 0; r < n; r++) {
        double s = 0.0;
        int any = 0;
        if (dir[r]) {
            for (p = rowptr[r]; p < rowptr[r + 1]; p++) val[p] = col[p] == r ? 1.0 : 0.0;
            if (b) b[r] = u[r];
            continue;
        }
        for (p = rowptr[r]; p < rowptr[r + 1]; p++)
            if (dir[col[p]]) {
                s += val[p] * u[col[p]];
                val[p] = 0.0;
                any = 1;
            }
        if (b && any && planted != 4) b[r] = b[r] - s;
    }
}
