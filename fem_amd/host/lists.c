/* Vector, index-list and sorting utilities of the LinA drop-in (host C),
 * with the results of the functions of LinA/linear_algebra.c cited at each
 * definition; callers (FEM2D, Crack) use them on their own host arrays. */
#include <math.h>
#include <string.h>

#include "dropin.h"

/* linear_algebra.c:13-19: exported as C functions (callers use max) */
int min(int a, int b) { return a < b ? a : b; }
int max(int a, int b) { return a >= b ? a : b; }

int* clone_list(int* x, int n) {                                   /* :84-89 */
    int* r = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (n > 0) memcpy(r, x, (size_t)n * sizeof(int));
    return r;
}

double* generate_vector(int size, double val) {                    /* :4044-4049 */
    int i;
    double* r = (double*)malloc((size_t)(size > 0 ? size : 1) * sizeof(double));
    for (i = 0; i < size; i++) r[i] = val;
    return r;
}

int* get_ID_index_map(int size) {                                  /* :2622-2627 */
    int i;
    int* r = zero_int_list(size);
    for (i = 0; i < size; i++) r[i] = i;
    return r;
}

double get_max_element_abs(double* X, int size) {                  /* :2642-2649 */
    int i;
    double m = fabs(X[0]);
    for (i = 1; i < size; i++)
        if (fabs(X[i]) > m) m = fabs(X[i]);
    return m;
}

double get_mean(double* List, int size) {                          /* :4166-4172 */
    int i;
    double s = 0;
    for (i = 0; i < size; i++) s += List[i];
    return s / (double)size;
}

int in_list(int* List, int size, int n) {                          /* :1821-1827: first position, -1 */
    int i;
    for (i = 0; i < size; i++)
        if (List[i] == n) return i;
    return -1;
}

double* join_vectors(double* V1, int n1, double* V2, int n2) {     /* :3420-3426 */
    double* r = (double*)malloc((size_t)(n1 + n2 > 0 ? n1 + n2 : 1) * sizeof(double));
    if (n1 > 0) memcpy(r, V1, (size_t)n1 * sizeof(double));
    if (n2 > 0) memcpy(r + n1, V2, (size_t)n2 * sizeof(double));
    return r;
}

void map_back(double* Solution, double* Restricted, int n_sol, int* index_map) { /* :3398-3405 */
    int i;
    for (i = 0; i < n_sol; i++)
        if (index_map[i] >= 0) Solution[i] = Restricted[index_map[i]];
}

/* :3370-3380: Res[index_map[i]] = Vector[i] for the mapped entries (others
 * left as allocated) */
double* restrict_vector(double* Vector, int* index_map, int n_sol, int n_res) {
    int i;
    double* r = (double*)malloc((size_t)(n_res > 0 ? n_res : 1) * sizeof(double));
    for (i = 0; i < n_sol; i++)
        if (index_map[i] >= 0) {
            if (index_map[i] >= n_res) fprintf(stderr, "restrict_vector: index %d maps beyond %d\n", i, n_res);
            else r[index_map[i]] = Vector[i];
        }
    return r;
}

double* restrict_vector_by(int i_min, int i_max, double* Vector) { /* :3382-3389 */
    const int m = i_max - i_min;
    double* r = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    if (m > 0) memcpy(r, Vector + i_min, (size_t)m * sizeof(double));
    return r;
}

double vec_dist(double* x1, double* x2, int n) {                   /* :288-297: long double sum */
    int i;
    long double s = 0;
    for (i = 0; i < n; i++) {
        const long double d = (long double)(x1[i] - x2[i]);
        s = s + d * d;
    }
    return (double)sqrtl(s);
}

double vector_contraction(double* x, int n) {                      /* :344-349: sum */
    int i;
    double s = 0;
    for (i = 0; i < n; i++) s += x[i];
    return s;
}

void vector_normalize(double* x, int n) { scalar_mult((double)1 / euklid_norm(x, n), x, n); } /* :306-309 */

void vector_pseudo_mult(double* x, double* y, int n) {             /* :351-354 */
    int i;
    for (i = 0; i < n; i++) x[i] *= y[i];
}

void vector_pseudo_div(double* x, double* y, int n) {              /* :356-359 */
    int i;
    for (i = 0; i < n; i++) x[i] /= y[i];
}

void vector_shift(double* x, int n, double a) {                    /* :361-364 */
    int i;
    for (i = 0; i < n; i++) x[i] += a;
}

/* :1058-1067 writes to a fixed directory of the author's machine; here the
 * file "vector" in the working directory, one value per line (%.15e) */
void print_vector(double* V, int n) { print_scalar_data("vector", V, n); }

void print_scalar_data(char* Name, double* V, int n) {            /* :1069-1081 */
    int i;
    FILE* f;
    if (V == NULL) return;
    f = fopen(Name, "w");
    if (!f) {
        printf("could not write to file %s -> skip\n", Name);
        return;
    }
    for (i = 0; i < n; i++) fprintf(f, "%.15e\n", V[i]);
    fclose(f);
}

/* ------------------------------------------------------------- sorting */

static int cmp_int_asc(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

void int_sort(int* List, int size) {                               /* :1635-1653: ascending */
    if (size > 1) qsort(List, (size_t)size, sizeof(int), cmp_int_asc);
}

/* :1655-1687: the distinct values of L1 and L2, ascending, in a new list */
void join_lists(int** List, int* size, int* L1, int* L2, int s1, int s2) {
    const int t = s1 + s2;
    int i, w = 0;
    int* r = (int*)malloc((size_t)(t > 0 ? t : 1) * sizeof(int));
    if (s1 > 0) memcpy(r, L1, (size_t)s1 * sizeof(int));
    if (s2 > 0) memcpy(r + s1, L2, (size_t)s2 * sizeof(int));
    int_sort(r, t);
    for (i = 0; i < t; i++)
        if (w == 0 || r[i] != r[w - 1]) r[w++] = r[i];
    *List = (int*)realloc(r, (size_t)(w > 0 ? w : 1) * sizeof(int));
    *size = w;
}

/* :1591-1622: List (indices into Values) reordered so that Compare(Values[a],
 * Values[b]) > 0 never holds for a before b.  The order among elements that
 * do not compare greater is LinA's: its recursion takes the first index of a
 * range as the pivot and keeps the other indices, in their order, to the
 * left when they do not compare greater than it and to the right otherwise,
 * the pivot between them.  The same partitions are formed here, range by
 * range from an explicit stack, with one scratch array. */
void divide_and_conquer_sort(int* List, void** Values, int size, int (*Compare)(void* X1, void* X2)) {
    int* scratch;
    int* stack;
    int top = 0;
    if (size <= 1) return;
    scratch = (int*)malloc((size_t)size * sizeof(int));
    stack = (int*)malloc((size_t)2 * (size + 1) * sizeof(int));
    stack[top++] = 0;
    stack[top++] = size;
    while (top > 0) {
        const int end = stack[--top], begin = stack[--top];
        int i, left = 0, right = 0;
        const int pivot = List[begin];
        const int n = end - begin;
        if (n <= 1) continue;
        for (i = begin + 1; i < end; i++)
            if (Compare(Values[List[i]], Values[pivot]) > 0) scratch[n - 1 - right++] = List[i];
            else scratch[left++] = List[i];
        /* left part in order, the pivot, right part (collected back to front) in order */
        memcpy(List + begin, scratch, (size_t)left * sizeof(int));
        List[begin + left] = pivot;
        for (i = 0; i < right; i++) List[begin + left + 1 + i] = scratch[n - 1 - i];
        stack[top++] = begin;
        stack[top++] = begin + left;
        stack[top++] = begin + left + 1;
        stack[top++] = end;
    }
    free(scratch);
    free(stack);
}

void insert_vector(double* Vector, double* vector, int start, int size) { /* :3391-3396 */
    if (size > 0) memcpy(Vector + start, vector, (size_t)size * sizeof(double));
}

int test_vector(double* X, int n, char* message) { /* :1881-1888: 1 (and message) on a NaN */
    int i;
    for (i = 0; i < n; i++)
        if (isnan(X[i])) {
            printf("%s\n", message);
            fflush(stdout);
            return 1;
        }
    return 0;
}

int test_negative(double* X, int n) { /* :1890-1897 */
    int i;
    for (i = 0; i < n; i++)
        if (X[i] < 0) {
            printf("negative component at index %d\n", i);
            return 1;
        }
    return 0;
}

void remove_negative(double* X, int n) { /* :1899-1902 */
    int i;
    for (i = 0; i < n; i++)
        if (X[i] < 0) X[i] = 0;
}

void vector_permutation(double* F, int* Map, int n) { /* :2857-2862: F[i] = F_old[Map[i]] */
    int i;
    double* old = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    if (n > 0) memcpy(old, F, (size_t)n * sizeof(double));
    for (i = 0; i < n; i++) F[i] = old[Map[i]];
    free(old);
}
