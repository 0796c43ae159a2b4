/* reflective_newton and its globals (LinA/reflective_newton.c) in the LinA
 * drop-in: the method runs on the GPU through liblina_hip's lina_rn handle
 * for the reference's default quadratic q(x) = x.H.x/2 + x.b; the trust
 * radius persists between calls like the reference's RN_Delta. */
#include "../../include/linalg/reflective_newton.h"

#include <string.h>
#include <time.h>

#include "dropin.h"

/* defaults of reflective_newton.c:6-25 */
static int RN_max_iter = 50;
static double RN_eps = 1e-8, RN_sigma_l = 0.1, RN_sigma_u = 0.9, RN_rho = 0.1, RN_Delta = 250., RN_Xi = 1.;
static double RN_trust_lower = 0.6, RN_trust_upper = 1.1, RN_trust_radius_min = 1e-3, RN_trust_radius_max = 1000.;
static int RN_solver_type = RN_BICGSTAB;
static int RN_info = RN_YES;
static int RN_dim = 0;
static int* RN_bound = NULL;
static int RN_bound_num = 0;
static double *RN_lower = NULL, *RN_upper = NULL;
static sparse_matrix* RN_Hesse = NULL;
static sparse_matrix* RN_M = NULL;
static lina_matrix* RN_M_dev = NULL;
static double* RN_b = NULL;
static double (*RN_Functional)(double* X, int n) = NULL;
static double* (*RN_gradient)(double* X) = NULL;
static sparse_matrix* (*RN_compute_Hesse)(double* X) = NULL;
static int (*RN_local_solver)(sparse_matrix* A, double* X, double* b) = NULL;
static amg_system_info* RN_AMG_setup_data = NULL;
static RN_statistics statistics;

lina_matrix* dropin_rn_default_matrix(void) { return RN_M_dev; }

void RN_set_params(int max_iter, double eps, double sigma_l, double sigma_u, double rho, double Delta, double Xi,
                   double trust_lower, double trust_upper, double radius_min, double radius_max) {
    RN_max_iter = max_iter;
    RN_eps = eps;
    RN_sigma_l = sigma_l;
    RN_sigma_u = sigma_u;
    RN_rho = rho;
    RN_Delta = Delta;
    RN_Xi = Xi;
    RN_trust_lower = trust_lower;
    RN_trust_upper = trust_upper;
    RN_trust_radius_min = radius_min;
    RN_trust_radius_max = radius_max;
}
void set_RN_params(int max_iter, double eps, double sigma_l, double sigma_u) {
    RN_max_iter = max_iter;
    RN_eps = eps;
    RN_sigma_l = sigma_l;
    RN_sigma_u = sigma_u;
}
void set_RN_constraints(int* Index_set, double* Lower, double* Upper, int n) {
    RN_bound = Index_set;
    RN_bound_num = n;
    RN_lower = Lower;
    RN_upper = Upper;
}
void set_RN_Hesse(sparse_matrix* A) {
    RN_Hesse = A;
    if (A) RN_dim = A->size;
}
void set_RN_linear(double* B) { RN_b = B; }
void set_RN_functional(double (*F)(double* X, int n)) { RN_Functional = F; }
void RN_set_gradient(double* (*G)(double* X)) { RN_gradient = G; }
void RN_set_Hesse_function(sparse_matrix* (*H)(double* X), int dim) {
    RN_compute_Hesse = H;
    RN_dim = dim;
    RN_Hesse = NULL;
}
void RN_set_default_matrix(sparse_matrix* M) {
    RN_M = M;
    if (RN_M_dev) lina_matrix_destroy(RN_M_dev);
    RN_M_dev = M ? dropin_device_matrix(M) : NULL;
}
void RN_default_mat_mult(int n, double* y, double* z) {
    double *dy, *dz;
    if (!RN_M_dev) {
        fprintf(stderr, "liblinalg_hip: RN_default_mat_mult without RN_set_default_matrix\n");
        return;
    }
    dy = dropin_dalloc(n);
    dz = dropin_dalloc(n);
    dropin_h2d(dy, y, n);
    dropin_check(lina_matrix_spmv_device(RN_M_dev, dy, dz), "RN_default_mat_mult");
    dropin_d2h(z, dz, n);
    dropin_dfree(dy);
    dropin_dfree(dz);
}
void RN_set_AMG_data(amg_system_info* AMG_info) { RN_AMG_setup_data = AMG_info; }
void RN_set_solver(int type) { RN_solver_type = type; }
void RN_set_local_solver(int (*S)(sparse_matrix* A, double* X, double* b)) {
    RN_local_solver = S;
    if (S) RN_solver_type = RN_UMF;
}
void RN_set_trust_region(double Delta) { RN_Delta = Delta; }
void RN_set_trust_limits(double mn, double mx) {
    RN_trust_radius_min = mn;
    RN_trust_radius_max = mx;
}
void RN_print_info(int mode) { RN_info = mode; }
int RN_info_set(void) { return RN_info; }
void RN_set_debug_mode(int mode) { (void)mode; }
void RN_set_kernel(double* K, int n) {
    (void)K;
    (void)n;
}
void RN_init_statistics(void) { memset(&statistics, 0, sizeof(statistics)); }
void print_RN_statistics(FILE* file) {
    if (!file) return;
    fprintf(file, "\nReflective Newton statistics:\n");
    if (statistics.calls)
        fprintf(file, "average number of Newton iterations: %f\n", (float)statistics.total_iter / statistics.calls);
    fprintf(file, "number of neg. curvature events: %d\n", statistics.number_neg_curv);
    fprintf(file, "line searches: %d\n", statistics.line_search_count);
}

int reflective_newton(double* X) {
    lina_rn_params p;
    lina_rn_stats st;
    dropin_csr c;
    lina_rn* h;
    int it, dof = 2, depth = 6;
    if (RN_Functional || RN_gradient || RN_compute_Hesse) {
        fprintf(stderr, "liblinalg_hip: reflective_newton runs the default quadratic functional only\n");
        return -1;
    }
    if (!RN_Hesse || !RN_b) {
        fprintf(stderr, "liblinalg_hip: reflective_newton needs set_RN_Hesse and set_RN_linear\n");
        return -1;
    }
    if (RN_AMG_setup_data) {
        dof = RN_AMG_setup_data->AMG_deg_of_freedom;
        depth = RN_AMG_setup_data->depth;
    }
    lina_rn_params_default(&p, dof, depth);
    p.max_iter = RN_max_iter;
    p.eps = RN_eps;
    p.sigma_l = RN_sigma_l;
    p.sigma_u = RN_sigma_u;
    p.rho = RN_rho;
    p.delta = RN_Delta;
    p.xi = RN_Xi;
    p.trust_lower = RN_trust_lower;
    p.trust_upper = RN_trust_upper;
    p.radius_min = RN_trust_radius_min;
    p.radius_max = RN_trust_radius_max;
    switch (RN_solver_type) {
        case RN_PCG: p.solver = LINA_RN_PCG; break;
        case RN_AMG: p.solver = LINA_RN_AMG; break;
        case RN_BICGSTAB: p.solver = LINA_RN_BICGSTAB; break;
        case RN_GMRES: p.solver = LINA_RN_GMRES; break;
        default:
            fprintf(stderr, "liblinalg_hip: reflective_newton inner solver %d is not built (use RN_PCG, RN_AMG, "
                            "RN_BICGSTAB or RN_GMRES)\n", RN_solver_type);
            return -1;
    }
    c = dropin_to_csr(RN_Hesse);
    h = lina_rn_create(c.n, c.rp, c.ci, c.v, RN_b, RN_bound_num, RN_bound, RN_lower, RN_upper, &p);
    dropin_free_csr(&c);
    if (!h) {
        dropin_check(-1, "reflective_newton setup");
        return -1;
    }
    it = lina_rn_solve(h, X, &st);
    dropin_check(it, "reflective_newton");
    lina_rn_destroy(h);
    if (it >= 0) {
        RN_Delta = st.delta;
        statistics.calls++;
        statistics.total_iter += it;
        statistics.res += st.dq;
        statistics.number_neg_curv += st.neg_curvature;
        statistics.line_search_count += st.line_searches;
        statistics.line_search_iter += st.line_search_iter;
        if (it == RN_max_iter) printf("Reflective Newton: no convergence after %d iterations\n", it);
    }
    return it;
}

/* reflective_newton.c:394-450 on the host: the coefficients of q(t) by
 * scalar and sparse_bilinear, the angle by liblina_hip's search over
 * [0, 2 pi] (lina_circle_minimum). */
void subspace_minimum(sparse_matrix* A, double* b, double* V1, double* V2, double* X) {
    const int n = A->size;
    double t = 0;
    const double bx = scalar(V1, b, n), by = scalar(V2, b, n);
    const double ax = sparse_bilinear(V1, A, V1), ay = sparse_bilinear(V2, A, V2);
    const double axy = sparse_bilinear(V1, A, V2) + sparse_bilinear(V2, A, V1);
    if (lina_circle_minimum(ax, ay, axy, bx, by, RN_Delta, &t) <= 0) {
        int i;
        printf("subspace_minimum: no extrema found (%s)\n", lina_last_error());
        for (i = 0; i < n; i++) X[i] = NAN;
        return;
    }
    copy_vector_to(V1, X, n);
    scalar_mult(RN_Delta * cos(t), X, n);
    vector_add(X, V2, RN_Delta * sin(t), n);
}

int* get_all_indices(int size) { return get_ID_index_map(size); } /* reflective_newton.c:64-69 */

/* reflective_newton.c:376-392 writes to a fixed directory of the author's
 * machine; here the file "func" in the working directory: n lines
 * "x<TAB>F(x)" for x evenly spaced over [min, max] */
void RN_print_function(double (*F)(double x), int n, double min, double max) {
    int i;
    FILE* f = fopen("func", "w");
    if (!f) {
        printf("could not open file func\n");
        return;
    }
    for (i = 0; i < n; i++) {
        const double x = min + (max - min) * ((double)i / (n - 1));
        fprintf(f, "%f\t%e\n", x, F(x));
    }
    fclose(f);
}

/* reflective_newton.c:719-757: n steps of a random search on the
 * functional of set_RN_functional: one component moved by +-0.002 inside
 * its bounds per step (every component must be bounded); a lower value is
 * kept with probability 1 - exp((curr - prev) / T), a higher one never; T
 * grows by the factor 1 + 10/n per step.  The random sequence starts from
 * the time like the reference's (whose step counter starts uninitialised;
 * here at 0); the result goes to print_vector as there. */
int simulated_annealing_min(double* X, int n) {
    int i = 0;
    const double dx = 2. / 1000;
    const double z = 10. / n;
    double T = 1e-4, prev;
    if (!RN_Functional || RN_bound_num != RN_dim) {
        printf("simulated_annealing_min: needs set_RN_functional and a bound on every component\n");
        return -1;
    }
    srand((unsigned)time(NULL));
    prev = RN_Functional(X, RN_dim);
    printf("start value: %f\n", prev);
    do {
        int j, sign;
        double curr, a, moved;
        do {
            j = rand() % RN_dim;
            sign = 1 - 2 * (rand() % 2);
            moved = X[j] + sign * dx;
        } while (moved < RN_lower[j] || moved > RN_upper[j]);
        X[j] = moved;
        curr = RN_Functional(X, RN_dim);
        a = exp((curr - prev) / T);
        if (a < 1) {
            if ((double)rand() / RAND_MAX < a) X[j] -= sign * dx;
            else prev = curr;
        } else {
            X[j] -= sign * dx;
        }
        i++;
        T *= 1. + z;
    } while (i < n);
    print_vector(X, RN_dim);
    printf("final value: %f\n", prev);
    return i;
}
