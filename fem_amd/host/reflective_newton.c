/* reflective_newton and its globals (LinA/reflective_newton.c) in the LinA
 * drop-in: the method runs on the GPU through liblina_hip's lina_rn handle
 * for a quadratic q(x) = x.H.x/2 + x.b + c -- the reference's default
 * functional, or one given by callbacks that is quadratic (checked); the
 * local solver of RN_set_local_solver is called on host copies; the trust
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

/* RN_UMF's local solver (reflective_newton.c:307-310) through the device
 * method's LINA_RN_HOST slot: M arrives as CSR and goes to the caller's
 * solver as a row list */
static int rn_host_trampoline(void* ctx, int n, const int* rp, const int* ci, const double* v, double* x,
                              const double* b) {
    int (*solver)(sparse_matrix*, double*, double*) = (int (*)(sparse_matrix*, double*, double*))ctx;
    sparse_matrix* M = sparse_zero(n);
    double* bb = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    int i, rc;
    for (i = 0; i < n; i++) {
        const int m = rp[i + 1] - rp[i];
        free(M->Values[i]);
        free(M->Indices[i]);
        M->Values[i] = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
        M->Indices[i] = (int*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int));
        memcpy(M->Values[i], v + rp[i], (size_t)m * sizeof(double));
        memcpy(M->Indices[i], ci + rp[i], (size_t)m * sizeof(int));
        M->Len[i] = m;
    }
    memcpy(bb, b, (size_t)n * sizeof(double));
    rc = solver(M, x, bb);
    Free_sparse(M);
    free(bb);
    return rc;
}

/* A functional given by callbacks (set_RN_functional, RN_set_gradient,
 * RN_set_Hesse_function, as Crack's damage step sets them, crack.c:9218-9220)
 * runs on the device when it is the quadratic q(x) = x.H.x/2 + b.x + c: H is
 * the Hessian callback's matrix at X, b = gradient(X) - H X, c = F(X) -
 * X.H.X/2 - b.X.  Checked at a second point (X moved on every row that is
 * not a unit row): the Hessian must be the same, the gradient H X' + b and
 * the functional q(X') + c to 1e-9 relative.  Crack's damage functional is
 * such a quadratic in the damage field (its weight g(h) = h (s + (1-s) h)
 * and every other term are polynomials of degree <= 2); its gradient sets
 * unit-row entries to X_i, which the unit rows of its Hessian reproduce. */
static sparse_matrix* rn_capture_H = NULL;
static double* rn_capture_b = NULL;

static double quad_value(sparse_matrix* H, const double* b, const double* x, int n) {
    int i, j;
    double xhx = 0, bx = 0;
    for (i = 0; i < n; i++) {
        double s = 0;
        for (j = 0; j < H->Len[i]; j++) s += H->Values[i][j] * x[H->Indices[i][j]];
        xhx += x[i] * s;
        bx += b[i] * x[i];
    }
    return xhx / 2. + bx;
}

static int capture_quadratic(double* X, int n, double* offset) {
    sparse_matrix* H;
    double *G, *Hx, *X2, *G2;
    double F, F2, q2, c, gscale = 0, gerr = 0;
    int i, j, ok = 1;
    H = RN_compute_Hesse ? RN_compute_Hesse(X) : clone(RN_Hesse);
    if (!H || H->size != n) {
        fprintf(stderr, "liblinalg_hip: reflective_newton: no Hessian of size %d\n", n);
        return -1;
    }
    Hx = sparse_mult(H, X);
    if (RN_gradient) G = RN_gradient(X);
    else {
        G = clone_vector(Hx, n);
        if (RN_b) vector_add(G, RN_b, 1., n);
    }
    rn_capture_b = (double*)malloc((size_t)n * sizeof(double));
    for (i = 0; i < n; i++) rn_capture_b[i] = G[i] - Hx[i];
    F = RN_Functional ? RN_Functional(X, n) : quad_value(H, rn_capture_b, X, n);
    c = F - quad_value(H, rn_capture_b, X, n);
    /* the second point */
    X2 = clone_vector(X, n);
    for (i = 0; i < n; i++) {
        const int unit = H->Len[i] == 1 && H->Indices[i][0] == i && H->Values[i][0] == 1.0;
        if (!unit) X2[i] += 1e-3 * (1.0 + fabs(X[i])) * sin(0.7 * i + 0.3);
    }
    if (RN_compute_Hesse) {
        sparse_matrix* H2 = RN_compute_Hesse(X2);
        double hs = 0, he = 0;
        if (!H2 || H2->size != n) ok = 0;
        for (i = 0; ok && i < n; i++) {
            if (H2->Len[i] != H->Len[i]) {
                ok = 0;
                break;
            }
            for (j = 0; j < H->Len[i]; j++) {
                if (H2->Indices[i][j] != H->Indices[i][j]) ok = 0;
                hs = fmax(hs, fabs(H->Values[i][j]));
                he = fmax(he, fabs(H2->Values[i][j] - H->Values[i][j]));
            }
        }
        if (ok && he > 1e-12 * hs) ok = 0;
        if (H2) {
            free_sparse(H2);
        }
    }
    if (ok && RN_gradient) {
        double* H2x = sparse_mult(H, X2);
        G2 = RN_gradient(X2);
        for (i = 0; i < n; i++) {
            const double g = H2x[i] + rn_capture_b[i];
            gscale = fmax(gscale, fabs(g));
            gerr = fmax(gerr, fabs(G2[i] - g));
        }
        if (gerr > 1e-9 * (gscale + 1e-300)) ok = 0;
        free(G2);
        free(H2x);
    }
    if (ok && RN_Functional) {
        F2 = RN_Functional(X2, n);
        q2 = quad_value(H, rn_capture_b, X2, n) + c;
        if (fabs(F2 - q2) > 1e-9 * (fabs(F2) + fabs(F) + 1e-300)) ok = 0;
    }
    free(X2);
    free(G);
    free(Hx);
    if (!ok) {
        fprintf(stderr, "liblinalg_hip: reflective_newton: the functional of the callbacks is not a quadratic "
                        "x.H.x/2 + b.x + c; the device method needs one\n");
        free_sparse(H);
        free(rn_capture_b);
        rn_capture_b = NULL;
        return -1;
    }
    rn_capture_H = H;
    *offset = c;
    return 0;
}

int reflective_newton(double* X) {
    lina_rn_params p;
    lina_rn_stats st;
    dropin_csr c;
    lina_rn* h;
    sparse_matrix* Hm;
    double* bvec;
    double offset = 0;
    int it, dof = 2, depth = 6;
    const int callbacks = RN_Functional || RN_gradient || RN_compute_Hesse;
    if (callbacks) {
        const int n = RN_compute_Hesse ? RN_dim : (RN_Hesse ? RN_Hesse->size : RN_dim);
        if (capture_quadratic(X, n, &offset) < 0) return -1;
        Hm = rn_capture_H;
        bvec = rn_capture_b;
    } else {
        if (!RN_Hesse || !RN_b) {
            fprintf(stderr, "liblinalg_hip: reflective_newton needs set_RN_Hesse and set_RN_linear\n");
            return -1;
        }
        Hm = RN_Hesse;
        bvec = RN_b;
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
    p.q_offset = offset;
    it = 0;
    switch (RN_solver_type) {
        case RN_PCG: p.solver = LINA_RN_PCG; break;
        case RN_AMG: p.solver = LINA_RN_AMG; break;
        case RN_BICGSTAB: p.solver = LINA_RN_BICGSTAB; break;
        case RN_GMRES: p.solver = LINA_RN_GMRES; break;
        case RN_UMF:
            if (!RN_local_solver) {
                p.solver = LINA_RN_GMRES;
                break;
            }
            p.solver = LINA_RN_HOST;
            p.host_solver = rn_host_trampoline;
            p.host_ctx = (void*)RN_local_solver;
            break;
        default:
            fprintf(stderr, "liblinalg_hip: reflective_newton inner solver %d is not built (use RN_PCG, RN_AMG, "
                            "RN_BICGSTAB, RN_GMRES or RN_set_local_solver)\n", RN_solver_type);
            it = -1;
    }
    if (it == 0) {
        c = dropin_to_csr(Hm);
        h = lina_rn_create(c.n, c.rp, c.ci, c.v, bvec, RN_bound_num, RN_bound, RN_lower, RN_upper, &p);
        dropin_free_csr(&c);
        if (!h) {
            dropin_check(-1, "reflective_newton setup");
            it = -1;
        } else {
            it = lina_rn_solve(h, X, &st);
            dropin_check(it, "reflective_newton");
            lina_rn_destroy(h);
        }
    }
    if (callbacks) {
        free_sparse(rn_capture_H);
        free(rn_capture_b);
        rn_capture_H = NULL;
        rn_capture_b = NULL;
    }
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
