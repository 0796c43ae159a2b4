/* BiCGStab_solve and its globals (LinA/BiCGStab.c) in the LinA drop-in:
 * the reference's iteration (BiCGStab.c:115-198) on device vectors, the AMG
 * preconditioner and the library's operators (RN_default_mat_mult,
 * PCG_mult) on the GPU, caller callbacks on host copies. */
#include "../../include/linalg/BiCGStab.h"

#include "../../include/linalg/PCGS.h"
#include "../../include/linalg/reflective_newton.h"
#include "dropin.h"

static int (*Precon)(int n, double* Y, double* X) = NULL;
static void (*Mat_mult)(int n, double* Y, double* X) = NULL;
static amg_system_info* BI_AMG_data = NULL;
static amg_statistics BI_AMG_stat, BI_stat;
static int BI_max_iter = 1000;
static double BI_eps = 1e-8;
static int BI_info = YES;

void BI_set_precon(int (*P)(int n, double* Y, double* X)) { Precon = P; }
void BI_set_mat_mult(void (*M)(int n, double* Y, double* X)) { Mat_mult = M; }
amg_system_info* BI_get_AMG_setup_data(void) { return BI_AMG_data; }
void BI_set_AMG_setup_data(amg_system_info* d) { BI_AMG_data = d; }
void BI_set_eps(double eps) { BI_eps = eps; }
void BI_set_max_iter(int m) { BI_max_iter = m; }
int BI_get_max_iter(void) { return BI_max_iter; }
void BI_print_info(int mode) { BI_info = mode; }
int BI_info_set(void) { return BI_info; }
void BI_set_stat(char* Out_AMG, char* Out_BiCGStab) {
    BI_AMG_stat = AMG_init_stat(Out_AMG);
    BI_stat = AMG_init_stat(Out_BiCGStab);
}
amg_statistics* BI_get_AMG_stat(void) { return &BI_AMG_stat; }
amg_statistics* BI_get_stat(void) { return &BI_stat; }

int BI_AMG_precon(int n, double* Y, double* X) {
    (void)n;
    return AMG_solve(BI_AMG_data, Y, X, &BI_AMG_stat);
}

int BI_PCG_precon(int n, double* Y, double* X) {
    const double eps = 1e-2;
    const int max_iter = 1000;
    int i = pcg(X, Y, n, eps, max_iter);
    if (i == max_iter) printf("warning PCG preconditioner could not solve to required residuum reduction\n");
    return i;
}

void BI_set_PCG_preconditioner(sparse_matrix* A) {
    int n = A->size;
    sparse_matrix* L = sparse_zero(n);
    sparse_matrix* U = sparse_zero(n);
    incomplete_LU_factorization(A, L, U);
    PCG_default(A, L, U);
    PCG_set_default_precon();
    BI_set_precon(&BI_PCG_precon);
}

void BI_set_AMG_preconditioner(sparse_matrix* A, int num_fields, int depth) {
    AMG_print_info(NO);
    AMG_set_matrix_tolerance(1e-1);
    AMG_set_max_iter(20);
    AMG_set_eps(1e-4);
    BI_AMG_data = AMG_setup(A, num_fields, depth, 2, 3, AGGRESSIVE, STAND_ALONE, DIRECT, NULL, NULL);
    BI_set_precon(&BI_AMG_precon);
}

double BI_residuum(double* X, double* b, int n) {
    double* R = zero_vector(n);
    double r;
    (*Mat_mult)(n, X, R);
    vector_add(R, b, -1., n);
    scalar_mult(-1, R, n);
    r = euklid_norm(R, n);
    free(R);
    return r;
}

int BiCGStab_solve(double* X, double* b, int n) {
    const double default_eps = 1e-10;
    dropin_op op = {NULL, NULL};
    dropin_prec pr = {NULL, NULL, NULL};
    dropin_stage sg;
    double *dX, *db, *dS, *dPS, *dPP, *dAPP, *dAPS, *dR, *dR0, *dP;
    double h, s, r = 0, r0, t, alpha, beta, omega;
    int iter = 0;
    if (Mat_mult == RN_default_mat_mult) op.dev = dropin_rn_default_matrix();
    else if (Mat_mult == PCG_mult) op.dev = dropin_pcg_default_matrix();
    else op.host = Mat_mult;
    if (Precon == BI_AMG_precon && BI_AMG_data) {
        dropin_amg_sync(BI_AMG_data);
        pr.amg = (lina_solver*)BI_AMG_data->device;
    } else {
        pr.host = Precon;
    }
    if ((!op.dev && !op.host) || (!pr.amg && !pr.host)) {
        fprintf(stderr, "liblinalg_hip: BiCGStab_solve without operator or preconditioner\n");
        return -1;
    }
    dropin_stage_init(&sg, n);
    dX = dropin_dalloc(n);
    db = dropin_dalloc(n);
    dS = dropin_dalloc(n);
    dPS = dropin_dalloc(n);
    dPP = dropin_dalloc(n);
    dAPP = dropin_dalloc(n);
    dAPS = dropin_dalloc(n);
    dR = dropin_dalloc(n);
    dR0 = dropin_dalloc(n);
    dP = dropin_dalloc(n);
    dropin_h2d(dX, X, n);
    dropin_h2d(db, b, n);
    /* R = -(A X - b), R0 = P = R */
    dropin_op_apply(&op, &sg, dX, dR);
    lina_dev_axpby(n, 1.0, dR, -1., db);
    lina_dev_scale(n, dR, -1);
    dropin_d2d(dR0, dR, n);
    dropin_d2d(dP, dR, n);
    r0 = sqrt(dropin_dot(dR, dR, n));
    if (r0 == 0 && BI_info == BI_YES) {
        printf("initial residuum zero -> continue ...\n");
    } else {
        do {
            dropin_prec_apply(&pr, &sg, dP, dPP);
            dropin_op_apply(&op, &sg, dPP, dAPP);
            h = dropin_dot(dR, dR0, n);
            t = dropin_dot(dAPP, dR0, n);
            if (t == 0) {
                if (fabs(r0) > default_eps) printf("Warning: BiCGStab stopped because alpha->infinity\n");
                break;
            }
            alpha = h / t;
            dropin_d2d(dS, dR, n);
            lina_dev_axpby(n, 1.0, dS, -alpha, dAPP);
            dropin_prec_apply(&pr, &sg, dS, dPS);
            dropin_op_apply(&op, &sg, dPS, dAPS);
            s = dropin_dot(dAPS, dAPS, n);
            if (s == 0) {
                if (fabs(r0) > default_eps) printf("Warning: BiCGStab stopped because APS=0\n");
                break;
            }
            omega = dropin_dot(dAPS, dS, n) / s;
            lina_dev_axpby(n, 1.0, dX, alpha, dPP);
            lina_dev_axpby(n, 1.0, dX, omega, dPS);
            dropin_d2d(dR, dS, n);
            lina_dev_axpby(n, 1.0, dR, -omega, dAPS);
            beta = dropin_dot(dR, dR0, n) * alpha / (h * omega);
            lina_dev_axpby(n, 1.0, dP, -omega, dAPP);
            lina_dev_scale(n, dP, beta);
            lina_dev_axpby(n, 1.0, dP, 1., dR);
            r = sqrt(dropin_dot(dR, dR, n));
            iter++;
        } while (r / r0 > BI_eps && iter <= BI_max_iter);
    }
    if (iter == BI_max_iter) printf("BiCGStab: no convergence in %d iterations\n\n", iter);
    AMG_add_stat(&BI_stat, iter, fabs(r / r0));
    dropin_d2h(X, dX, n);
    dropin_dfree(dX);
    dropin_dfree(db);
    dropin_dfree(dS);
    dropin_dfree(dPS);
    dropin_dfree(dPP);
    dropin_dfree(dAPP);
    dropin_dfree(dAPS);
    dropin_dfree(dR);
    dropin_dfree(dR0);
    dropin_dfree(dP);
    dropin_stage_free(&sg);
    return iter;
}
