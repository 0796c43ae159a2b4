/* GPU solver callbacks in the form of Crack's UMFPACK drivers (host C of the
 * LinA drop-in).  Crack solves its elastic step with
 * multifrontal_solver(A, X, b, Caller, &time) (Crack/crack.c:5424-5525,
 * called at :9340) and hands multifrontal_solver_RN (:5664) to
 * RN_set_local_solver (:9207); both factor A with UMFPACK.  The systems are
 * SPD apart from the Dirichlet rows Set_dirichlet_BC (:6445-6459) resets to
 * unit rows, which leave their columns in the other rows.
 *
 * lina_spd_solver / lina_spd_local_solver take the same arguments and
 * return SUCCESS (1) or FAILED (0) like the drivers they replace:
 *   1. rows whose only nonzero is the diagonal are constraints
 *      x_i = b_i / a_ii; their columns move to the right-hand side
 *      (symmetric elimination, the pattern kept with zeros);
 *   2. the system goes to an AMG-PCG on the GPU; the hierarchy is cached and
 *      refreshed with AMG_recompute_coarse_matrices' values-only path while
 *      the pattern holds (the staggered loop's case), rebuilt otherwise;
 *   3. X is the initial guess (the previous step's solution in the loop) and
 *      is overwritten with the solution; the time is added to *time.
 * The elastic and the local (damage) solver keep one hierarchy each. */
#include <string.h>
#include <time.h>

#include "dropin.h"
#include "../../include/linalg/spd_solver.h"

typedef struct {
    lina_solver* amg;
    int n, nnz;
    int* rp;
    int* ci;
    double tol;
    int max_iter, depth, dof, precision;
    int last_iters;
    double last_relres;
} spd_state;

static spd_state spd_elastic = {NULL, 0, 0, NULL, NULL, 1e-12, 20000, 8, 2, LINA_PREC_MIXED, 0, 0};
static spd_state spd_local = {NULL, 0, 0, NULL, NULL, 1e-12, 20000, 6, 1, LINA_PREC_FP64, 0, 0};

static spd_state* pick(int which) { return which == LINA_SPD_LOCAL ? &spd_local : &spd_elastic; }

void lina_spd_solver_settings(int which, double tol, int max_iter, int depth, int dof, int precision) {
    spd_state* s = pick(which);
    s->tol = tol;
    s->max_iter = max_iter;
    s->depth = depth;
    s->dof = dof;
    s->precision = precision;
    if (s->amg) { /* new settings take effect with the next setup */
        lina_amg_destroy(s->amg);
        s->amg = NULL;
    }
}

void lina_spd_solver_free(void) {
    spd_state* all[2] = {&spd_elastic, &spd_local};
    int k;
    for (k = 0; k < 2; k++) {
        if (all[k]->amg) lina_amg_destroy(all[k]->amg);
        free(all[k]->rp);
        free(all[k]->ci);
        all[k]->amg = NULL;
        all[k]->rp = all[k]->ci = NULL;
        all[k]->n = all[k]->nnz = 0;
    }
}

int lina_spd_last_iterations(int which, double* relres) {
    spd_state* s = pick(which);
    if (relres) *relres = s->last_relres;
    return s->last_iters;
}

/* the eliminated system in CSR: constraint rows become unit rows, their
 * columns are zeroed in the other rows and moved to rhs (row order, column
 * order within a row) */
static void eliminate(const dropin_csr* c, const double* b, int* is_con, double* u, double* rhs) {
    const int n = c->n;
    int i, p;
    for (i = 0; i < n; i++) {
        int nz = 0, diag = -1;
        for (p = c->rp[i]; p < c->rp[i + 1]; p++)
            if (c->v[p] != 0.0) {
                nz++;
                if (c->ci[p] == i) diag = p;
            }
        is_con[i] = nz == 1 && diag >= 0;
        u[i] = is_con[i] ? b[i] / c->v[diag] : 0.0;
    }
    for (i = 0; i < n; i++) {
        if (is_con[i]) {
            for (p = c->rp[i]; p < c->rp[i + 1]; p++) c->v[p] = c->ci[p] == i ? 1.0 : 0.0;
            rhs[i] = u[i];
            continue;
        }
        {
            double s = 0.0;
            int any = 0;
            for (p = c->rp[i]; p < c->rp[i + 1]; p++)
                if (is_con[c->ci[p]]) {
                    s += c->v[p] * u[c->ci[p]];
                    c->v[p] = 0.0;
                    any = 1;
                }
            rhs[i] = any ? b[i] - s : b[i];
        }
    }
}

static int same_pattern(const spd_state* s, const dropin_csr* c) {
    return s->amg && s->n == c->n && s->nnz == c->rp[c->n] &&
           memcmp(s->rp, c->rp, (size_t)(c->n + 1) * sizeof(int)) == 0 &&
           memcmp(s->ci, c->ci, (size_t)s->nnz * sizeof(int)) == 0;
}

static int spd_solve(spd_state* s, sparse_matrix* A, double* X, double* b, const char* caller, double* time) {
    struct timespec t0, t1;
    dropin_csr c;
    int n, it, *is_con;
    double *u, *rhs;
    lina_stats st;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    c = dropin_to_csr(A);
    n = c.n;
    is_con = (int*)malloc((size_t)n * sizeof(int));
    u = (double*)malloc((size_t)n * sizeof(double));
    rhs = (double*)malloc((size_t)n * sizeof(double));
    eliminate(&c, b, is_con, u, rhs);
    if (same_pattern(s, &c)) {
        if (lina_amg_recompute(s->amg, c.v) < 0) {
            lina_amg_destroy(s->amg);
            s->amg = NULL;
        }
    } else if (s->amg) {
        lina_amg_destroy(s->amg);
        s->amg = NULL;
    }
    if (!s->amg) {
        lina_amg_params p;
        lina_params_vcycle(&p, s->dof, s->depth);
        p.precision = s->precision;
        s->amg = lina_amg_create(n, c.rp, c.ci, c.v, &p, 1);
        free(s->rp);
        free(s->ci);
        s->rp = s->ci = NULL;
        if (s->amg) {
            s->n = n;
            s->nnz = c.rp[n];
            s->rp = (int*)malloc((size_t)(n + 1) * sizeof(int));
            s->ci = (int*)malloc((size_t)(s->nnz > 0 ? s->nnz : 1) * sizeof(int));
            memcpy(s->rp, c.rp, (size_t)(n + 1) * sizeof(int));
            memcpy(s->ci, c.ci, (size_t)s->nnz * sizeof(int));
        }
    }
    it = -1;
    if (s->amg) {
        int i;
        for (i = 0; i < n; i++)
            if (is_con[i]) X[i] = u[i];
        it = lina_pcg(s->amg, X, rhs, s->tol, s->max_iter, LINA_STOP_RELRES, &st);
        s->last_iters = it;
        s->last_relres = st.relres;
    }
    if (it < 0) printf("GPU solver failed in %s: %s\n", caller ? caller : "?", lina_last_error());
    else if (it >= s->max_iter) printf("GPU solver in %s: no convergence after %d iterations\n", caller ? caller : "?", it);
    dropin_free_csr(&c);
    free(is_con);
    free(u);
    free(rhs);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (time) *time += (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    return it >= 0 && it < s->max_iter ? 1 : 0;
}

int lina_spd_solver(sparse_matrix* A, double* X, double* b, char* Caller, double* time) {
    return spd_solve(&spd_elastic, A, X, b, Caller, time);
}

int lina_spd_local_solver(sparse_matrix* A, double* X, double* b) {
    return spd_solve(&spd_local, A, X, b, "local solver", NULL);
}
