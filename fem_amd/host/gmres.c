/* gmres of the LinA drop-in: GMRES(m) with right and left preconditioning
 * and classical Gram-Schmidt (the method of LinA/gmres.c:308-501, which
 * FEM2D/FEM.c:737-744 drives with PCG_mult-like operators and the ILU(0)
 * preconditioner PCG_preconlr).
 *
 * The Krylov basis is one (m + 2) x n block in HBM; every vector operation
 * runs on the GPU (all Gram-Schmidt dots of a step in one pass, the
 * combinations in one pass).  The library's own operator and
 * preconditioners (PCG_mult, RN_default_mat_mult, PCG_preconlr,
 * PCG_preconl, PCG_preconr) are recognised by address and run on device
 * data; other callbacks run on host copies.  The (m + 1) x m Hessenberg
 * matrix is reduced on the host by Givens rotations as its columns arrive;
 * the iterate is formed at the end of each restart cycle (CheckOnRestart)
 * or after every step (CheckEachIter).  Option checks, defaults and the
 * result block follow LinA's (utils.c:330-420, gmres.c:357-380). */
#include <math.h>
#include <string.h>

#include "../../include/linalg/PCGS.h"
#include "../../include/linalg/itlin.h"
#include "../../include/linalg/reflective_newton.h"
#include "dropin.h"

/* an operator or preconditioner: a device matrix, a device ILU sweep, a
 * host callback, or (prec only) none = identity */
typedef struct {
    lina_matrix* mat;
    lina_ilu* ilu;
    MATVEC* host;
} gm_map;

typedef struct {
    int n;
    gm_map A, right, left;
    double *hx, *hy; /* host staging for callbacks */
} gm_ctx;

static void gm_apply(gm_ctx* c, const gm_map* f, const double* dx, double* dy, int is_prec) {
    if (f->mat) {
        dropin_check(lina_matrix_spmv_device(f->mat, dx, dy), "gmres operator");
    } else if (f->ilu) {
        dropin_check(lina_ilu_apply_device(f->ilu, dx, dy), "gmres preconditioner");
    } else if (f->host) {
        if (!c->hx) {
            c->hx = (double*)calloc((size_t)c->n, sizeof(double));
            c->hy = (double*)calloc((size_t)c->n, sizeof(double));
        }
        dropin_d2h(c->hx, dx, c->n);
        if (is_prec) dropin_d2h(c->hy, dy, c->n); /* a callback may read its output as a start value */
        f->host(c->n, c->hx, c->hy);
        dropin_h2d(dy, c->hy, c->n);
    } else {
        dropin_d2d(dy, dx, c->n);
    }
}

static gm_map gm_prec(PRECON* p) {
    gm_map m = {NULL, NULL, NULL};
    if (!p) return m;
    if (p == (PRECON*)(void*)PCG_preconlr) m.ilu = dropin_pcg_ilu('B');
    else if (p == (PRECON*)(void*)PCG_preconl) m.ilu = dropin_pcg_ilu('L');
    else if (p == (PRECON*)(void*)PCG_preconr) m.ilu = dropin_pcg_ilu('U');
    if (!m.ilu) m.host = (MATVEC*)p;
    return m;
}

/* Givens rotation (c, s) with c a - s b = r, s a + c b = 0 */
static void givens(double a, double b, double* c, double* s) {
    if (b == 0.0) {
        *c = 1.0;
        *s = 0.0;
    } else {
        const double r = hypot(a, b);
        *c = a / r;
        *s = -b / r;
    }
}

static int gm_check(int n, MATVEC* matvec, struct ITLIN_OPT* opt) {
    if (!matvec) return -99;
    if (n <= 0) return 20;
    if (opt->tol <= 0) return 21;
    if (opt->tol > 1e-1) opt->tol = 1e-1;
    if (opt->tol < 1e-15) opt->tol = 1e-15;
    if (opt->i_max <= 0 || opt->i_max > n) opt->i_max = n < 10 ? n : 10;
    return 0;
}

void gmres(int n, double* y, MATVEC* matvec, PRECON* preconr, PRECON* preconl, double* b, struct ITLIN_OPT* opt,
           struct ITLIN_INFO* info) {
    gm_ctx c;
    int m, maxit, iter = 0, i = 0, k, rc, cycles_left = 1;
    int nmat = 0, nprl = 0, nprr = 0;
    double *V, *dy, *dy0, *db, *dr, *dw;
    double *H, *cs, *sn, *g, *z, *h;
    double beta0 = 0, beta1 = 0, eta = 0, tol;
    FILE* err = opt->errorfile ? opt->errorfile : stdout;
    memset(&c, 0, sizeof(c));
    rc = gm_check(n, matvec, opt);
    if (rc != 0) {
        if (opt->errorlevel > 0) fprintf(err, "\n gmres: invalid arguments, code=%d\n", rc);
        info->rcode = rc;
        return;
    }
    m = opt->i_max;
    maxit = opt->maxiter > 0 ? opt->maxiter : 100;
    tol = opt->tol;
    c.n = n;
    c.A.mat = matvec == (MATVEC*)PCG_mult ? dropin_pcg_default_matrix()
              : matvec == (MATVEC*)RN_default_mat_mult ? dropin_rn_default_matrix() : NULL;
    if (!c.A.mat) c.A.host = matvec;
    c.right = gm_prec(preconr);
    c.left = gm_prec(preconl);
    V = lina_dev_alloc((long)(m + 2) * n);
    dy = dropin_dalloc(n);
    dy0 = dropin_dalloc(n);
    db = dropin_dalloc(n);
    dr = dropin_dalloc(n);
    dw = dropin_dalloc(n);
    if (!V || !dy || !dy0 || !db || !dr || !dw) {
        info->rcode = -999;
        goto out;
    }
    dropin_check(lina_dev_fill(V, (long)(m + 2) * n, 0.0), "fill");
    H = (double*)calloc((size_t)(m + 1) * m, sizeof(double)); /* column j at H + j*(m+1) */
    cs = (double*)calloc((size_t)m, sizeof(double));
    sn = (double*)calloc((size_t)m, sizeof(double));
    g = (double*)calloc((size_t)m + 1, sizeof(double));
    z = (double*)calloc((size_t)m + 1, sizeof(double));
    h = (double*)calloc((size_t)m + 1, sizeof(double));
    dropin_h2d(dy, y, n);
    dropin_h2d(db, b, n);
    rc = 0;
    while (cycles_left) {
        /* restart: r = b - M_l^-1 A M_r^-1 y */
        dropin_d2d(dy0, dy, n);
        gm_apply(&c, &c.right, dy, V, 1);
        gm_apply(&c, &c.A, V, dw, 0);
        gm_apply(&c, &c.left, dw, dr, 1);
        nprr++;
        nmat++;
        nprl++;
        dropin_check(lina_dev_axpby(n, -1.0, dr, 1.0, db), "axpby");
        beta1 = sqrt(dropin_dot(dr, dr, n));
        if (iter == 0) {
            beta0 = beta1;
            eta = beta0 == 0.0 ? 0.0 : 1.0;
        } else {
            eta = beta1 / beta0;
        }
        dropin_d2d(V, dr, n);
        lina_dev_scale(n, V, 1.0 / beta1);
        memset(g, 0, (size_t)(m + 1) * sizeof(double));
        g[0] = beta1;
        for (i = 1; i <= m && eta > tol; i++) {
            double* vi = V + (size_t)(i - 1) * n;
            double* vn = V + (size_t)i * n;
            double nv, t;
            double* col = H + (size_t)(i - 1) * (m + 1);
            gm_apply(&c, &c.right, vi, vn, 1);
            gm_apply(&c, &c.A, vn, dw, 0);
            gm_apply(&c, &c.left, dw, dr, 1);
            nprr++;
            nmat++;
            nprl++;
            /* classical Gram-Schmidt against the basis so far, then normalise */
            dropin_check(lina_dev_multidot(n, i, V, n, dr, h), "multidot");
            dropin_check(lina_dev_combine(n, i, V, n, h, dr, -1.0, vn), "combine");
            nv = sqrt(dropin_dot(vn, vn, n));
            lina_dev_scale(n, vn, 1.0 / nv);
            /* the new Hessenberg column, reduced by the rotations so far and a new one */
            for (k = 0; k < i; k++) col[k] = h[k];
            col[i] = nv;
            for (k = 0; k < i - 1; k++) {
                t = cs[k] * col[k] - sn[k] * col[k + 1];
                col[k + 1] = sn[k] * col[k] + cs[k] * col[k + 1];
                col[k] = t;
            }
            givens(col[i - 1], col[i], &cs[i - 1], &sn[i - 1]);
            col[i - 1] = cs[i - 1] * col[i - 1] - sn[i - 1] * col[i];
            col[i] = 0.0;
            t = cs[i - 1] * g[i - 1] - sn[i - 1] * g[i];
            g[i] = sn[i - 1] * g[i - 1] + cs[i - 1] * g[i];
            g[i - 1] = t;
            if (col[i - 1] == 0.0) {
                rc = 1;
                if (opt->errorlevel > 0) fprintf(err, "\n gmres: singular Hessenberg matrix at step %d\n", i);
                goto done;
            }
            if (opt->termcheck == CheckEachIter || i == m) {
                /* iterate y = y0 + V z with R z = g (upper triangle of H) */
                for (k = i - 1; k >= 0; k--) {
                    int j;
                    double s = g[k];
                    for (j = k + 1; j < i; j++) s -= H[(size_t)j * (m + 1) + k] * z[j];
                    z[k] = s / H[(size_t)k * (m + 1) + k];
                }
                dropin_check(lina_dev_combine(n, i, V, n, z, dy0, 1.0, dy), "combine");
                if (opt->termcheck == CheckEachIter) {
                    double* vx = V + (size_t)(i + 1) * n;
                    gm_apply(&c, &c.right, dy, vx, 1);
                    gm_apply(&c, &c.A, vx, dw, 0);
                    gm_apply(&c, &c.left, dw, dr, 1);
                    nprr++;
                    nmat++;
                    nprl++;
                    dropin_check(lina_dev_axpby(n, -1.0, dr, 1.0, db), "axpby");
                    beta1 = sqrt(dropin_dot(dr, dr, n));
                    eta = beta1 / beta0;
                }
            }
            iter++;
        }
        cycles_left = eta > tol && iter < maxit;
    }
    rc = eta <= tol ? 0 : 2;
    if (rc == 2 && opt->errorlevel > 0) fprintf(err, "\n gmres: iteration cap %d reached (n = %d)\n", maxit, n);
done:
    dropin_d2h(y, dy, n);
    info->rcode = rc;
    info->precision = eta;
    info->iter = iter;
    info->nomatvec = nmat;
    info->noprecl = nprl;
    info->noprecr = nprr;
    free(H);
    free(cs);
    free(sn);
    free(g);
    free(z);
    free(h);
out:
    if (V) lina_dev_free(V);
    dropin_dfree(dy);
    dropin_dfree(dy0);
    dropin_dfree(db);
    dropin_dfree(dr);
    dropin_dfree(dw);
    free(c.hx);
    free(c.hy);
}
