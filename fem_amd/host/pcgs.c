/* pcg and its global configuration (LinA/PCGS.c) in the LinA drop-in.
 *
 * The iteration is the reference's, operation for operation (PCGS.c:
 * 163-233); every vector lives in HBM and every vector operation, product
 * and preconditioner application runs on the GPU through liblina_hip's C
 * ABI.  Only the scalars steering the loop come back to the host.  The
 * library's own operator and preconditioners (PCG_mult, PCG_preconlr,
 * PCG_AMG_precon) are recognised by address and run natively; callbacks the
 * caller registers run on host copies of the vectors. */
#include "../../include/linalg/PCGS.h"

#include "../../include/linalg/Multigrid.h"
#include "dropin.h"

static sparse_matrix *default_A = NULL, *default_L = NULL, *default_U = NULL;
static lina_matrix* dev_A = NULL;
static lina_ilu* dev_LU = NULL;
static lina_ilu *dev_Lonly = NULL, *dev_Uonly = NULL; /* one sweep each: the other factor the identity */
static int (*precon)(int n, double* R, double* Z) = NULL;
static void (*mat_multi)(int n, double* x, double* y) = NULL;
static amg_system_info* PCG_AMG_data = NULL;
static double pcg_residuum = 0;

lina_matrix* dropin_pcg_default_matrix(void) { return dev_A; }

void PCG_default(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U) {
    default_A = A;
    default_L = L;
    default_U = U;
    if (dev_A) lina_matrix_destroy(dev_A);
    dev_A = A ? dropin_device_matrix(A) : NULL;
    if (dev_LU) lina_ilu_destroy(dev_LU);
    if (dev_Lonly) lina_ilu_destroy(dev_Lonly);
    if (dev_Uonly) lina_ilu_destroy(dev_Uonly);
    dev_LU = dev_Lonly = dev_Uonly = NULL;
    if (L && U) {
        dropin_csr l = dropin_to_csr(L), u = dropin_to_csr(U);
        dev_LU = lina_ilu_create_factors(l.n, l.rp, l.ci, l.v, u.rp, u.ci, u.v);
        if (!dev_LU) dropin_check(-1, "PCG_default factors");
        dropin_free_csr(&l);
        dropin_free_csr(&u);
    }
    PCG_set_default_precon();
}

void PCG_mult(int n, double* y, double* z) {
    double *dy, *dz;
    if (!dev_A) {
        fprintf(stderr, "liblinalg_hip: PCG_mult without PCG_default\n");
        return;
    }
    dy = dropin_dalloc(n);
    dz = dropin_dalloc(n);
    dropin_h2d(dy, y, n);
    dropin_check(lina_matrix_spmv_device(dev_A, dy, dz), "PCG_mult");
    dropin_d2h(z, dz, n);
    dropin_dfree(dy);
    dropin_dfree(dz);
}

int PCG_preconlr(int n, double* z, double* w) {
    (void)n;
    if (!dev_LU) {
        fprintf(stderr, "liblinalg_hip: PCG_preconlr without PCG_default factors\n");
        return 0;
    }
    dropin_check(lina_ilu_apply(dev_LU, z, w), "PCG_preconlr");
    return 1;
}

/* the sweeps of PCG_default's factors on the device: 'B' both (PCG_preconlr),
 * 'L' the forward one, 'U' the backward one (created on first use, with the
 * other factor the identity: a unit row costs one exact division by 1) */
lina_ilu* dropin_pcg_ilu(char which) {
    lina_ilu** h;
    dropin_csr f, id;
    int i, n;
    if (which == 'B') return dev_LU;
    h = which == 'L' ? &dev_Lonly : &dev_Uonly;
    if (*h || !default_L || !default_U) return *h;
    f = dropin_to_csr(which == 'L' ? default_L : default_U);
    n = f.n;
    id.n = n;
    id.rp = (int*)malloc((size_t)(n + 1) * sizeof(int));
    id.ci = (int*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    id.v = (double*)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (i = 0; i <= n; i++) id.rp[i] = i;
    for (i = 0; i < n; i++) {
        id.ci[i] = i;
        id.v[i] = 1.0;
    }
    *h = which == 'L' ? lina_ilu_create_factors(n, f.rp, f.ci, f.v, id.rp, id.ci, id.v)
                      : lina_ilu_create_factors(n, id.rp, id.ci, id.v, f.rp, f.ci, f.v);
    if (!*h) dropin_check(-1, which == 'L' ? "PCG_preconl" : "PCG_preconr");
    dropin_free_csr(&f);
    dropin_free_csr(&id);
    return *h;
}

void PCG_preconl(int n, double* z, double* w) {
    lina_ilu* h = dropin_pcg_ilu('L');
    (void)n;
    if (!h) {
        fprintf(stderr, "liblinalg_hip: PCG_preconl without PCG_default factors\n");
        return;
    }
    dropin_check(lina_ilu_apply(h, z, w), "PCG_preconl");
}

void PCG_preconr(int n, double* z, double* w) {
    lina_ilu* h = dropin_pcg_ilu('U');
    (void)n;
    if (!h) {
        fprintf(stderr, "liblinalg_hip: PCG_preconr without PCG_default factors\n");
        return;
    }
    dropin_check(lina_ilu_apply(h, z, w), "PCG_preconr");
}

int PCG_AMG_precon(int n, double* Y, double* X) {
    (void)n;
    return AMG_solve(PCG_AMG_data, Y, X, NULL);
}

void PCG_set_precon(int (*Precon)(int n, double* R, double* Z)) { precon = Precon; }
void PCG_set_mat_mult(void (*Mat_mult)(int n, double* x, double* y)) { mat_multi = Mat_mult; }
void* PCG_get_precon(void) { return (void*)precon; }
void* PCG_get_mat_mult(void) { return (void*)mat_multi; }
void PCG_set_default_precon(void) {
    PCG_set_mat_mult(&PCG_mult);
    PCG_set_precon(&PCG_preconlr);
}

void PCG_set_AMG_preconditioner(sparse_matrix* A, int num_fields, int depth) {
    AMG_print_info(NO);
    AMG_set_SOR_relaxation_coeff(1.6);
    AMG_set_eps(1e-2);
    AMG_set_max_iter(100);
    PCG_AMG_data = AMG_setup(A, num_fields, depth, 2, 3, AGGRESSIVE, STAND_ALONE, DIRECT, NULL, NULL);
    PCG_set_precon(&PCG_AMG_precon);
}

double get_pcg_residuum(void) { return pcg_residuum; }

int pcg(double* X, double* B, int n, double tol, int max_iter) {
    dropin_op op = {NULL, NULL};
    dropin_prec pr = {NULL, NULL, NULL};
    dropin_stage sg;
    double *dX, *dB, *dR, *dP, *dQ, *dZ, *dW;
    double h, r0, r, s = 0, s_old = 0, a, b;
    int i = 0;
    if (mat_multi == PCG_mult) op.dev = dev_A;
    else op.host = mat_multi;
    if (precon == PCG_AMG_precon && PCG_AMG_data) {
        dropin_amg_sync(PCG_AMG_data);
        pr.amg = (lina_solver*)PCG_AMG_data->device;
    } else if (precon == PCG_preconlr) {
        pr.ilu = dev_LU;
    } else {
        pr.host = precon;
    }
    if ((!op.dev && !op.host) || (!pr.amg && !pr.ilu && !pr.host)) {
        fprintf(stderr, "liblinalg_hip: pcg without operator or preconditioner\n");
        return -1;
    }
    dropin_stage_init(&sg, n);
    dX = dropin_dalloc(n);
    dB = dropin_dalloc(n);
    dR = dropin_dalloc(n);
    dP = dropin_dalloc(n);
    dQ = dropin_dalloc(n);
    dZ = dropin_dalloc(n);
    dW = dropin_dalloc(n);
    dropin_h2d(dX, X, n);
    dropin_h2d(dB, B, n);
    h = sqrt(dropin_dot(dB, dB, n));
    if (h == 0) h = 1;
    lina_dev_scale(n, dB, (double)1 / h);
    lina_dev_scale(n, dX, (double)1 / h);
    dropin_d2d(dR, dB, n);
    dropin_op_apply(&op, &sg, dX, dQ);
    lina_dev_axpby(n, 1.0, dR, -1., dQ);
    dropin_prec_apply(&pr, &sg, dR, dW);
    r0 = sqrt(dropin_dot(dW, dW, n));
    if (r0 == 0) {
        /* the reference returns here with B and X still scaled */
        dropin_d2h(X, dX, n);
        dropin_d2h(B, dB, n);
        goto done;
    }
    r = r0;
    do {
        dropin_prec_apply(&pr, &sg, dR, dZ);
        s_old = s;
        s = dropin_dot(dR, dZ, n);
        if (i > 0) {
            b = s / s_old;
            lina_dev_scale(n, dP, b);
            lina_dev_axpby(n, 1.0, dP, 1, dZ);
        } else {
            dropin_d2d(dP, dZ, n);
        }
        dropin_op_apply(&op, &sg, dP, dQ);
        a = s / dropin_dot(dP, dQ, n);
        lina_dev_axpby(n, 1.0, dX, a, dP);
        lina_dev_axpby(n, 1.0, dR, -a, dQ);
        r = sqrt(dropin_dot(dR, dR, n));
        i++;
    } while (r / r0 > tol && i < max_iter);
    if (i == max_iter) printf("pcg: maximum of %d iterations reached\n", i);
    dropin_prec_apply(&pr, &sg, dR, dW);
    r = sqrt(dropin_dot(dW, dW, n));
    if (isnan(r / r0)) fprintf(stderr, "liblinalg_hip: pcg produced NaN\n");
    lina_dev_scale(n, dB, h);
    lina_dev_scale(n, dX, h);
    dropin_d2h(X, dX, n);
    dropin_d2h(B, dB, n);
    pcg_residuum = r / r0;
done:
    dropin_dfree(dX);
    dropin_dfree(dB);
    dropin_dfree(dR);
    dropin_dfree(dP);
    dropin_dfree(dQ);
    dropin_dfree(dZ);
    dropin_dfree(dW);
    dropin_stage_free(&sg);
    return i;
}
