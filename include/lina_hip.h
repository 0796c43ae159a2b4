/*
 * lina_hip.h -- handle-based C ABI of liblina_hip, the MI355X (gfx950)
 * implementation of the LinA sparse-solver path (MRadszuweit/FEM, LinA/).
 *
 * The LinA drop-in entry points (pcg, AMG_setup, AMG_solve, BiCGStab_solve,
 * reflective_newton, ... with the reference's row-list sparse_matrix) are
 * declared in the headers next to this one (linear_algebra.h, PCGS.h,
 * Multigrid.h, BiCGStab.h, reflective_newton.h, gmres.h).  This header adds
 * the handle API they are built on: plain pointers and sizes, no globals, so
 * several systems can live on the device at once.
 *
 * Conventions: CSR with int32 indices, columns sorted inside each row and
 * every row holding at least one entry (the reference's sparse_zero seeds
 * the diagonal).  Functions return >= 0 on success and a negative LINA_E*
 * code on failure; lina_last_error() returns the message.  Nothing calls
 * exit(): where the reference prints and exits (e.g. NaN in pcg, PCGS.c:
 * 220-223) these functions return LINA_ENUMERIC instead.
 */
#ifndef LINA_HIP_H
#define LINA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define LINA_OK 0
#define LINA_EARG -1     /* invalid argument / shape */
#define LINA_EHIP -2     /* HIP runtime failure */
#define LINA_ESETUP -3   /* AMG setup failed (zero diagonal, no interpolation, ...) */
#define LINA_ENUMERIC -4 /* NaN / breakdown */
#define LINA_ENODEV -5   /* no device available */

/* smoothers of the V-cycle (AMG_set_smoother, LinA/Multigrid.c:1967-1969) */
#define LINA_SMOOTHER_SOR 0       /* AMG_SOR_smoother (Multigrid.c:1345-1365): lexicographic, exact */
#define LINA_SMOOTHER_JACOBI 1    /* AMG_parallel_Jacobi_smoother (Multigrid.c:1367-1400) */
#define LINA_SMOOTHER_MC_SOR 2    /* multicolour SOR, same sweep direction before and after */
#define LINA_SMOOTHER_L1_JACOBI 3 /* l1-Jacobi x += (f-Ax)/sum_j|a_ij| */
#define LINA_SMOOTHER_MC_SSOR 4   /* multicolour SOR, colours reversed on the post-smoothing call */
#define LINA_SMOOTHER_SSOR 5      /* lexicographic SOR, rows reversed on the post-smoothing call */

#define LINA_COARSEN_AGGRESSIVE 1 /* AGGRESSIVE (Multigrid.h:38) on the finest level */
#define LINA_COARSEN_STANDARD 2   /* STANDARD (Multigrid.h:37) */

#define LINA_PREC_FP64 0          /* hierarchy, smoothers and Krylov vectors in FP64 */
#define LINA_PREC_MIXED 1         /* FP32 hierarchy + smoothers; FP64 outer Krylov,
                                     dots/norms and coarsest solve */

#define LINA_CYCLE_V 0            /* STAND_ALONE (Multigrid.h:35): V-cycles until amg_eps / amg_max_iter */
#define LINA_CYCLE_FULL 2         /* FULL_AMG (Multigrid.h:38): every AMG_solve (a Krylov
                                     method's preconditioner calls included) is one
                                     full-multigrid pass (AMG_full_multigrid,
                                     Multigrid.c:1582-1611) */

#define LINA_STOP_REFERENCE 0     /* pcg's test |R_k| / |M^-1 R_0| <= tol (PCGS.c:182,213) */
#define LINA_STOP_RELRES 1        /* |R_k| <= tol with B scaled to unit norm, i.e. the
                                     relative residual |b-Ax|/|b| of the CG recurrence */

typedef struct lina_amg_params {
    /* AMG_setup arguments (Multigrid.c:1613-1624) */
    int dof;                       /* degrees of freedom per node, blocked ordering */
    int depth;                     /* number of levels */
    int smoothing_iterations;      /* sweeps per smoothing call */
    int aggr_smoothing_iterations; /* sweeps on an aggressively coarsened finest level */
    int coarsening;                /* LINA_COARSEN_* */
    double eps_minus, eps_plus;    /* set_connection_thresholds (Multigrid.c:125-128) */
    double matrix_tol;             /* AMG_set_matrix_tolerance: P truncation */
    /* AMG_solve controls */
    int smoother;                  /* LINA_SMOOTHER_* */
    double omega;                  /* AMG_set_SOR_relaxation_coeff */
    double amg_eps;                /* AMG_set_eps */
    int amg_max_iter;              /* AMG_set_max_iter */
    int zero_init;                 /* 1: the preconditioner starts every call from Z = 0 */
    int precision;                 /* LINA_PREC_* */
    int coarse_smoother;           /* smoother of the levels below the finest, -1: `smoother` */
    int coarse_sweeps;             /* sweeps per call there (when coarse_smoother >= 0) */
    int autotune;                  /* 1: time the row-partition candidates of every level's kernels at upload
                                    * and keep the fastest (changes summation order of multi-lane rows) */
    int cycle;                     /* LINA_CYCLE_*: AMG_setup's cycle_structure (AMG_solve's cycle) */
    unsigned skip_levels;          /* bit l (level l counted from the coarsest, 0 < l < depth - 1): no
                                    * smoothing on that level in the V-cycle -- its right-hand side is
                                    * restricted as it comes and the coarse correction prolongated back
                                    * (x = P V(R f)), reading no A of that level; the cycle stays symmetric
                                    * whenever the full cycle is (0) */
} lina_amg_params;

typedef struct lina_stats {
    int iterations;      /* Krylov iterations (pcg's return value) */
    int precon_calls;    /* AMG_solve calls */
    int vcycles;         /* V-cycles summed over all calls */
    double residual;     /* get_pcg_residuum(): |M^-1 R_final| / |M^-1 R_0| */
    double relres;       /* |R_final| of the recurrence with B scaled to 1 */
    double t_solve_ms;   /* device time of the solve (HIP events on the library stream) */
    double spmv_ms;      /* mean duration of the outer-loop A*p launch (kernel timing on), else 0 */
    int spmv_launches;   /* launches averaged into spmv_ms (every fourth iteration's) */
} lina_stats;

typedef struct lina_solver lina_solver;

const char* lina_last_error(void);
int lina_device_count(void);

/* parameters of PCG_set_AMG_preconditioner (PCGS.c:118-129): SOR(1.6), 2/3
 * sweeps, aggressive first coarsening, AMG solve to 1e-2, <= 100 V-cycles */
void lina_params_reference(lina_amg_params* p, int dof, int depth);
/* one symmetric V-cycle per Krylov iteration: multicolour SOR (forward
 * before, colours reversed after the coarse correction) on the finest level,
 * one l1-Jacobi sweep on the coarser levels */
void lina_params_vcycle(lina_amg_params* p, int dof, int depth);

/* Build the hierarchy for A (host CSR).  upload=0 runs the host setup only
 * (for inspection on machines without a GPU). */
lina_solver* lina_amg_create(int n, const int* rowptr, const int* col, const double* val,
                             const lina_amg_params* p, int upload);
void lina_amg_destroy(lina_solver* s);
int lina_amg_set_params(lina_solver* s, const lina_amg_params* p); /* smoother/eps/precision only */

/* hierarchy inspection: depth; per level sizes; copies of A/P/R ('A','P','R'),
 * labels, and the coarse LU factors ('L','U' on level 0) */
int lina_amg_depth(const lina_solver* s);
/* rows of the coarsest level solved through the dense inverse (AMG_Direct_
 * solver's dense part) and rows solved by their diagonal alone */
int lina_amg_coarse_rows(const lina_solver* s, int* dense, int* trivial);
int lina_amg_level_info(const lina_solver* s, int level, int* fine, int* coarse, int* nnz_a, int* nnz_p, int* nnz_r);
int lina_amg_level_nnz(const lina_solver* s, int level, char which);
int lina_amg_level_copy(const lina_solver* s, int level, char which, int* rowptr, int* col, double* val);
int lina_amg_labels(const lina_solver* s, int level, int* labels);
int lina_amg_colours(const lina_solver* s, int level, int* colours);

/* Values-only update with fixed P/R (AMG_recompute_coarse_matrices,
 * Multigrid.c:1756-1780).  The pattern of A must be unchanged. */
int lina_amg_recompute(lina_solver* s, const double* val);
/* the same with a new finest matrix of any pattern (P and R kept) */
int lina_amg_recompute_csr(lina_solver* s, int n, const int* rowptr, const int* col, const double* val);
/* lina_amg_recompute from finest values already in device memory (d_val:
 * the setup pattern's nnz values): the Galerkin products run on the device
 * from a per-pattern plan in LinA's summation order (bit-identical to the
 * host products) while every level's pattern holds, else the host path
 * takes over.  The host copy of the hierarchy (lina_amg_level_copy) is
 * brought up to date when read.  For the staggered loop's elastic step
 * (the new damage field's stiffness from lina_assembly_values). */
int lina_amg_recompute_device(lina_solver* s, const double* d_val);

/* pcg (PCGS.c:163-233) preconditioned by this hierarchy; x (initial guess,
 * overwritten) and b follow the reference's in/out semantics, including the
 * in-place scaling of b by 1/|b| and back.  Host pointers. */
int lina_pcg(lina_solver* s, double* x, double* b, double tol, int max_iter, int stop, lina_stats* st);
/* the same on device-resident vectors (length n, FP64) */
int lina_pcg_device(lina_solver* s, double* d_x, double* d_b, double tol, int max_iter, int stop, lina_stats* st);

/* AMG_solve (Multigrid.c:1863-1931): V-cycles on A x = f from the initial
 * guess x until |f-Ax|/|f-Ax0| <= amg_eps or amg_max_iter.  Host pointers;
 * returns the number of V-cycles. */
int lina_amg_solve(lina_solver* s, const double* f, double* x);
/* the same on device vectors */
int lina_amg_solve_device(lina_solver* s, const double* f, double* x);
/* lina_amg_solve that also reports |r|/|r0| after the last V-cycle, the
 * value AMG_solve records in its statistics (Multigrid.c:1918): 0 when r0
 * is 0, 1 after a full-multigrid pass */
int lina_amg_solve_stat(lina_solver* s, const double* f, double* x, double* relres);
/* the row signs make_positive_diag applied to the finest matrix
 * (linear_algebra.c:2612-2620): +1 or -1 per row; returns n */
int lina_amg_signs(const lina_solver* s, double* signs);

/* BiCGStab_solve (BiCGStab.c:115-198) preconditioned by AMG_solve with this
 * hierarchy (warm-started PP/PS as in the reference); stops at
 * |R|/|R0| <= eps or after max_iter+1 iterations.  Host pointers. */
int lina_bicgstab(lina_solver* s, double* x, const double* b, double eps, int max_iter, double* res);

/* ILU(0) systems: incomplete_LU_factorization (linear_algebra.c:1952-1982)
 * on the host, level-scheduled triangular solves on the device. */
typedef struct lina_ilu lina_ilu;
lina_ilu* lina_ilu_create(int n, const int* rowptr, const int* col, const double* val, int upload);
void lina_ilu_destroy(lina_ilu* h);
/* Row orders of an ILU(0) system.  NATURAL: the reference's factors of A
 * (level-scheduled sweeps: ~2 sqrt(n) levels on a 2-D mesh, one workgroup).
 * MULTICOLOUR (no counterpart in the reference; for callers that can trade
 * the natural order's factors for sweeps on the whole GPU): A's rows and
 * columns permuted by a greedy colouring of its pattern, explicit zeros
 * included (colour by colour,
 * natural order within a colour); the factors are exactly the reference's
 * incomplete LU of P A P^T, whose sweeps have as many levels as colours;
 * pcg / gmres iterate on P A P^T.  Vectors are always passed in A's order;
 * the factors (lina_ilu_factor_copy) are P A P^T's. */
#define LINA_ILU_NATURAL 0
#define LINA_ILU_MULTICOLOUR 1
lina_ilu* lina_ilu_create_ordered(int n, const int* rowptr, const int* col, const double* val, int ordering,
                                  int upload);
/* perm[r] = the row of A at position r of the factors (identity when natural) */
int lina_ilu_permutation(const lina_ilu* h, int* perm);
int lina_ilu_ok(const lina_ilu* h); /* 0: zero pivot, as the reference's return flag */
/* incomplete_LU_factorization (linear_algebra.c:1952-1982) of a new matrix on
 * the device, bit-identical to the host factorisation (the CSR's rows sorted
 * by column, as everywhere in this API).  While the pattern
 * (rowptr, col) stays the same as the previous refactor's, only the values
 * are uploaded and the numeric factorisation runs (the reflective-Newton
 * steps' use).  lina_ilu_ok reports a zero pivot. */
int lina_ilu_refactor(lina_ilu* h, int n, const int* rowptr, const int* col, const double* val);
int lina_ilu_factor_nnz(const lina_ilu* h, char which);
int lina_ilu_factor_copy(const lina_ilu* h, char which, int* rowptr, int* col, double* val);
/* PCG_default(A, L, U) (PCGS.c:29-34) with factors given by the caller (L:
 * diagonal last in each row, U: diagonal first); NULL on error */
lina_ilu* lina_ilu_create_factors(int n, const int* lrp, const int* lci, const double* lv, const int* urp,
                                  const int* uci, const double* uv);
/* w = U^-1 L^-1 z on device vectors */
int lina_ilu_apply_device(lina_ilu* h, const double* z, double* w);
/* w = U^-1 L^-1 z (PCG_preconlr, PCGS.c:61-108) */
int lina_ilu_apply(lina_ilu* h, const double* z, double* w);
/* pcg with PCG_default's ILU(0) preconditioner (PCGS.c:29-34, 163-233) */
int lina_ilu_pcg(lina_ilu* h, double* x, double* b, double tol, int max_iter, int stop, double* residuum);
/* restarted GMRES(m) with right ILU(0) preconditioning and the restart-only
 * convergence test FEM.c uses (FEM.c:730-745); x in/out in the unknowns'
 * space (x = M^-1 y); returns Arnoldi steps, *eta = final |r|/|r0| */
int lina_ilu_gmres(lina_ilu* h, double* x, const double* b, int m, double tol, int max_iter, double* eta);

/* device vectors owned by the library (FP64): the solve entry points above
 * take pointers from here when inputs should stay resident in HBM.
 * kind: 0 host->device, 1 device->host, 2 device->device */
double* lina_dev_alloc(long n);
int lina_dev_free(double* d);
int lina_dev_copy(double* dst, const double* src, long n, int kind);
int lina_dev_fill(double* d, long n, double value);
/* device vector primitives for host code that drives its own Krylov loop
 * (the LinA drop-in layer): each returns after completion.
 * scalar (linear_algebra.c:256-263) on device vectors, deterministic tree order */
double lina_dev_dot(const double* a, const double* b, long n);
/* lina_dev_dot that allocates nothing.  No counterpart in the reference: for
 * callers that cannot allocate device memory (allocation-free serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and result as
 * lina_dev_dot; the block partials and the result live in a file-scope
 * device scratch of 8,256 bytes per device (at most 1,024 partials), held
 * until the process ends.  n in [0, 2^31-1], else NaN and lina_last_error().
 * Every call on one device uses that device's scratch: calls that overlap
 * on one device race. */
double lina_dev_dot_noalloc(const double* a, const double* b, long n);
/* classical Gram-Schmidt pieces of GMRES (gmres.c:401-414): out[l] =
 * V_l . r for the m <= 64 vectors V_l = V + l*ld (out: host array); and
 * y = base + sign * sum_l c[l] V_l (c: host array; y may alias base) */
int lina_dev_multidot(long n, int m, const double* V, long ld, const double* r, double* out);
int lina_dev_combine(long n, int m, const double* V, long ld, const double* c, const double* base, double sign,
                     double* y);
/* x = a*x + b*y  (vector_add with a = 1, linear_algebra.c:311-316) */
int lina_dev_axpby(long n, double a, double* x, double b, const double* y);
/* x = a*x  (scalar_mult, linear_algebra.c:299-304) */
int lina_dev_scale(long n, double* x, double a);

/* single kernels on the finest level for verification and timing:
 * y = A x (precision FP64 or FP32 copy of A), host pointers */
int lina_spmv(lina_solver* s, int level, int precision, const double* x, double* y);
/* average device time (ms) of one finest-level SpMV launch over reps
 * launches, HIP events recorded on the library's stream */
double lina_time_spmv(lina_solver* s, int level, int precision, int reps);
/* per-launch algorithmic byte count used for roofline reporting, from the
 * uploaded arrays: (value size + index size) nnz + 4 (n+1) row pointers +
 * 8 (blocks+1) block bounds (+ 4 per block of base columns) + x and y once;
 * the index size is 2 where every row block's columns span < 65536, else 4 */
double lina_spmv_bytes(const lina_solver* s, int level, int precision);
/* on != 0: pcg brackets the outer-loop A*p launch of every fourth iteration
 * with HIP events and reports the mean launch duration in lina_stats.spmv_ms */
int lina_amg_kernel_timing(lina_solver* s, int on);
/* the autotuner's choices (params.autotune): per tuned matrix its level, kind
 * (A, P, R, C = colour-permuted A) and lanes per row for the FP64 / FP32
 * copies; returns the number of entries (arrays of size cap) */
int lina_amg_tuned(lina_solver* s, int* level, char* which, int* lpr64, int* lpr32, int cap);

/* A device-resident CSR matrix outside any hierarchy: the library's own
 * y = A x of sparse_mult (linear_algebra.c:134-147) and PCG_mult
 * (PCGS.c:36-59), summed per row in column order like the reference.
 * Rows need >= 1 entry with sorted unique columns.  NULL on error. */
typedef struct lina_matrix lina_matrix;
lina_matrix* lina_matrix_create(int n, const int* rowptr, const int* col, const double* val);
void lina_matrix_destroy(lina_matrix* h);
/* host pointers; precision LINA_PREC_MIXED uses the FP32 copy */
int lina_matrix_spmv(lina_matrix* h, int precision, const double* x, double* y);
/* device pointers (FP64); returns after the product is complete */
int lina_matrix_spmv_device(lina_matrix* h, const double* x, double* y);
double lina_matrix_time_spmv(lina_matrix* h, int precision, int reps);
/* power_method (LinA/linear_algebra.c:4802-4822), the eigenvalue estimate of
 * the reflective Newton's negative-curvature search (get_neg_curvature_direction,
 * :4849-4865): power iterations v <- A v / |A v| from the start vector v
 * (host, length n; ends as the last normalised iterate), the estimate
 * v_new.v_old / |v_old|^2 with |v_old| = 1, until |change| / matrix_norm(A) <= eps.
 * *eig: on entry the estimate the first test compares with (the reference
 * reads its caller's variable), on exit the last estimate.  Returns the
 * iteration count, 0 when max_iter is reached, negative on error. */
int lina_matrix_power(lina_matrix* h, double* v, int max_iter, double eps, double* eig);
/* the algorithmic bytes of one product, counted as for lina_spmv_bytes */
double lina_matrix_spmv_bytes(const lina_matrix* h, int precision);
/* pcg (PCGS.c:163-233) with the operator op (PCG_set_mat_mult with a device
 * matrix, PCGS.c:137-139) and this hierarchy as the preconditioner: solves
 * op x = b; op has the hierarchy's size and at most as many row blocks as its
 * finest matrix */
int lina_pcg_operator(lina_solver* s, lina_matrix* op, double* x, double* b, double tol, int max_iter, int stop,
                      lina_stats* st);

/* Backward entry point of P1 plane-strain elasticity with per-element Young's
 * moduli theta (K = sum_e elem_scale_e theta_e K_e(E=1)): the adjoint solve
 * K lambda = dL/du with this hierarchy (K symmetric), then
 *   grad_e = dL/dtheta_e = -elem_scale_e lambda_e^T K_e(E=1) u_e
 * per element on the device.  No counterpart in the reference.  xy: node
 * coordinates (x,y pairs), tri: 3 node indices per element, blocked != 0:
 * dofs [x of all nodes; y of all nodes] (else interleaved), u: forward
 * solution with the prescribed values on constrained dofs, dLdu: zero on
 * constrained dofs; elem_scale may be NULL (1).  Host pointers; returns the
 * adjoint pcg's iterations (relative-residual stop at tol). */
int lina_elasticity_adjoint(lina_solver* s, int n_nodes, const double* xy, int n_elem, const int* tri, double nu,
                            const double* elem_scale, int blocked, const double* u, const double* dLdu, double tol,
                            int max_iter, double* lambda, double* grad, lina_stats* st);

/* ------------------------------------------------------------------
 * Batched AMG-PCG (no counterpart in the reference, whose API holds one
 * system in globals): nsys independent SPD systems given as one
 * block-diagonal CSR whose system k owns rows [row_off[k], row_off[k+1]).
 * Each system gets its own AMG_setup hierarchy (params p, equal depth); the
 * hierarchies are joined into one block-diagonal hierarchy so that every
 * kernel of the V-cycle serves all systems.  pcg per system with the
 * relative-residual stop |b_k - A_k x_k| <= tol |b_k|; graph_iters > 0
 * captures that many iterations into a hipGraph and replays it.
 * p->autotune = 1 picks each level's row partition by timing.  Every CG
 * iteration is preconditioned by one V-cycle from a zero guess: p->cycle
 * must be LINA_CYCLE_V, p->amg_max_iter 1 and p->zero_init 1 (the vcycle
 * defaults), else lina_batch_create returns NULL (LINA_EARG message).
 * ------------------------------------------------------------------ */
typedef struct lina_batch_stats {
    int iterations_launched; /* iterations run (graph replays x graph_iters) */
    int active;              /* systems not converged at the end */
    double t_solve_ms;       /* device time of the solve */
} lina_batch_stats;
typedef struct lina_batch lina_batch;
lina_batch* lina_batch_create(int nsys, const int* row_off, const int* rowptr, const int* col, const double* val,
                              const lina_amg_params* p);
void lina_batch_destroy(lina_batch* h);
/* host arrays: x (guess in, solution out) and b of the union; iters[k] and
 * relres[k] per system (may be NULL); returns the iterations launched */
int lina_batch_pcg(lina_batch* h, double* x, const double* b, double tol, int max_iter, int graph_iters,
                   double* iters, double* relres, lina_batch_stats* st);
/* the same on device vectors (lina_dev_alloc) */
int lina_batch_pcg_device(lina_batch* h, double* x, const double* b, double tol, int max_iter, int graph_iters,
                          double* iters, double* relres, lina_batch_stats* st);
/* the autotuner's choices: per tuned matrix its level, kind (A, P, R, C =
 * colour-permuted A) and lanes per row for the FP64 / FP32 copies; returns
 * the number of entries (arrays of size cap) */
int lina_batch_tuned(lina_batch* h, int* level, char* which, int* lpr64, int* lpr32, int cap);

/* ------------------------------------------------------------------
 * Reflective Newton (reflective_newton.c:660-711): minimise
 *   q(x) = x.H.x/2 + x.b   subject to lower_k <= x_idx[k] <= upper_k
 * (RN_default_func / RN_default_gradient, NaN bound = unbounded side) by the
 * trust-region reflective Newton method with a two-dimensional subspace
 * minimum (subspace_minimum) and the piecewise-linear line search of
 * RN_iterate (:536-627).  Every length-n vector operation, SpMV and inner
 * Krylov solve runs on the device; the scaled Hessian M = D H D + diag(|g| J)
 * (RN_get_M_D, with sparse_product's zero deletion) is assembled on the host,
 * as are its ILU(0) factors or Galerkin coarse matrices.
 * ------------------------------------------------------------------ */
#define LINA_RN_BICGSTAB 1 /* BiCGStab + AMG (RN_init_BiCGStab), eps bi_eps, <= bi_max_iter */
#define LINA_RN_AMG 2      /* AMG_solve with the recomputed hierarchy */
#define LINA_RN_PCG 3      /* pcg with ILU(0) of M, tol = eps, <= pcg_iter */
#define LINA_RN_GMRES 4    /* local solver in the RN_UMF slot (RN_set_local_solver): GMRES(gmres_m) with right
                            * ILU(0) as FEM.c drives it; M indefinite when it misses gmres_tol */
#define LINA_RN_HOST 5     /* the RN_UMF slot with the caller's own solver (host_solver): M as host CSR, x (the
                            * start -D g, overwritten with the solution) and b = -D g on the host; its return
                            * value is the status (1 positive definite, 0 indefinite) */

/* a host local solver for LINA_RN_HOST (rowptr/col/val: M in CSR, sorted rows) */
typedef int (*lina_rn_host_solver)(void* ctx, int n, const int* rowptr, const int* col, const double* val, double* x,
                                   const double* b);

typedef struct lina_rn_params {
    int max_iter;        /* RN_max_iter (50) */
    double eps;          /* RN_eps (1e-8): outer |dq/q| and the inner PCG tolerance */
    double sigma_l, sigma_u, rho;  /* line search (0.1, 0.9, 0.1) */
    double delta;        /* initial trust radius RN_Delta (250) */
    double xi;           /* boundary shift RN_Xi (1) */
    double trust_lower, trust_upper;   /* contraction thresholds (0.6, 1.1) */
    double radius_min, radius_max;     /* (1e-3, 1000) */
    int solver;          /* LINA_RN_* (reference default: BiCGStab) */
    int pcg_iter;        /* RN_PCG_iter (100) */
    double bi_eps;       /* BI_eps of RN_init_BiCGStab (1e-3) */
    int bi_max_iter;     /* BI_max_iter (10) */
    int gmres_m;         /* restart length (30) */
    int gmres_max_iter;  /* Arnoldi steps (600) */
    double gmres_tol;    /* |r|/|r0| of the preconditioned system (1e-10) */
    lina_amg_params amg; /* hierarchy of H for the BiCGStab / AMG inner solvers */
    unsigned long long seed; /* start vectors of the negative-curvature power method */
    double q_offset;     /* constant added to q (0): a functional x.H.x/2 + b.x + c keeps its values, which
                          * the relative reduction |dq/q| of the outer test depends on */
    lina_rn_host_solver host_solver; /* LINA_RN_HOST */
    void* host_ctx;
    int ilu_ordering;    /* LINA_ILU_* of the PCG / GMRES inner solvers' ILU(0) (natural); MULTICOLOUR
                          * factors M on H's pattern plus the diagonal (M's absent entries as zeros), its
                          * values assembled on the device every step */
} lina_rn_params;

typedef struct lina_rn_stats {
    int iterations;      /* outer Newton steps (the return value) */
    int line_searches;   /* steps that needed the bisection line search */
    int line_search_iter;
    int neg_curvature;   /* inner solves that reported an indefinite M */
    int inner_iterations;/* summed inner solver iterations */
    double dq;           /* last |(q - q0)/q0| */
    double q;            /* final functional value */
    double delta;        /* trust radius after the solve */
} lina_rn_stats;

typedef struct lina_rn lina_rn;
void lina_rn_params_default(lina_rn_params* p, int dof, int depth);
/* H (n x n CSR), b (n), nb bounded indices idx with lower/upper (NaN =
 * unbounded); the arrays are copied.  NULL on error. */
lina_rn* lina_rn_create(int n, const int* rowptr, const int* col, const double* val, const double* b, int nb,
                        const int* idx, const double* lower, const double* upper, const lina_rn_params* p);
void lina_rn_destroy(lina_rn* h);
/* reflective_newton(X): X (host, n) in/out; the trust radius persists
 * between calls like the reference's RN_Delta */
int lina_rn_solve(lina_rn* h, double* x, lina_rn_stats* st);
/* RN_set_trust_region */
int lina_rn_set_trust_region(lina_rn* h, double delta);
/* the scaled Hessian M = D H D + diag(|G| J) of RN_get_M_D
 * (reflective_newton.c:250-262) with sparse_product's zero deletion, on the
 * host; mci/mv need room for rowptr[n] + n entries; returns nnz(M) */
int lina_rn_scaled_hessian(int n, const int* rowptr, const int* col, const double* val, const double* D,
                           const double* G, const double* J, int* mrp, int* mci, double* mv);
/* ---------------------------------------------------------- re-assembly
 *
 * The degraded stiffness of Crack's elastic step on the device: the
 * contraction of the third-order tensor Strain_grad with the nodal rank-4
 * moduli, sparse3D_times_rank4_2D_right_use_map (FEM2D/FEM2D.c:8534-8600,
 * map: :8504-8532).  n nodes; the tensor as rows k = 0..n-1 of triples
 * q in [bptr[k], bptr[k+1]) with indices bi1[q] = n b + j, bi2[q] = n d + k'
 * (b, d in {0, 1}) and values bval[q]; the 2n x 2n target pattern rowptr /
 * col (sorted columns) must hold every entry (n a + j, n c + k').  Each
 * entry's value is the sum, from +0.0 and in the reference's order (node,
 * triple, then (a,c) = (0,0), (0,1), (1,1), (1,0)), of
 * C[n (8a + 4b + 2c + d) + k] * bval[q]: bit-identical to the reference.
 * Host arrays at creation (copied); NULL on error. */
typedef struct lina_assembly lina_assembly;
lina_assembly* lina_assembly_create(int n, const int* bptr, const int* bi1, const int* bi2, const double* bval,
                                    const int* rowptr, const int* col);
void lina_assembly_destroy(lina_assembly* h);
int lina_assembly_nnz(const lina_assembly* h);
/* the number of products one assembly sums (4 per tensor triple) */
long lina_assembly_terms(const lina_assembly* h);
/* constrained DOFs (host, 2n flags; NULL: none) for lina_assembly_values */
int lina_assembly_set_dirichlet(lina_assembly* h, const unsigned char* is_dirichlet);
/* val (device, nnz) = the contraction with the moduli C (device, 16 n);
 * then, with constrained DOFs set, the symmetric Dirichlet treatment of
 * lina_dirichlet_symmetric with the values u_dir (device, 2n) on val and on
 * b (device, 2n, may be NULL).  Returns after completion. */
int lina_assembly_values(lina_assembly* h, const double* C, double* val, const double* u_dir, double* b);
/* Symmetric Dirichlet treatment of a CSR system keeping its pattern (device
 * arrays): a constrained row i becomes the unit row with b[i] = u[i]; every
 * other row r gets b[r] -= sum over its constrained columns c, in column
 * order from +0.0, of a_rc u[c], and those a_rc become 0.  The matrix stays
 * symmetric where it was: the SPD-preserving replacement of Crack's
 * Set_dirichlet_BC (Crack/crack.c:6445-6459), which only resets the rows.
 * b may be NULL. */
int lina_dirichlet_symmetric(int n, const int* rowptr, const int* col, const unsigned char* is_dirichlet,
                             const double* u, double* val, double* b);

/* The angle search of subspace_minimum (reflective_newton.c:394-450): the
 * stationary points of q(t) = c (c ax + s axy + bx) + s (s ay + by),
 * c = delta cos t, s = delta sin t, found on 1000 intervals of [0, 2 pi]
 * (GMRES_Newton.c:402-439); *theta receives the one of least q.  Returns
 * the number of stationary points, negative (lina_last_error) when there
 * is none. */
int lina_circle_minimum(double ax, double ay, double axy, double bx, double by, double delta, double* theta);

#ifdef __cplusplus
}
#endif
#endif
