/*
 * lina_oracle.h -- CPU restatement of the LinA solver path (MRadszuweit/FEM,
 * LinA/) used as the parity oracle for liblina_hip.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg through oracle/bind.py; the product library
 * never links or calls it.
 *
 * Matrices are CSR (int32, sorted rows).  Every function restates one
 * reference routine, cited at its definition in lina_oracle.c.
 */
#ifndef LINA_ORACLE_H
#define LINA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_amg orc_amg;

/* AMG_setup (Multigrid.c:1613-1754) */
orc_amg* orc_amg_setup(int n, const int* rp, const int* ci, const double* v, int dof, int depth, int smiter,
                       int aggr_smiter, int coarsening, double eps_minus, double eps_plus, double matrix_tol);
void orc_amg_free(orc_amg* h);
int orc_amg_level(const orc_amg* h, int l, int* fine, int* coarse, int* nnz_a, int* nnz_p, int* nnz_r);
int orc_amg_copy(const orc_amg* h, int l, char which, int* rp, int* ci, double* v);
int orc_amg_labels(const orc_amg* h, int l, int* labels);
/* AMG_recompute_coarse_matrices (Multigrid.c:1756-1780) with new finest values */
int orc_amg_recompute(orc_amg* h, const double* v);

/* smoother on the levels below the finest (sm < 0: same as the finest) */
void orc_set_coarse_smoother(orc_amg* h, int sm, int sweeps);

/* AMG_solve (Multigrid.c:1863-1931); smoother ids as lina_hip.h */
int orc_amg_solve(orc_amg* h, const double* F, double* X, int smoother, double omega, double eps, int max_iter);
int orc_amg_solve_fmg(orc_amg* h, const double* F, double* X, int smoother, double omega);
/* cycle structure of the preconditioner inside orc_pcg_amg: 0 STAND_ALONE
 * (orc_amg_solve), 2 FULL_AMG (one orc_amg_solve_fmg pass per call) */
void orc_set_cycle(orc_amg* h, int cycle);

/* pcg (PCGS.c:163-233) preconditioned by AMG_solve; stop 0: reference test
 * r/r0 <= tol, 1: |R| <= tol.  Returns iterations; *res = pcg residuum,
 * *vcycles = V-cycles summed over all preconditioner calls. */
int orc_pcg_amg(orc_amg* h, double* X, double* B, double tol, int max_iter, int smoother, double omega,
                double amg_eps, int amg_max_iter, int zero_init, int stop, double* res, int* vcycles);

/* one smoother call of the V-cycle on a CSR matrix: `sweeps` sweeps of
 * smoother sm (ids as lina_hip.h; colour/ncolour: the colouring of the
 * multicolour sweeps, NULL otherwise; dof: the parallel Jacobi's row block;
 * post = 1: the post-smoothing call, which reverses the symmetric sweeps) */
void orc_smooth(int n, const int* rp, const int* ci, const double* v, const int* colour, int ncolour, int dof,
                const double* F, double* X, int sweeps, double w, int sm, int post);
/* planted mistakes for the tests that show the checks catch them (0 = off):
 * 1 stale gathers in the multicolour sweep, 2 forward colour order in the
 * symmetric sweep's post-smoothing call, 3 and 4 see orc_rank4_assembly and
 * orc_dirichlet_symmetric */
void orc_plant(int which);

/* sparse3D_times_rank4_2D_right_use_map (FEM2D/FEM2D.c:8534-8600) with
 * its map (:8504-8532): val (nnz of the 2n x 2n pattern rowptr/col) =
 * the contraction of the tensor rows (bptr, bi1, bi2, bval) with the nodal
 * moduli C (16 n), accumulated from zero in the reference's order.
 * Returns 0, -1 when a target entry is not in the pattern.  Planted
 * mistake 3 swaps the tensor's b and d in the moduli index. */
int orc_rank4_assembly(int n, const int* bptr, const int* bi1, const int* bi2, const double* bval, const double* C,
                       const int* rowptr, const int* col, double* val);
/* symmetric Dirichlet treatment keeping the pattern (the replacement of
 * Crack's Set_dirichlet_BC, Crack/crack.c:6445-6459, that keeps the matrix
 * symmetric): constrained rows -> unit rows with b = u; other rows subtract
 * their constrained columns' a_rc u_c (column order, summed from +0.0) from
 * b and zero them.  Planted mistake 4 leaves b unchanged in those rows. */
void orc_dirichlet_symmetric(int n, const int* rowptr, const int* col, const unsigned char* dir, const double* u,
                             double* val, double* b);

/* greedy colouring of the symmetrised nonzero pattern (natural order) */
int orc_colouring(int n, const int* rp, const int* ci, const double* v, int* colour);

/* y = A x with sequential row sums (PCG_mult, PCGS.c:36-59) */
void orc_spmv(int n, const int* rp, const int* ci, const double* v, const double* x, double* y);

/* incomplete_LU_factorization (linear_algebra.c:1958-1984): L (lower, its
 * diagonal last in each row) and U (unit diagonal first, then the strict
 * upper part) on A's pattern; lci/lv and uci/uv hold nnz(A) + n entries.
 * Returns 1, or 0 at a zero pivot. */
int orc_ilu0(int n, const int* rp, const int* ci, const double* v, int* lrp, int* lci, double* lv, int* urp,
             int* uci, double* uv);

/* FEM.c's GMRES(m) path: ILU(0), gmres (gmres.c:308-502) right-preconditioned,
 * CheckOnRestart, x = M^-1 y on return.  Returns Arnoldi steps (negative on
 * error), *eta = residual reduction at the last restart. */
int orc_gmres_ilu(int n, const int* rp, const int* ci, const double* v, double* x, const double* b, int m,
                  double tol, int max_iter, double* eta);

/* BiCGStab_solve (BiCGStab.c:115-198) preconditioned by AMG_solve on h,
 * operator (rp, ci, v).  Returns iterations; *res = |R|/|R0|. */
int orc_bicgstab_amg(orc_amg* h, int n, const int* rp, const int* ci, const double* v, double* X, const double* b,
                     int sm, double w, double amg_eps, int amg_max_iter, double eps, int max_iter, double* res);

/* pcg (PCGS.c:163-233) with PCG_default's ILU(0) preconditioner; *res =
 * get_pcg_residuum.  Returns iterations, -2 for a zero pivot. */
int orc_pcg_ilu(int n, const int* rp, const int* ci, const double* v, double* X, double* B, double tol, int max_iter,
                double* res);

/* PCG_preconlr (PCGS.c:102-108): w = U^-1 L^-1 z with orc_ilu0's factors */
void orc_ilu_apply(int n, const int* lrp, const int* lci, const double* lv, const int* urp, const int* uci,
                   const double* uv, const double* z, double* w);

#ifdef __cplusplus
}
#endif
#endif
