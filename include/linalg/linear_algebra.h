/* linear_algebra.h of the liblinalg_hip drop-in: the LinA row-list matrix
 * and the vector/matrix utilities the solver path and its callers use, with
 * the types and signatures of LinA/linear_algebra.h (cited per function) so
 * that code written against LinA compiles and links unchanged.  Host C
 * (fem_amd/host/linalg.c); the solvers (PCGS.h, Multigrid.h, BiCGStab.h,
 * reflective_newton.h) run on the GPU through liblina_hip. */
#ifndef LINALG_HIP_LINEAR_ALGEBRA_H
#define LINALG_HIP_LINEAR_ALGEBRA_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

/* row-list sparse matrix (linear_algebra.h:29-34): row i holds Len[i]
 * entries with sorted column indices Indices[i][] and values Values[i][] */
typedef struct SPARSE {
    double** Values;
    int** Indices;
    int* Len;
    int size;
} sparse_matrix;

sparse_matrix* sparse_zero(int length);                          /* linear_algebra.c:501-508: (i, 0.0) per row */
int get_position(int* list, int size, int val, int* Flag);       /* :1515-1537 */
int insert_sparse(sparse_matrix* A, double a, int i, int j);     /* :1124-1133: adds to an existing entry */
void set_sparse(sparse_matrix* A, double a, int i, int j);       /* :1135-1143 */
double get_matrix_element(sparse_matrix* A, int i, int j);       /* :1236-1243 */
double get_diag(sparse_matrix* A, int i);                        /* :3434-3439 */
void delete_zeros(sparse_matrix* A);                             /* :793-805 */
sparse_matrix* clone(sparse_matrix* A);                          /* :1453-1471 */
void Free_sparse(sparse_matrix* A);                              /* :807-816 */
#define free_sparse(A) Free_sparse(A);free(A)                    /* linear_algebra.h:375 */

double* sparse_mult(sparse_matrix* A, double* x);                /* :134-147: y = A x (new vector) */
void linear_map(double* x, double fac, sparse_matrix* A, double* y); /* :149-172: x += fac A y */
void sparse_multiplication(sparse_matrix* A, double* x);         /* :174-184: x = A x */
double sparse_bilinear(double* y, sparse_matrix* A, double* x);  /* :442-447: y . A x */
double matrix_residuum(sparse_matrix* A, double* Sol, double* b); /* :2431-2438: |b - A Sol| */

double scalar(double* x1, double* x2, int n);                    /* :256-263 */
double euklid_norm(double* x1, int n);                           /* :274-277 */
void scalar_mult(double a, double* x, int n);                    /* :299-304 */
void vector_add(double* x, double* y, double factor, int n);     /* :311-316 */
double* zero_vector(int n);                                      /* :3158-3166 */
double* clone_vector(double* x, int n);                          /* :56-65 */
void copy_vector_to(double* x, double* y, int n);                /* :67-77 */

/* ILU(0) (linear_algebra.c:1952-1982): L and U created by sparse_zero(n);
 * factors computed bit-identically to the reference; returns 0 on a zero
 * pivot */
int incomplete_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U);
void L_triang_invert(sparse_matrix* L, double* x);               /* :2008-2025 (on the GPU) */
void U_triang_invert(sparse_matrix* U, double* x);               /* :2027-2043 (on the GPU) */

int* zero_int_list(int n);                                        /* :3168-3174 */
void copy_vector_content(double* x, double* y, int x_pos, int y_pos, int size); /* :79-82 */
sparse_matrix* get_transpose(sparse_matrix* A, int new_size);    /* :1290-1300 */
void enlarge_matrix(sparse_matrix* A, int mesh_size, int new_size, int equ_index, int var_index); /* :4109-4137 */
void simple_enlarge_matrix(sparse_matrix* A, int add_size);      /* :4139-4156 */

/* primal-dual active-set method for min/max of the quadratic of A under
 * L <= B x <= U (linear_algebra.c:4933-5042): each pass solves the system of
 * A augmented with the active constraints through the caller's Solver
 * (add_constraints_default :4867-4884, or add_constraints_flux :4886-4918
 * for mode ASS_FLUX), drops constraints whose multiplier has the wrong sign
 * (definiteness +1: minimum, -1: maximum) and activates violated ones (tol
 * 1e-10), until none change or max_iter passes; state in *solver_info
 * (created on first use). */
#define ASS_FLUX 1
#define ASS_DEFAULT 0
typedef struct ASS_IFO {                                         /* linear_algebra.h:81-88, same layout */
    int active_size;
    int* Active;
    double* Signs;
    double* Constraint;
    int* Iter_history;
    int calls;
} ass_info;
sparse_matrix* add_constraints_default(sparse_matrix* A, sparse_matrix* Constraints, int* ActiveSet,
                                       int active_size);
sparse_matrix* add_constraints_flux(sparse_matrix* A, sparse_matrix* Constraints, int* ActiveSet, int active_size);
int active_set_solver(sparse_matrix* A, double* Sol, double* b, sparse_matrix* B, double* L, double* U,
                      int (*Solver)(sparse_matrix* A, double* X, double* b), int max_iter, double definiteness,
                      int mode, ass_info** solver_info);
void free_ass_info(ass_info** info);                             /* :4920-4928 */

/* products, sweeps and direct factorisations on caller-owned host row lists */
sparse_matrix* sparse_product(sparse_matrix* A, sparse_matrix* B);        /* :186-222: A B, zeros deleted */
void SOR(sparse_matrix* A, double* F, double* Sol, double alpha, int iter); /* :2464-2512 (iter <= 0: to 1e-10) */
int gauss_seidel(int (*Inv_NB_matvec)(int n, double* x, double* b), void (*NB_matvec)(int n, double* y, double* z),
                 double* Solution, double* b, int size, int iterations, double tol); /* :2514-2544 */
double residuum(void (*Matvec)(int n, double* y, double* z), double* Sol, double* b, int n); /* :1999-2006 */
void sparse_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U);   /* :3040-3062 */
void complete_LU_factorization(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U); /* :3122-3139 */
void convert_array_to_sparse(double** M_A, sparse_matrix* A, int col_size, int row_size); /* :3082-3090 */
void diagonal_preconditioner(sparse_matrix* A, sparse_matrix* L, sparse_matrix* U); /* :1984-1997 */
long double hp_scalar(double* x1, double* x2, int n);                     /* :265-272 */
long double hp_euklid_norm(double* x1, int n);                            /* :279-282 */
#define AUTOMATIC -1                                                      /* linear_algebra.h:14 */
void set_thread_num(int n);                                               /* :23-32: host threads of the setup */
void set_omp_chunk_size(int chunk);                                       /* :34-36 */

/* eigenvalue searches (power iterations on the GPU) */
double matrix_norm(sparse_matrix* A);                                     /* linear_algebra.c:3441-3450 */
void set_random_normal_vector(double* X, int n);                          /* :4824-4829 */
int power_method(sparse_matrix* A, double* max_eig, double* Max_vec, int max_iter, double eps,
                 sparse_matrix* Original, double eps_neg);                /* :4802-4822 */
int get_most_negative_eigenvector(sparse_matrix* A, double* eigenvalue, double* Eigenvector, int max_iter,
                                  double eps);                            /* :4831-4847 */
int get_neg_curvature_direction(sparse_matrix* A, double* curvature, double* Direction, int max_iter, double eps,
                                double neg_eps);                          /* :4849-4865 */

/* sparse-matrix files: binary rows (int n, Len[n], indices, values), "(j|v)"
 * text lines, MatrixMarket coordinate output (<Name>.mtx) */
int write_sparse(sparse_matrix* A, char* path, char* name);               /* linear_algebra.c:832-844 */
sparse_matrix* read_sparse(char* path, char* name);                       /* :846-868 */
sparse_matrix* read_sparse_textfile(char* path, char* name);              /* :870-916 */
void write_sparse_textfile(char* path, char* name, sparse_matrix* A);     /* :983-1002 */
void print_to_matrix_market_format(sparse_matrix* A, char* Name);        /* :1023-1056 */

#ifdef __cplusplus
}
#endif
#endif
