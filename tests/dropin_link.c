/* One translation unit that takes the address of every LinA function the
 * callers in Crack/crack.c and FEM2D/{FEM,FEM2D,K2_tree,conditionfile,
 * geometry2D}.c use and of every function declared in LinA's solver headers
 * (PCGS.h, Multigrid.h, BiCGStab.h, reflective_newton.h), plus gmres
 * (gmres.c:308-310), sparse_LU_factorization and Free_sparse.  Built once against the drop-in's headers and
 * liblinalg_hip.so and once against LinA's own headers and the library built
 * from LinA's sources: both must compile and link with no undefined symbol.
 * Prints the number of entries. */
#include <stdio.h>

#include "BiCGStab.h"
#include "File_stuff.h"
#include "Multigrid.h"
#include "PCGS.h"
#include "linear_algebra.h"
#include "red_black_tree.h"
#include "reflective_newton.h"
#ifdef LINA_REFERENCE_HEADERS
void gmres(); /* LinA declares it in no installed header; FEM.c calls it undeclared */
#else
#include "itlin.h"
#endif

typedef void (*entry)(void);

static const entry table[] = {
    (entry)&AMG_SOR_smoother, (entry)&AMG_add_stat, (entry)&AMG_default_iterations,
    (entry)&AMG_free_data, (entry)&AMG_get_max_iter, (entry)&AMG_info_set,
    (entry)&AMG_init_stat, (entry)&AMG_parallel_Jacobi_smoother, (entry)&AMG_print_info,
    (entry)&AMG_print_stat, (entry)&AMG_recompute_coarse_matrices, (entry)&AMG_set_SOR_relaxation_coeff,
    (entry)&AMG_set_deg_of_freedom, (entry)&AMG_set_eps, (entry)&AMG_set_matrix_tolerance,
    (entry)&AMG_set_max_iter, (entry)&AMG_set_smoother, (entry)&AMG_setup,
    (entry)&AMG_solve, (entry)&BI_AMG_precon, (entry)&BI_get_AMG_setup_data,
    (entry)&BI_get_AMG_stat, (entry)&BI_get_max_iter, (entry)&BI_get_stat,
    (entry)&BI_info_set, (entry)&BI_print_info, (entry)&BI_residuum,
    (entry)&BI_set_AMG_preconditioner, (entry)&BI_set_AMG_setup_data, (entry)&BI_set_PCG_preconditioner,
    (entry)&BI_set_eps, (entry)&BI_set_mat_mult, (entry)&BI_set_max_iter,
    (entry)&BI_set_precon, (entry)&BI_set_stat, (entry)&BiCGStab_solve,
    (entry)&Free_sparse, (entry)&L_triang_invert, (entry)&Open_file,
    (entry)&PCG_default, (entry)&PCG_get_mat_mult, (entry)&PCG_get_precon,
    (entry)&PCG_set_AMG_preconditioner, (entry)&PCG_set_default_precon, (entry)&PCG_set_mat_mult,
    (entry)&PCG_set_precon, (entry)&RBDelete, (entry)&RBEnumerate,
    (entry)&RBExactQuery, (entry)&RBLargest, (entry)&RBTreeCount,
    (entry)&RBTreeCreate, (entry)&RBTreeDestroy, (entry)&RBTreeInsert,
    (entry)&RBTreePrint, (entry)&RN_default_mat_mult, (entry)&RN_info_set,
    (entry)&RN_print_function, (entry)&RN_print_info, (entry)&RN_set_AMG_data,
    (entry)&RN_set_Hesse_function, (entry)&RN_set_debug_mode, (entry)&RN_set_default_matrix,
    (entry)&RN_set_gradient, (entry)&RN_set_kernel, (entry)&RN_set_local_solver,
    (entry)&RN_set_params, (entry)&RN_set_solver, (entry)&RN_set_trust_limits,
    (entry)&RN_set_trust_region, (entry)&TreePredecessor, (entry)&TreeSuccessor,
    (entry)&U_triang_invert, (entry)&active_set_solver, (entry)&add_rank4_tensor_product,
    (entry)&add_subsparse_to_sparse, (entry)&append_data, (entry)&clone,
    (entry)&clone_list, (entry)&clone_vector, (entry)&complete_LU_factorization,
    (entry)&convert_array_to_sparse, (entry)&convert_sparse_to_UMFPACK, (entry)&copy_vector_content,
    (entry)&copy_vector_to, (entry)&diagonal_preconditioner, (entry)&divide_and_conquer_sort,
    (entry)&enlarge_matrix, (entry)&euklid_norm, (entry)&fast_sparse_product,
    (entry)&find_position_in_row, (entry)&free_UMFPACK_matrix_info, (entry)&free_ass_info,
    (entry)&free_sparse_index_map, (entry)&gauss_seidel, (entry)&generate_vector,
    (entry)&get_ID_index_map, (entry)&get_all_indices, (entry)&get_col_sums,
    (entry)&get_mean, (entry)&get_most_negative_eigenvector, (entry)&get_pcg_residuum,
    (entry)&get_position, (entry)&get_test_matrix, (entry)&get_transpose,
    (entry)&get_transpose3D_13, (entry)&gmres, (entry)&hp_euklid_norm,
    (entry)&in_list, (entry)&incomplete_LU_factorization, (entry)&insert_sparse,
    (entry)&int_sort, (entry)&join_lists, (entry)&join_vectors,
    (entry)&linear_map, (entry)&map_back, (entry)&matrix_norm,
    (entry)&matrix_residuum, (entry)&max, (entry)&new_sparse_general,
    (entry)&normalize_row_sum, (entry)&open_file, (entry)&pcg,
    (entry)&print_RN_statistics, (entry)&print_scalar_data, (entry)&print_sparse,
    (entry)&print_to_matrix_market_format, (entry)&print_vector, (entry)&rank4_product,
    (entry)&rank4_tensor_product, (entry)&readLine, (entry)&read_sparse_textfile,
    (entry)&readline, (entry)&reflective_newton, (entry)&remove_element_at,
    (entry)&remove_zeros, (entry)&reset_row, (entry)&restrict_matrix_cols,
    (entry)&restrict_matrix_rows, (entry)&restrict_vector, (entry)&restrict_vector_by,
    (entry)&scalar, (entry)&scalar_mult, (entry)&scalar_sparse_mult,
    (entry)&set_RN_Hesse, (entry)&set_RN_constraints, (entry)&set_RN_functional,
    (entry)&set_RN_linear, (entry)&set_RN_params, (entry)&set_connection_thresholds,
    (entry)&set_omp_chunk_size, (entry)&set_product_pattern, (entry)&set_thread_num,
    (entry)&simulated_annealing_min, (entry)&sparse3D_Bv, (entry)&sparse3D_Bvv,
    (entry)&sparse3D_bilinear, (entry)&sparse3D_merge, (entry)&sparse3D_transpose13,
    (entry)&sparse3D_vB, (entry)&sparse_LU_factorization, (entry)&sparse_add,
    (entry)&sparse_bilinear, (entry)&sparse_diagonal, (entry)&sparse_general_append,
    (entry)&sparse_identity, (entry)&sparse_mult, (entry)&sparse_multiplication,
    (entry)&sparse_product, (entry)&sparse_zero, (entry)&sparse_zero3D,
    (entry)&split, (entry)&subspace_minimum, (entry)&vec_dist,
    (entry)&vector_add, (entry)&vector_contraction, (entry)&vector_normalize,
    (entry)&vector_pseudo_div, (entry)&vector_pseudo_mult, (entry)&vector_shift,
    (entry)&zero_int_list, (entry)&zero_vector,
};

int main(void) {
    size_t i, n = 0;
    for (i = 0; i < sizeof(table) / sizeof(table[0]); i++) n += table[i] != NULL;
    printf("%zu\n", n);
    return 0;
}
