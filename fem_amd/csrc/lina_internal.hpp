// Internal host-side types of liblina_hip (not part of the C ABI).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace lina {

// Label values of the C/F splitting (LinA/Multigrid.h:25-29)
enum { UNDECIDED = 0, COARSE = 1, FINE = 2, ISOLATED = 3 };
enum { AGGRESSIVE = 1, STANDARD = 2 };

// Compressed sparse rows, int32 indices, columns sorted inside each row.
struct Csr {
    int n = 0;        // rows
    int ncols = 0;    // columns
    std::vector<int> rp;
    std::vector<int> ci;
    std::vector<double> v;
    int nnz() const { return rp.empty() ? 0 : rp[n]; }
};

// One level of the AMG hierarchy as built by the host setup.  Index `l`
// follows the reference: depth-1 is the finest level, 0 the coarsest.
struct HostLevel {
    Csr A;                    // Coarse_A[l]
    Csr P;                    // Prolongation[l]  (fine x coarse), l >= 1
    Csr R;                    // Restriction[l]   (coarse x fine), l >= 1
    std::vector<int> labels;  // Coarsening[l], l >= 1
    std::vector<int> levels;  // dof component of each row (Levels[] of the reference)
    int fine = 0, coarse = 0;
};

// Dense LU of the coarsest level (sparse_LU_factorization, linear_algebra.c:3040-3062)
struct CoarseLU {
    int n = 0;                       // full size of level 0
    std::vector<int> dense_idx;      // rows/cols that enter the dense system
    std::vector<int> trivial_idx;    // rows with <= 1 nonzero (L=1, U=diag)
    std::vector<double> trivial_diag;
    std::vector<double> dense;       // dense system, row-major, nd x nd
    // block-diagonal coarsest level (a batch of independent systems): block b
    // owns dense_idx[blk_off[b] .. blk_off[b+1]) and the row-major square
    // matrix dense[mat_off[b] ..); empty: one block
    std::vector<int> blk_off;
    std::vector<size_t> mat_off;
};

struct SetupParams {
    int dof = 1;
    int depth = 4;
    int smoothing_iterations = 2;
    int aggr_smoothing_iterations = 3;
    int coarsening = AGGRESSIVE;
    double eps_minus = 0.3;          // set_connection_thresholds defaults (Multigrid.c:35-36)
    double eps_plus = 0.5;
    double matrix_tol = 0.1;         // AMG_tol (Multigrid.c:37)
};

struct HostHierarchy {
    SetupParams prm;
    std::vector<HostLevel> lv;       // size depth
    std::vector<double> signs;       // diagonal of Signs (make_positive_diag)
    std::vector<int> iso_rows;       // ISOLATED rows of the finest level
    std::vector<double> iso_a;       // first nonzero of those rows (AMG_fix_isolated)
    CoarseLU lu;
    std::vector<std::string> warnings;
};

// setup (amg_setup.cpp)
void amg_setup(const Csr& A, const SetupParams& prm, HostHierarchy& H);
void amg_recompute(HostHierarchy& H, const Csr& A_new);  // AMG_recompute_coarse_matrices
double get_diag(const Csr& A, int i);
void sparse_product(const Csr& A, const Csr& B, int b_ncols, Csr& C);   // sparse_product semantics
void galerkin(const Csr& A, const Csr& R, const Csr& P, Csr& Ac);        // get_restricted_matrix
void coarse_lu(const Csr& A0, CoarseLU& lu);
std::vector<int> greedy_colouring(const Csr& A, int* ncolours);
int ilu0(const Csr& A, Csr& L, Csr& U);                                  // incomplete_LU_factorization

// error reporting shared by all entry points
void set_error(const char* fmt, ...);

}  // namespace lina
