// ILU(0)-preconditioned Krylov solvers on the device (krylov.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

namespace lina {

class IluSystem {
  public:
    IluSystem();
    ~IluSystem();
    // incomplete_LU_factorization on the host (bit-identical factors), then upload
    void build(const Csr& A, bool upload = true);
    // the same factorisation on the device (bit-identical factors) for a
    // matrix whose rows are sorted by column: the symbolic structure (L/U
    // patterns, level schedules, device buffers) is kept while A's pattern
    // stays the same, so a refactorisation uploads A's values and runs the
    // numeric kernel only.  Unsorted rows fall back to build().
    void refactor(const Csr& A);
    // the same numeric factorisation from values already in device memory
    // (d_val: the last refactor's matrix, in its own order and pattern);
    // false, changing nothing, when no device structure is in place
    bool refactor_device(const double* d_val);
    // PCG_default(A, L, U) with factors computed elsewhere (L: diagonal last
    // in each row, U: diagonal first); only the preconditioner is available
    void set_factors(const Csr& L, const Csr& U, bool upload = true);
    void upload(bool pattern_levels = false);
    // 0: the natural order (the reference's factors); 1: multicolour order --
    // the rows and columns permuted by a greedy colouring of A's pattern
    // (colour by colour, natural order within a colour), so that both
    // triangular sweeps and the factorisation run in a few wide levels on
    // the whole GPU.  The factors, the operator and the Krylov iterations are
    // those of P A P^T (factors exactly the reference's incomplete LU of
    // P A P^T); every vector entry point takes and returns the caller's
    // (natural) order.  Takes effect at the next build / refactor.
    void set_ordering(int ordering);
    int ordering() const { return ordering_; }
    const std::vector<int>& permutation() const { return perm_; }  // empty in the natural order
    bool ilu_ok() const { return ilu_ok_ != 0; }
    // the factors (downloaded first when they were computed on the device)
    const Csr& L() { sync_factors(); return L_; }
    const Csr& U() { sync_factors(); return U_; }
    int n() const { return n_; }

    void apply(const double* z, double* w);  // w = U^-1 L^-1 z (device, the caller's order)
    void matvec(const double* x, double* y);
    double dot(const double* a, const double* b);
    int pcg_dev(double* X, double* B, double tol, int max_iter, int stop, double* residuum);
    int gmres_dev(double* Y, const double* B, int m, double tol, int max_iter, double* eta);
    int pcg_host(double* x, double* b, double tol, int max_iter, int stop, double* res);
    int gmres_host(double* x, const double* b, int m, double tol, int max_iter, double* eta);
    void apply_host(const double* z, double* w);
    hipStream_t stream() const { return stream_; }
    int levels_l() const { return nlv_l_; }
    int levels_u() const { return nlv_u_; }

    // a factor packed by level for dev::tri_solve_pair
    struct TriDev {
        dev::TriPack p;
        int nov = 0;
        int *vmap = nullptr, *dmap = nullptr, *ovmap = nullptr;
        void* rec = nullptr;
        double* ovval = nullptr;
    };

  private:
    void build_p(const Csr& A, bool upload);
    void refactor_p(const Csr& A);
    void apply_p(const double* z, double* w);  // in the factors' (permuted) order
    int pcg_core(double* X, double* B, double tol, int max_iter, int stop, double* residuum);
    int gmres_core(double* Y, const double* B, int m, double tol, int max_iter, double* eta);
    Csr permuted(const Csr& A);  // P A P^T (the permutation kept while A's pattern holds)
    int ordering_ = 0;
    std::vector<int> perm_, vpos_;  // row r of P A P^T is row perm_[r] of A; its entry k is A's entry vpos_[k]
    Csr opat_, ppat_;               // A's pattern the permutation was built for, and P A P^T's pattern
    int* dperm_ = nullptr;
    int* dvpos_ = nullptr;          // vpos_ on the device (values of A into P A P^T)
    void numeric();
    double *vXo_ = nullptr, *vBo_ = nullptr, *vPz_ = nullptr, *vPw_ = nullptr;
    bool wide_ = false;             // few, wide levels: a launch per level on the whole GPU
    std::vector<int> hoff_u_;
    void sync_factors();
    void pack_values();
    bool build_plan();
    // device factorisation plan (dev::ilu0_planned), built with the symbolic structure
    bool plan_ = false;
    int4 *pm1_ = nullptr, *pm2_ = nullptr;
    int *pmask_ = nullptr, *ppu_ = nullptr;
    std::vector<int> hoff_l_, hrows_l_;  // L's level schedule (host copy)
    bool packed_ = false;          // tf_/tb_ hold the level packs (else the k_trsv sweeps)
    TriDev tf_, tb_;
    int* permL_ = nullptr;         // row of each forward phase position
    double *vZp_ = nullptr, *vYu_ = nullptr;
    Csr A_, L_, U_;
    int ilu_ok_ = 0;
    bool device_ready_ = false;    // refactor's symbolic structure is uploaded for A_'s pattern
    bool host_factors_ = true;     // L_, U_ hold the current values
    int* fail_ = nullptr;          // zero-pivot flag of the device factorisation
    int n_ = 0;
    hipStream_t stream_ = nullptr;
    double* host_ = nullptr;
    std::unique_ptr<DevBuffers> bufs_;
    dev::DMat<double> dA_, dL_, dU_;
    int nlv_l_ = 0, nlv_u_ = 0;
    int *lvl_off_ = nullptr, *lvl_rows_ = nullptr, *lvu_off_ = nullptr, *lvu_rows_ = nullptr;
    double *vX_ = nullptr, *vB_ = nullptr, *vR_ = nullptr, *vP_ = nullptr, *vQ_ = nullptr, *vZ_ = nullptr,
           *vW_ = nullptr, *vY_ = nullptr, *vT_ = nullptr;
    double* part_ = nullptr;
    double* scal_ = nullptr;
    double* vbasis_ = nullptr;
    double* hcols_ = nullptr;  // a GMRES cycle's Gram-Schmidt columns on the device, (m+2) doubles each
    int mb_ = 0;
};

}  // namespace lina
