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
    // PCG_default(A, L, U) with factors computed elsewhere (L: diagonal last
    // in each row, U: diagonal first); only the preconditioner is available
    void set_factors(const Csr& L, const Csr& U, bool upload = true);
    void upload(bool pattern_levels = false);
    bool ilu_ok() const { return ilu_ok_ != 0; }
    // the factors (downloaded first when they were computed on the device)
    const Csr& L() { sync_factors(); return L_; }
    const Csr& U() { sync_factors(); return U_; }
    int n() const { return n_; }

    void apply(const double* z, double* w);  // w = U^-1 L^-1 z (device)
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
