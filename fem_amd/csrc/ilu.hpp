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
    // PCG_default(A, L, U) with factors computed elsewhere (L: diagonal last
    // in each row, U: diagonal first); only the preconditioner is available
    void set_factors(const Csr& L, const Csr& U, bool upload = true);
    void upload();
    bool ilu_ok() const { return ilu_ok_ != 0; }
    const Csr& L() const { return L_; }
    const Csr& U() const { return U_; }
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

  private:
    Csr A_, L_, U_;
    int ilu_ok_ = 0;
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
    int mb_ = 0;
};

}  // namespace lina
