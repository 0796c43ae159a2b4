// Device-resident solver object behind the lina_solver handle.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lina_hip.h"
#include "kernels.hpp"
#include "lina_internal.hpp"

namespace lina {

// argument checks of an lina_amg_params block (throws std::invalid_argument)
void check_params(const lina_amg_params* p);


struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NumericError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);
const char* last_error();

struct DevBuffers {
    std::vector<void*> ptrs;
    ~DevBuffers();
};

// device arrays owned by a DevBuffers (freed with it)
template <typename T>
T* dalloc(size_t n, DevBuffers& owner) {
    if (n == 0) n = 1;
    void* p = nullptr;
    hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    owner.ptrs.push_back(p);
    return (T*)p;
}

template <typename T>
T* dupload(const std::vector<T>& h, DevBuffers& owner, hipStream_t st) {
    T* d = dalloc<T>(h.size(), owner);
    if (!h.empty()) hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
    return d;
}

// The plans of the device Galerkin products (rap.cpp) on the device: per
// fine level l >= 1 the product lists of A P and R (A P), P's and R's
// values at the plans' positions, the coarse entries' CSR positions; the
// colour-permuted copies' source positions; the coarsest level's dense and
// trivial-row maps; a pattern-change flag.
struct DeviceRap {
    struct Lv {
        int nap = 0, nac = 0;
        int64_t *ap_ptr = nullptr, *ac_ptr = nullptr;
        int *ap_a = nullptr, *ap_p = nullptr, *ac_r = nullptr, *ac_ap = nullptr, *pos = nullptr;
        uint8_t *ap_seed = nullptr, *ac_seed = nullptr, *lone = nullptr;
        double *pv = nullptr, *rv = nullptr, *ap_v = nullptr, *ac_v = nullptr;
    };
    std::vector<Lv> lv;
    std::vector<int*> acmap;
    int* dense_map = nullptr;
    int* tdiag_map = nullptr;
    int* flag = nullptr;
    int* host_flag = nullptr;  // pinned
    double bytes = 0;
    DevBuffers bufs;
    ~DeviceRap() {
        if (host_flag) (void)hipHostFree(host_flag);
    }
};

struct DevLevel {
    int n = 0;
    dev::DMat<double> A, P, R, Ac;
    dev::DMat<float> Af, Pf, Rf, Acf;
    // multicolour sweeps: Ac holds the rows grouped by colour
    int ncolours = 0;
    int* perm = nullptr;
    std::vector<int> colour_rows;  // ncolours+1 row offsets into Ac
    std::vector<int> colour_blk;   // ncolours+1 block offsets into Ac's blocks
    std::vector<int> colour_blk32; // the same for Acf (FP32 blocks hold more entries)
    // exact lexicographic sweeps
    int nlv_f = 0, nlv_b = 0;
    int *lvf_off = nullptr, *lvf_rows = nullptr, *lvb_off = nullptr, *lvb_rows = nullptr;
    int nlim = 0;  // rows touched by AMG_parallel_Jacobi_smoother: (n/dof)*dof
    // level-packed exact lexicographic sweeps (FP64), forward and backward
    bool sor_packed = false;
    dev::SorPack sorf, sorb;
    std::vector<int> sor_sizes_f, sor_sizes_b;  // overflow sizes the packs were allocated for
    // work vectors
    double *x = nullptr, *f = nullptr, *r = nullptr, *xo = nullptr;
    float *xf = nullptr, *ff = nullptr, *rf = nullptr, *xof = nullptr;
};

int upload_pair(const Csr& A0, dev::DMat<double>& D, dev::DMat<float>& F, DevBuffers& B, hipStream_t st,
                const std::vector<int>* seg = nullptr);
void build_blocks(const Csr& A, int begin, int end, int max_nnz, std::vector<int>& starts,
                  int max_rows = dev::BLOCK_ROWS);
std::vector<int2> block_pairs(const Csr& A, const std::vector<int>& starts_with_end);
// algorithmic bytes of one SpMV with 32-bit indices (before upload): 12 nnz + 4 (n+1) + 16 n (FP64),
// 8 nnz + 4 (n+1) + 8 n (FP32)
double spmv_bytes(double n, double nnz, int precision);

// A device-resident matrix (FP64 and FP32 copies) for y = A x outside a
// hierarchy: the library's own matrix-vector product of the Krylov drivers.
class DevMatrix {
  public:
    explicit DevMatrix(const Csr& A);
    ~DevMatrix();
    // new values of the same pattern: uploads the FP64 values and converts
    // the FP32 copy on the device, keeping the row blocks and column codes
    // (the rows' l1 norms, which no product reads, keep the first values);
    // returns false, changing nothing, for another pattern
    bool update_values(const Csr& A);
    // the same from nnz values in device memory, in the uploaded pattern's
    // entry order (the caller vouches for the pattern); false when nnz differs
    bool update_values_device(const double* d_val, long nnz);
    int n() const { return n_; }
    long nnz() const { return nnz_; }
    void apply(const double* x, double* y);  // enqueue on stream(), FP64
    // y = A x / b - A x / y + A x (mode) on stream(); part != nullptr: block
    // partials of sum_r y_r w_r, returns their number
    int launch(const double* x, const double* b, double* y, int mode, double* part, const double* w);
    int launch_on(hipStream_t st, const double* x, const double* b, double* y, int mode, double* part,
                  const double* w);
    int nblocks() const { return A_.nblk; }
    void spmv_dev(const double* x, double* y);
    void spmv_host(int precision, const double* x, double* y);
    double time_spmv(int precision, int reps);
    double bytes(int precision) const;
    // matrix_norm (linear_algebra.c:3441-3450): sum of |a_ij| in row order / n
    double matrix_norm() const;
    hipStream_t stream() const { return stream_; }

  private:
    int n_ = 0;
    long nnz_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    std::unique_ptr<DevBuffers> bufs_;
    dev::DMat<double> A_;
    dev::DMat<float> Af_;
    std::vector<int> prp_, pci_;  // the uploaded pattern
    double *x_ = nullptr, *y_ = nullptr;
    float *xf_ = nullptr, *yf_ = nullptr;
};

// power_method (linear_algebra.c:4802-4822) on the device (rn.cpp)
int power_iteration(hipStream_t st, DevMatrix& A, double q, double* v, double* tmp, double* part, double* scal,
                    double* host, int max_iter, double eps, double* eig);

class Solver {
  public:
    Solver();
    ~Solver();
    void setup_host(const Csr& A, const lina_amg_params& p);
    // a hierarchy built elsewhere (the block-diagonal union of a batch)
    void set_hierarchy(HostHierarchy&& H, const lina_amg_params& p);
    // row-partition autotuning (run by upload when params.autotune != 0)
    void autotune();
    struct Tuned {
        int level;
        char which;  // A, P, R, C (colour-permuted copy of A)
        int lpr64, lpr32;
    };
    const std::vector<Tuned>& tuned() const { return tuned_; }
    void upload();
    void recompute(const std::vector<double>& val);
    // the same from finest values already on the device (d_val: nnz values
    // in the setup's pattern, before its diagonal signs); the host copy of
    // the hierarchy catches up when read (host_synced)
    void recompute_device_values(const double* d_val);
    // AMG_recompute_coarse_matrices with a new finest matrix of any pattern
    void recompute_matrix(const Csr& M);
    void set_params(const lina_amg_params& p);  // takes effect at the next solve
    // a batch's row offsets of its systems on every level (before upload), and
    // the systems' active flags (device, after upload): row blocks whose
    // systems are all inactive are skipped by the V-cycle's row kernels
    void set_segments(std::vector<std::vector<int>> off);
    void set_active(const double* active);
    const lina_amg_params& params() const { return prm_; }
    const HostHierarchy& host() const { return H_; }
    // the host hierarchy with the values of a device recompute copied back
    const HostHierarchy& host_synced();
    const std::vector<int>& colours(int l) const { return colours_.at(l); }
    bool uploaded() const { return bufs_ != nullptr; }
    int n() const { return n_; }
    // rows of the coarsest level solved by the dense inverse / by their diagonal
    int coarse_dense_rows() const { return nd_; }
    int coarse_trivial_rows() const { return nt_; }

    int pcg_dev(double* X, double* B, double tol, int max_iter, int stop, lina_stats* st);
    int pcg_host(double* x, double* b, double tol, int max_iter, int stop, lina_stats* st);
    int amg_solve_dev(const double* F, double* X, double* relres = nullptr);
    int amg_solve_host(const double* f, double* x, double* relres = nullptr);
    // Z = M^-1 R; w != nullptr: block partials of w.Z in part(), their number in *np
    int precon(const double* R, double* Z, const double* w = nullptr, int* np = nullptr, bool prepared = false);
    int vcycle_once(const double* F, double* X, const double* w = nullptr, int* np = nullptr, bool prepared = false);
    bool precon_async() const;
    bool entry_prepared_ok() const;
    int spmv_host(int level, int precision, const double* x, double* y);
    double time_spmv(int level, int precision, int reps);
    double spmv_bytes(int level, int precision) const;
    double pcg_residuum() const { return pcg_residuum_; }
    hipStream_t stream() const { return stream_; }

    // state of a level's iterate on entry to smooth / vcycle
    enum XInit { X_GUESS = 0, X_ZERO = 1, X_J0 = 2, X_SPARE = 3 };
    void level_smoother(int l, int& smoother, int& sweeps) const;
    void vcycle_smoother(int l, int& smoother, int& sweeps) const;
    template <typename T>
    void smooth(int l, const T* f, T* x, int sweeps, bool post, int xinit = X_GUESS);
    template <typename T>
    void vcycle(int l, const T* f, T* x, int xinit = X_GUESS);
    template <typename T>
    void fmg(const T* f, T* x);
    template <typename T>
    double residual_norm(const T* f, const T* x, T* r, int slot);
    double read_scalar(int slot);

    // device buffers shared with other drivers (BiCGStab, GMRES, RN)
    double* part() { return part_; }
    double* scal() { return scal_; }
    DevLevel& finest() { return L_.back(); }
    double* dalloc_public(size_t n);
    double dot_host(const double* a, const double* b);
    void apply_A(const double* x, double* y);
    int bicgstab_dev(double* X, const double* b, double eps, int max_iter, double* res);
    int bicgstab_host(double* x, const double* b, double eps, int max_iter, double* res);
    int last_vcycles() const { return vcycles_; }
    void set_kernel_timing(bool on);
    // pcg's operator (PCG_set_mat_mult with a device matrix): the hierarchy
    // stays the preconditioner; nullptr restores the hierarchy's own matrix
    void set_operator(DevMatrix* op);
    void elasticity_gradient(int n_nodes, const double* xy, int n_elem, const int* tri, double nu,
                             const double* scale, int blocked, const double* u, const double* lam, double* grad);

  private:
    // recompute on the device (rap.cpp): the Galerkin products from a plan
    // of each level's products, then every value-derived array; false (the
    // caller takes the host path) when it does not apply or a level's
    // pattern changes
    bool device_rap_ok() const;
    bool recompute_device(const Csr& A);
    bool recompute_device_core();
    void fetch_isolated_rows();
    void build_rap();
    void sync_host_levels();
    // the isolated rows' divisors (AMG_fix_isolated) from the finest values;
    // false when they need a full upload (a zero divisor appears or goes)
    bool refresh_isolated();
    std::unique_ptr<DeviceRap> rap_;
    bool host_stale_ = false;
    bool host_top_stale_ = false;  // the finest host values lag the device's
    int* neg_rows_ = nullptr;      // rows with a negative setup sign (device)
    int nneg_ = 0;
    int* iso_pos_ = nullptr;       // entry positions of the isolated rows (device)
    double* iso_val_ = nullptr;
    int niso_pos_ = 0;
    std::vector<int> iso_pos_host_;
    // recompute: colour the new levels and re-upload them (values only when
    // every pattern-derived structure matches what is on the device)
    void refresh_after_recompute();
    bool refresh_values();
    bool lex_smoothing() const;
    void build_sor_packs(int l, DevBuffers& B, bool reuse);
    struct UpLevel {  // host copy of what upload() derived from a level's pattern
        std::vector<int> rp, ci, colour, perm, lvf_off, lvf_rows, lvb_off, lvb_rows;
    };
    std::vector<UpLevel> up_;  // kept only once a recompute has happened (track_)
    bool track_ = false;
    std::vector<int> up_dense_, up_trivial_;
    double* cpiv_ = nullptr;
    bool fused_gather() const;
    void ensure_ctiles();
    template <typename VT, typename XT>
    int tune_spmv(dev::DMat<VT>& M, const XT* x, XT* y);
    template <typename VT, typename XT>
    int tune_mcsor(DevLevel& d, dev::DMat<VT>& Ac, const XT* f, XT* x);
    std::vector<Tuned> tuned_;
    std::vector<std::vector<int>> seg_off_;
    const double* seg_active_ = nullptr;
    lina_amg_params prm_{};
    HostHierarchy H_;
    std::vector<std::vector<int>> colours_;
    std::vector<int> ncolours_;
    int n_ = 0;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    double* host_scal_ = nullptr;
    std::unique_ptr<DevBuffers> bufs_;
    std::vector<DevLevel> L_;
    double* signs_ = nullptr;
    bool has_neg_signs_ = false;
    int* iso_rows_ = nullptr;
    double* iso_a_ = nullptr;
    int niso_ = 0;
    int nd_ = 0, nt_ = 0;
    int *didx_ = nullptr, *tidx_ = nullptr;
    double *inv_ = nullptr, *tdiag_ = nullptr;
    bool cbatched_ = false;
    double* cbd_ = nullptr;
    int* dinv_ = nullptr;     // dense index of each coarsest-level row (-1: not dense)
    bool ccover_ = false;     // dense and trivial rows cover the coarsest level
    double* ctiles_ = nullptr;  // upper-triangle FP64 tiles of the coarse inverse
    int nct_ = 0, ntiles_ = 0;  // tiles a side, stored tiles
    int *cti_ = nullptr, *ctj_ = nullptr;  // tile coordinates
    double* cpart_ = nullptr;   // nct_ x nct_ x CT tile partials
    double* isa_ = nullptr;   // isolated rows' diagonals by row (0 elsewhere); nullptr without isolated rows
    uint32_t* isb_ = nullptr; // bit i: row i is isolated (read instead of isa_ on every row)
    bool isa_ok_ = true;      // no isolated row has a zero diagonal (fused preconditioner entry/exit)
    int *crblk_ = nullptr, *coff_ = nullptr;
    size_t* cmoff_ = nullptr;
    double *ftop_ = nullptr, *rtop_ = nullptr;
    double *kR_ = nullptr, *kP_ = nullptr, *kQ_ = nullptr, *kZ_ = nullptr, *kW_ = nullptr, *kX_ = nullptr,
           *kB_ = nullptr;
    double* part_ = nullptr;
    double* scal_ = nullptr;
    int vcycles_ = 0, precon_calls_ = 0;
    double pcg_residuum_ = 0;
    bool ktime_ = false;
    DevMatrix* op_ = nullptr;
    int part_cap_ = 0;
    hipEvent_t evk_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // outer SpMV timing, per ring slot
    hipEvent_t evr_[2] = {nullptr, nullptr};                            // pcg iteration done, per ring slot
    double* host_ring_ = nullptr;                                       // pinned: 2 x S_COUNT scalars
    double *bS_ = nullptr, *bPS_ = nullptr, *bPP_ = nullptr, *bAPP_ = nullptr, *bAPS_ = nullptr, *bR_ = nullptr,
           *bR0_ = nullptr, *bP_ = nullptr;
};

}  // namespace lina
