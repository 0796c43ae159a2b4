// C ABI of the handle API (include/lina_hip.h).  Every entry point catches
// C++ exceptions and turns them into a negative return code plus a message
// for lina_last_error(); nothing here exits the process.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>

#include "../../include/lina_hip.h"
#include "lina_internal.hpp"
#include "solver.hpp"

using lina::Csr;
using lina::Solver;

struct lina_solver {
    Solver s;
};
struct lina_matrix {
    explicit lina_matrix(const Csr& A) : m(A) {}
    lina::DevMatrix m;
};

namespace {

int fail(const std::exception& e, int code) {
    lina::set_error("%s", e.what());
    return code;
}

template <typename F>
int guard(F&& f) {
    try {
        return f();
    } catch (const lina::HipError& e) {
        return fail(e, LINA_EHIP);
    } catch (const lina::NumericError& e) {
        return fail(e, LINA_ENUMERIC);
    } catch (const std::invalid_argument& e) {
        return fail(e, LINA_EARG);
    } catch (const std::out_of_range& e) {
        return fail(e, LINA_EARG);
    } catch (const std::runtime_error& e) {
        return fail(e, LINA_ESETUP);
    } catch (const std::exception& e) {
        return fail(e, LINA_ESETUP);
    }
}

Csr make_csr(int n, const int* rowptr, const int* col, const double* val) {
    if (n < 0 || !rowptr || (n > 0 && (!col || !val))) throw std::invalid_argument("null CSR array");
    Csr A;
    A.n = n;
    A.ncols = n;
    A.rp.assign(rowptr, rowptr + n + 1);
    if (A.rp[0] != 0) throw std::invalid_argument("rowptr[0] != 0");
    for (int i = 0; i < n; i++) {
        if (A.rp[i + 1] <= A.rp[i]) throw std::invalid_argument("every row needs at least one entry");
    }
    int nnz = A.rp[n];
    A.ci.assign(col, col + nnz);
    A.v.assign(val, val + nnz);
    for (int i = 0; i < n; i++)
        for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
            if (A.ci[p] < 0 || A.ci[p] >= n) throw std::invalid_argument("column index out of range");
            if (p > A.rp[i] && A.ci[p] <= A.ci[p - 1]) throw std::invalid_argument("columns not sorted/unique");
        }
    return A;
}

}  // namespace

namespace lina {
void check_params(const lina_amg_params* p) {
    if (!p) throw std::invalid_argument("null params");
    if (p->dof < 1 || p->depth < 2) throw std::invalid_argument("dof >= 1 and depth >= 2 required");
    if (p->smoother < 0 || p->smoother > 5) throw std::invalid_argument("unknown smoother");
    if (p->coarse_smoother < -1 || p->coarse_smoother > 5) throw std::invalid_argument("unknown coarse smoother");
    if (p->coarse_smoother >= 0 && p->coarse_sweeps < 0) throw std::invalid_argument("negative sweeps");
    if (p->precision != LINA_PREC_FP64 && p->precision != LINA_PREC_MIXED)
        throw std::invalid_argument("unknown precision");
    if (p->cycle != LINA_CYCLE_V && p->cycle != LINA_CYCLE_FULL) throw std::invalid_argument("unknown cycle");
}
}  // namespace lina

extern "C" {

const char* lina_last_error(void) { return lina::last_error(); }

int lina_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void lina_params_reference(lina_amg_params* p, int dof, int depth) {
    std::memset(p, 0, sizeof(*p));
    p->dof = dof;
    p->depth = depth;
    p->smoothing_iterations = 2;
    p->aggr_smoothing_iterations = 3;
    p->coarsening = LINA_COARSEN_AGGRESSIVE;
    p->eps_minus = 0.3;
    p->eps_plus = 0.5;
    p->matrix_tol = 0.1;
    p->smoother = LINA_SMOOTHER_SOR;
    p->omega = 1.6;
    p->amg_eps = 1e-2;
    p->amg_max_iter = 100;
    p->zero_init = 0;
    p->precision = LINA_PREC_FP64;
    p->coarse_smoother = -1;
    p->coarse_sweeps = 1;
}

void lina_params_vcycle(lina_amg_params* p, int dof, int depth) {
    lina_params_reference(p, dof, depth);
    p->smoothing_iterations = 1;
    p->aggr_smoothing_iterations = 1;
    p->smoother = LINA_SMOOTHER_MC_SSOR;
    p->omega = 1.0;
    p->amg_max_iter = 1;
    p->zero_init = 1;
    p->coarse_smoother = LINA_SMOOTHER_L1_JACOBI;
    p->coarse_sweeps = 1;
}

lina_solver* lina_amg_create(int n, const int* rowptr, const int* col, const double* val, const lina_amg_params* p,
                             int upload) {
    lina_solver* h = nullptr;
    int rc = guard([&]() {
        lina::check_params(p);
        Csr A = make_csr(n, rowptr, col, val);
        h = new lina_solver();
        h->s.setup_host(A, *p);
        if (upload) {
            if (lina_device_count() < 1) throw lina::HipError("no HIP device available");
            h->s.upload();
        }
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void lina_amg_destroy(lina_solver* s) { delete s; }

int lina_amg_set_params(lina_solver* s, const lina_amg_params* p) {
    return guard([&]() {
        lina::check_params(p);
        const lina_amg_params& o = s->s.params();
        if (p->dof != o.dof || p->depth != o.depth || p->coarsening != o.coarsening ||
            p->eps_minus != o.eps_minus || p->eps_plus != o.eps_plus || p->matrix_tol != o.matrix_tol)
            throw std::invalid_argument("setup parameters cannot change after AMG setup");
        s->s.set_params(*p);
        return 0;
    });
}

int lina_amg_depth(const lina_solver* s) { return (int)s->s.host().lv.size(); }

int lina_amg_coarse_rows(const lina_solver* s, int* dense, int* trivial) {
    return guard([&]() {
        if (!s) throw std::invalid_argument("null solver");
        if (!s->s.uploaded()) throw std::invalid_argument("the hierarchy is not on the device");
        if (dense) *dense = s->s.coarse_dense_rows();
        if (trivial) *trivial = s->s.coarse_trivial_rows();
        return 0;
    });
}

int lina_amg_level_info(const lina_solver* s, int level, int* fine, int* coarse, int* nnz_a, int* nnz_p, int* nnz_r) {
    return guard([&]() {
        const auto& L = s->s.host().lv.at(level);
        if (fine) *fine = L.A.n;
        if (coarse) *coarse = L.coarse;
        if (nnz_a) *nnz_a = L.A.nnz();
        if (nnz_p) *nnz_p = level > 0 ? L.P.nnz() : 0;
        if (nnz_r) *nnz_r = level > 0 ? L.R.nnz() : 0;
        return 0;
    });
}

static const Csr& pick(const lina_solver* s, int level, char which) {
    // values as the device holds them (a device recompute leaves the host copy behind until read)
    const auto& H = const_cast<lina_solver*>(s)->s.host_synced();
    const auto& L = H.lv.at(level);
    switch (which) {
        case 'A': return L.A;
        case 'P': if (level > 0) return L.P; break;
        case 'R': if (level > 0) return L.R; break;
        default: break;
    }
    throw std::invalid_argument("unknown matrix selector");
}

int lina_amg_level_nnz(const lina_solver* s, int level, char which) {
    return guard([&]() { return pick(s, level, which).nnz(); });
}

int lina_amg_level_copy(const lina_solver* s, int level, char which, int* rowptr, int* col, double* val) {
    return guard([&]() {
        const Csr& M = pick(s, level, which);
        std::memcpy(rowptr, M.rp.data(), (M.n + 1) * sizeof(int));
        std::memcpy(col, M.ci.data(), M.nnz() * sizeof(int));
        std::memcpy(val, M.v.data(), M.nnz() * sizeof(double));
        return 0;
    });
}

int lina_amg_labels(const lina_solver* s, int level, int* labels) {
    return guard([&]() {
        const auto& L = s->s.host().lv.at(level);
        if (level == 0) throw std::invalid_argument("level 0 has no labels");
        std::memcpy(labels, L.labels.data(), L.labels.size() * sizeof(int));
        return 0;
    });
}

int lina_amg_colours(const lina_solver* s, int level, int* colours) {
    return guard([&]() {
        const auto& c = s->s.colours(level);
        std::memcpy(colours, c.data(), c.size() * sizeof(int));
        int m = 0;
        for (int v : c) m = v + 1 > m ? v + 1 : m;
        return m;
    });
}

int lina_amg_recompute(lina_solver* s, const double* val) {
    return guard([&]() {
        int nnz = s->s.host().lv.back().A.nnz();
        std::vector<double> v(val, val + nnz);
        s->s.recompute(v);
        return 0;
    });
}

int lina_amg_recompute_device(lina_solver* s, const double* d_val) {
    return guard([&]() {
        if (!s || !d_val) throw std::invalid_argument("null argument");
        s->s.recompute_device_values(d_val);
        return 0;
    });
}

int lina_amg_recompute_csr(lina_solver* s, int n, const int* rowptr, const int* col, const double* val) {
    return guard([&]() {
        if (!s) throw std::invalid_argument("null solver");
        Csr M = make_csr(n, rowptr, col, val);
        s->s.recompute_matrix(M);
        return 0;
    });
}

int lina_pcg(lina_solver* s, double* x, double* b, double tol, int max_iter, int stop, lina_stats* st) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.pcg_host(x, b, tol, max_iter, stop, st);
    });
}

int lina_pcg_device(lina_solver* s, double* d_x, double* d_b, double tol, int max_iter, int stop, lina_stats* st) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.pcg_dev(d_x, d_b, tol, max_iter, stop, st);
    });
}

int lina_amg_solve(lina_solver* s, const double* f, double* x) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.amg_solve_host(f, x);
    });
}

int lina_amg_solve_stat(lina_solver* s, const double* f, double* x, double* relres) {
    return guard([&]() {
        if (!s || !relres) throw std::invalid_argument("null argument");
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.amg_solve_host(f, x, relres);
    });
}

int lina_amg_signs(const lina_solver* s, double* signs) {
    return guard([&]() {
        if (!s || !signs) throw std::invalid_argument("null argument");
        const auto& H = s->s.host();
        const int n = H.lv.back().A.n;
        for (int i = 0; i < n; i++) signs[i] = H.signs.empty() ? 1.0 : H.signs[i];
        return n;
    });
}

int lina_amg_solve_device(lina_solver* s, const double* f, double* x) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        int it = s->s.amg_solve_dev(f, x);
        lina::hip_check(hipStreamSynchronize(s->s.stream()), "sync");
        return it;
    });
}

int lina_spmv(lina_solver* s, int level, int precision, const double* x, double* y) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.spmv_host(level, precision, x, y);
    });
}

double lina_time_spmv(lina_solver* s, int level, int precision, int reps) {
    double out = -1.0;
    int rc = guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        out = s->s.time_spmv(level, precision, reps);
        return 0;
    });
    return rc < 0 ? -1.0 : out;
}

double* lina_dev_alloc(long n) {
    void* p = nullptr;
    if (n < 1) n = 1;
    if (hipMalloc(&p, (size_t)n * sizeof(double)) != hipSuccess) {
        lina::set_error("hipMalloc of %ld doubles failed", n);
        return nullptr;
    }
    return (double*)p;
}

int lina_dev_free(double* d) { return hipFree(d) == hipSuccess ? 0 : LINA_EHIP; }

int lina_dev_copy(double* dst, const double* src, long n, int kind) {
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    if (n < 0 || kind < 0 || kind > 2) {
        lina::set_error("lina_dev_copy: bad arguments");
        return LINA_EARG;
    }
    if (hipMemcpy(dst, src, (size_t)n * sizeof(double), k) != hipSuccess ||
        (kind == 2 && hipDeviceSynchronize() != hipSuccess)) {
        lina::set_error("hipMemcpy failed");
        return LINA_EHIP;
    }
    return 0;
}

int lina_dev_fill(double* d, long n, double value) {
    return guard([&]() {
        lina::dev::fill<double>(nullptr, (int)n, d, value);
        lina::hip_check(hipDeviceSynchronize(), "fill");
        return 0;
    });
}

double lina_spmv_bytes(const lina_solver* s, int level, int precision) {
    return s->s.spmv_bytes(level, precision);
}

int lina_amg_tuned(lina_solver* s, int* level, char* which, int* lpr64, int* lpr32, int cap) {
    return guard([&]() {
        if (!s) throw std::invalid_argument("null solver");
        const auto& t = s->s.tuned();
        int m = std::min(cap, (int)t.size());
        for (int i = 0; i < m; i++) {
            level[i] = t[i].level;
            which[i] = t[i].which;
            lpr64[i] = t[i].lpr64;
            lpr32[i] = t[i].lpr32;
        }
        return (int)t.size();
    });
}

int lina_amg_kernel_timing(lina_solver* s, int on) {
    return guard([&]() {
        if (!s) throw std::invalid_argument("null solver");
        s->s.set_kernel_timing(on != 0);
        return 0;
    });
}

// ---- device vector primitives on the library's primitive stream
namespace {
// the primitives' stream and buffers: block partials, scalars, multidot
// partials (MULTI_MAX dots x its 512-block grid), coefficients; pinned host
// staging for results and coefficients
constexpr int MULTI_MAX = 64;
struct Prim {
    hipStream_t st = nullptr;
    double* part = nullptr;
    double* scal = nullptr;
    double* mpart = nullptr;
    double* coef = nullptr;
    double* host = nullptr;
    void init() {
        if (st) return;
        lina::hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        lina::hip_check(hipMalloc((void**)&part, 4096 * sizeof(double)), "hipMalloc");
        lina::hip_check(hipMalloc((void**)&scal, 8 * sizeof(double)), "hipMalloc");
        lina::hip_check(hipMalloc((void**)&mpart, (size_t)MULTI_MAX * 512 * sizeof(double)), "hipMalloc");
        lina::hip_check(hipMalloc((void**)&coef, MULTI_MAX * sizeof(double)), "hipMalloc");
        lina::hip_check(hipHostMalloc((void**)&host, (8 + MULTI_MAX) * sizeof(double)), "hipHostMalloc");
    }
};
Prim g_prim;
}  // namespace

double lina_dev_dot(const double* a, const double* b, long n) {
    double out = NAN;
    int rc = guard([&]() {
        if (n < 0 || n > INT32_MAX || (n > 0 && (!a || !b))) throw std::invalid_argument("bad arguments");
        g_prim.init();
        int np = lina::dev::dot_part<double>(g_prim.st, (int)n, a, b, g_prim.part);
        lina::dev::finalize(g_prim.st, np, g_prim.part, g_prim.scal, 0, 0);
        lina::hip_check(hipMemcpyAsync(g_prim.host, g_prim.scal, sizeof(double), hipMemcpyDeviceToHost, g_prim.st),
                        "d2h");
        lina::hip_check(hipStreamSynchronize(g_prim.st), "sync");
        out = g_prim.host[0];
        return 0;
    });
    return rc == 0 ? out : NAN;
}

int lina_dev_multidot(long n, int m, const double* V, long ld, const double* r, double* out) {
    return guard([&]() {
        if (n < 0 || n > INT32_MAX || m < 0 || m > MULTI_MAX || ld < n || (m > 0 && (!V || !r || !out)))
            throw std::invalid_argument("lina_dev_multidot: bad arguments (at most 64 dots, ld >= n)");
        if (m == 0) return 0;
        g_prim.init();
        lina::dev::multidot(g_prim.st, (int)n, m, V, (size_t)ld, r, g_prim.mpart, g_prim.coef);
        lina::hip_check(hipMemcpyAsync(g_prim.host, g_prim.coef, m * sizeof(double), hipMemcpyDeviceToHost,
                                       g_prim.st), "d2h");
        lina::hip_check(hipStreamSynchronize(g_prim.st), "sync");
        std::memcpy(out, g_prim.host, m * sizeof(double));
        return 0;
    });
}

int lina_dev_combine(long n, int m, const double* V, long ld, const double* c, const double* base, double sign,
                     double* y) {
    return guard([&]() {
        if (n < 0 || n > INT32_MAX || m < 0 || m > MULTI_MAX || ld < n || !base || !y || (m > 0 && (!V || !c)))
            throw std::invalid_argument("lina_dev_combine: bad arguments (at most 64 vectors, ld >= n)");
        g_prim.init();
        std::memcpy(g_prim.host, c, m * sizeof(double));
        lina::hip_check(hipMemcpyAsync(g_prim.coef, g_prim.host, m * sizeof(double), hipMemcpyHostToDevice,
                                       g_prim.st), "h2d");
        lina::dev::combine(g_prim.st, (int)n, m, V, (size_t)ld, g_prim.coef, base, sign, y);
        lina::hip_check(hipStreamSynchronize(g_prim.st), "sync");
        return 0;
    });
}

int lina_dev_axpby(long n, double a, double* x, double b, const double* y) {
    return guard([&]() {
        if (n < 0 || n > INT32_MAX || (n > 0 && (!x || !y))) throw std::invalid_argument("bad arguments");
        g_prim.init();
        lina::dev::axpby(g_prim.st, (int)n, a, x, b, y);
        lina::hip_check(hipStreamSynchronize(g_prim.st), "sync");
        return 0;
    });
}

int lina_dev_scale(long n, double* x, double a) {
    return guard([&]() {
        if (n < 0 || n > INT32_MAX || (n > 0 && !x)) throw std::invalid_argument("bad arguments");
        g_prim.init();
        lina::dev::scale(g_prim.st, (int)n, x, a);
        lina::hip_check(hipStreamSynchronize(g_prim.st), "sync");
        return 0;
    });
}

lina_matrix* lina_matrix_create(int n, const int* rowptr, const int* col, const double* val) {
    lina_matrix* h = nullptr;
    int rc = guard([&]() {
        Csr A = make_csr(n, rowptr, col, val);
        h = new lina_matrix(A);
        return 0;
    });
    return rc == 0 ? h : nullptr;
}

void lina_matrix_destroy(lina_matrix* h) { delete h; }

int lina_matrix_spmv(lina_matrix* h, int precision, const double* x, double* y) {
    return guard([&]() {
        if (!h || !x || !y) throw std::invalid_argument("null argument");
        if (precision != LINA_PREC_FP64 && precision != LINA_PREC_MIXED) throw std::invalid_argument("precision");
        h->m.spmv_host(precision, x, y);
        return 0;
    });
}

int lina_matrix_spmv_device(lina_matrix* h, const double* x, double* y) {
    return guard([&]() {
        if (!h || !x || !y) throw std::invalid_argument("null argument");
        h->m.spmv_dev(x, y);
        return 0;
    });
}

int lina_matrix_power(lina_matrix* h, double* v, int max_iter, double eps, double* eig) {
    return guard([&]() {
        if (!h || !v || !eig || max_iter < 1) throw std::invalid_argument("bad arguments");
        const int n = h->m.n();
        const double q = h->m.matrix_norm();
        hipStream_t st = h->m.stream();
        double *dv = nullptr, *tmp = nullptr, *part = nullptr, *scal = nullptr, *host = nullptr;
        auto release = [&]() {
            if (dv) (void)hipFree(dv);
            if (tmp) (void)hipFree(tmp);
            if (part) (void)hipFree(part);
            if (scal) (void)hipFree(scal);
            if (host) (void)hipHostFree(host);
        };
        try {
            lina::hip_check(hipMalloc((void**)&dv, n * sizeof(double)), "hipMalloc");
            lina::hip_check(hipMalloc((void**)&tmp, n * sizeof(double)), "hipMalloc");
            lina::hip_check(hipMalloc((void**)&part, ((size_t)n + 8192) * sizeof(double)), "hipMalloc");
            lina::hip_check(hipMalloc((void**)&scal, 16 * sizeof(double)), "hipMalloc");
            lina::hip_check(hipHostMalloc((void**)&host, 16 * sizeof(double)), "hipHostMalloc");
            lina::hip_check(hipMemcpyAsync(dv, v, n * sizeof(double), hipMemcpyHostToDevice, st), "h2d");
            const int it = lina::power_iteration(st, h->m, q, dv, tmp, part, scal, host, max_iter, eps, eig);
            lina::hip_check(hipMemcpyAsync(v, dv, n * sizeof(double), hipMemcpyDeviceToHost, st), "d2h");
            lina::hip_check(hipStreamSynchronize(st), "sync");
            release();
            return it;
        } catch (...) {
            release();
            throw;
        }
    });
}

double lina_matrix_time_spmv(lina_matrix* h, int precision, int reps) {
    double out = -1.0;
    int rc = guard([&]() {
        if (!h || reps < 1) throw std::invalid_argument("bad arguments");
        out = h->m.time_spmv(precision, reps);
        return 0;
    });
    return rc == 0 ? out : -1.0;
}

double lina_matrix_spmv_bytes(const lina_matrix* h, int precision) { return h ? h->m.bytes(precision) : 0.0; }

int lina_pcg_operator(lina_solver* s, lina_matrix* op, double* x, double* b, double tol, int max_iter, int stop,
                      lina_stats* st) {
    return guard([&]() {
        if (!s || !op || !x || !b) throw std::invalid_argument("null argument");
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        s->s.set_operator(&op->m);
        int it;
        try {
            it = s->s.pcg_host(x, b, tol, max_iter, stop, st);
        } catch (...) {
            s->s.set_operator(nullptr);
            throw;
        }
        s->s.set_operator(nullptr);
        return it;
    });
}

int lina_elasticity_adjoint(lina_solver* s, int n_nodes, const double* xy, int n_elem, const int* tri, double nu,
                            const double* elem_scale, int blocked, const double* u, const double* dLdu, double tol,
                            int max_iter, double* lambda, double* grad, lina_stats* st) {
    return guard([&]() {
        if (!s || !xy || !tri || !u || !dLdu || !lambda || !grad || n_nodes <= 0 || n_elem < 0)
            throw std::invalid_argument("null or empty argument");
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        const int n = 2 * n_nodes;
        if (n != s->s.n()) throw std::invalid_argument("2 * n_nodes differs from the system size");
        for (long k = 0; k < 3L * n_elem; k++)
            if (tri[k] < 0 || tri[k] >= n_nodes) throw std::invalid_argument("element node index out of range");
        if (!(nu > -1.0 && nu < 0.5)) throw std::invalid_argument("Poisson ratio outside (-1, 0.5)");
        // adjoint: K lambda = dL/du (K symmetric: the forward hierarchy)
        std::vector<double> lam((size_t)n, 0.0), rhs(dLdu, dLdu + n);
        int it = s->s.pcg_host(lam.data(), rhs.data(), tol, max_iter, LINA_STOP_RELRES, st);
        std::copy(lam.begin(), lam.end(), lambda);
        s->s.elasticity_gradient(n_nodes, xy, n_elem, tri, nu, elem_scale, blocked, u, lam.data(), grad);
        return it;
    });
}

}  // extern "C"

// ------------------------------------------------------- ILU(0) systems

#include "ilu.hpp"

struct lina_ilu {
    lina::IluSystem s;
};

extern "C" {

lina_ilu* lina_ilu_create(int n, const int* rowptr, const int* col, const double* val, int upload) {
    lina_ilu* h = nullptr;
    int rc = guard([&]() {
        Csr A = make_csr(n, rowptr, col, val);
        h = new lina_ilu();
        h->s.build(A, false);
        if (upload) {
            if (lina_device_count() < 1) throw lina::HipError("no HIP device available");
            h->s.upload();
        }
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

lina_ilu* lina_ilu_create_ordered(int n, const int* rowptr, const int* col, const double* val, int ordering,
                                  int upload) {
    lina_ilu* h = nullptr;
    int rc = guard([&]() {
        if (ordering != LINA_ILU_NATURAL && ordering != LINA_ILU_MULTICOLOUR)
            throw std::invalid_argument("unknown ILU ordering");
        Csr A = make_csr(n, rowptr, col, val);
        h = new lina_ilu();
        h->s.set_ordering(ordering);
        h->s.build(A, false);
        if (upload) {
            if (lina_device_count() < 1) throw lina::HipError("no HIP device available");
            h->s.upload();
        }
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

int lina_ilu_permutation(const lina_ilu* h, int* perm) {
    return guard([&]() {
        if (!h || !perm) throw std::invalid_argument("null argument");
        const std::vector<int>& p = h->s.permutation();
        const int n = h->s.n();
        for (int r = 0; r < n; r++) perm[r] = p.empty() ? r : p[r];
        return 0;
    });
}

void lina_ilu_destroy(lina_ilu* h) { delete h; }

int lina_ilu_ok(const lina_ilu* h) { return h->s.ilu_ok() ? 1 : 0; }

int lina_ilu_refactor(lina_ilu* h, int n, const int* rowptr, const int* col, const double* val) {
    return guard([&]() {
        if (!h) throw std::invalid_argument("null handle");
        Csr A = make_csr(n, rowptr, col, val);
        h->s.refactor(A);
        return 0;
    });
}

int lina_ilu_factor_nnz(const lina_ilu* h, char which) {
    lina::IluSystem& s = const_cast<lina_ilu*>(h)->s;  // downloading device factors changes no result
    return which == 'L' ? s.L().nnz() : s.U().nnz();
}

int lina_ilu_factor_copy(const lina_ilu* h, char which, int* rowptr, int* col, double* val) {
    lina::IluSystem& s = const_cast<lina_ilu*>(h)->s;
    const Csr& M = which == 'L' ? s.L() : s.U();
    std::memcpy(rowptr, M.rp.data(), (M.n + 1) * sizeof(int));
    std::memcpy(col, M.ci.data(), M.nnz() * sizeof(int));
    std::memcpy(val, M.v.data(), M.nnz() * sizeof(double));
    return 0;
}

lina_ilu* lina_ilu_create_factors(int n, const int* lrp, const int* lci, const double* lv, const int* urp,
                                  const int* uci, const double* uv) {
    lina_ilu* h = nullptr;
    int rc = guard([&]() {
        auto mk = [&](const int* rp, const int* ci, const double* v, bool diag_last) {
            Csr T;
            T.n = T.ncols = n;
            T.rp.assign(rp, rp + n + 1);
            if (T.rp[0] != 0) throw std::invalid_argument("rowptr[0] != 0");
            for (int i = 0; i < n; i++) {
                if (T.rp[i + 1] <= T.rp[i]) throw std::invalid_argument("every factor row needs its diagonal");
                const int d = diag_last ? T.rp[i + 1] - 1 : T.rp[i];
                if (ci[d] != i) throw std::invalid_argument("factor diagonal not in place");
            }
            T.ci.assign(ci, ci + T.rp[n]);
            T.v.assign(v, v + T.rp[n]);
            for (int c : T.ci)
                if (c < 0 || c >= n) throw std::invalid_argument("factor column out of range");
            return T;
        };
        Csr L = mk(lrp, lci, lv, true), U = mk(urp, uci, uv, false);
        h = new lina_ilu();
        h->s.set_factors(L, U, true);
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

int lina_ilu_apply_device(lina_ilu* h, const double* z, double* w) {
    return guard([&]() {
        h->s.apply(z, w);
        lina::hip_check(hipStreamSynchronize(h->s.stream()), "sync");
        return 0;
    });
}

int lina_ilu_apply(lina_ilu* h, const double* z, double* w) {
    return guard([&]() {
        h->s.apply_host(z, w);
        return 0;
    });
}

int lina_ilu_pcg(lina_ilu* h, double* x, double* b, double tol, int max_iter, int stop, double* residuum) {
    return guard([&]() { return h->s.pcg_host(x, b, tol, max_iter, stop, residuum); });
}

int lina_ilu_gmres(lina_ilu* h, double* x, const double* b, int m, double tol, int max_iter, double* eta) {
    return guard([&]() {
        if (m < 1) throw std::invalid_argument("restart length m must be >= 1");
        return h->s.gmres_host(x, b, m, tol, max_iter, eta);
    });
}

int lina_bicgstab(lina_solver* s, double* x, const double* b, double eps, int max_iter, double* res) {
    return guard([&]() {
        if (!s->s.uploaded()) throw std::invalid_argument("solver was created without device upload");
        return s->s.bicgstab_host(x, b, eps, max_iter, res);
    });
}

}  // extern "C"
