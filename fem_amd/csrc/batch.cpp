// Batched AMG-PCG: many independent systems solved together.
//
// The systems (specimens of different sizes) are set up one by one with the
// reference's AMG_setup (bit-identical per system, in parallel on the host)
// and joined level by level into one block-diagonal hierarchy: level l of the
// union holds every system's level l on the diagonal, P and R are
// block-diagonal, the coarsest level is a batch of dense inverses.  The
// V-cycle kernels then process all systems in each launch; multicolour
// sweeps run colour c of every system at once (rows of one colour are
// uncoupled inside a system and systems are uncoupled from each other).
//
// The Krylov iteration is pcg per system: dots are segmented reductions over
// row chunks that never straddle two systems, every step length is a
// per-system device scalar, and a system that has converged keeps its
// iterate while the others go on; the V-cycle's row kernels skip the row
// blocks of converged systems.  The iteration
// body contains no host synchronisation, so `graph_iters` iterations are
// captured once into a hipGraph and replayed; the host reads the number of
// still-active systems between replays.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/lina_hip.h"
#include "kernels.hpp"
#include "lina_internal.hpp"
#include "solver.hpp"

namespace lina {

namespace {

// block-diagonal union of per-system hierarchies of equal depth
HostHierarchy join(std::vector<HostHierarchy>& H) {
    const int ns = (int)H.size();
    const int depth = (int)H[0].lv.size();
    for (auto& h : H)
        if ((int)h.lv.size() != depth) throw std::invalid_argument("batch: systems reached different depths");
    HostHierarchy U;
    U.prm = H[0].prm;
    U.lv.resize(depth);
    auto diag_join = [&](auto get, std::vector<int> row_off, std::vector<int> col_off, int nrows, int ncols) {
        Csr C;
        C.n = nrows;
        C.ncols = ncols;
        C.rp.assign(nrows + 1, 0);
        size_t nnz = 0;
        for (int k = 0; k < ns; k++) nnz += get(k).ci.size();
        C.ci.reserve(nnz);
        C.v.reserve(nnz);
        for (int k = 0; k < ns; k++) {
            const Csr& A = get(k);
            const int colmax = col_off[k + 1] - col_off[k];
            for (int i = 0; i < A.n; i++) {
                for (int p = A.rp[i]; p < A.rp[i + 1]; p++) {
                    if (A.ci[p] >= colmax) {
                        // sparse_zero's (i, 0.0) seed of a row left empty
                        if (A.v[p] != 0.0) throw std::logic_error("batch: nonzero outside a system's block");
                        continue;
                    }
                    C.ci.push_back(A.ci[p] + col_off[k]);
                    C.v.push_back(A.v[p]);
                }
                C.rp[row_off[k] + i + 1] = (int)C.ci.size();
            }
        }
        return C;
    };
    for (int l = 0; l < depth; l++) {
        std::vector<int> fo(ns + 1, 0), co(ns + 1, 0);
        for (int k = 0; k < ns; k++) {
            fo[k + 1] = fo[k] + H[k].lv[l].A.n;
            co[k + 1] = co[k] + (l > 0 ? H[k].lv[l - 1].A.n : 0);
        }
        HostLevel& u = U.lv[l];
        u.A = diag_join([&](int k) -> const Csr& { return H[k].lv[l].A; }, fo, fo, fo[ns], fo[ns]);
        u.fine = fo[ns];
        if (l > 0) {
            u.P = diag_join([&](int k) -> const Csr& { return H[k].lv[l].P; }, fo, co, fo[ns], co[ns]);
            u.R = diag_join([&](int k) -> const Csr& { return H[k].lv[l].R; }, co, fo, co[ns], fo[ns]);
            u.coarse = co[ns];
            for (int k = 0; k < ns; k++) {
                u.labels.insert(u.labels.end(), H[k].lv[l].labels.begin(), H[k].lv[l].labels.end());
                u.levels.insert(u.levels.end(), H[k].lv[l].levels.begin(), H[k].lv[l].levels.end());
            }
        }
    }
    // finest-level extras and the coarsest level's dense blocks
    int off = 0;
    for (int k = 0; k < ns; k++) {
        U.signs.insert(U.signs.end(), H[k].signs.begin(), H[k].signs.end());
        for (size_t t = 0; t < H[k].iso_rows.size(); t++) {
            U.iso_rows.push_back(H[k].iso_rows[t] + off);
            U.iso_a.push_back(H[k].iso_a[t]);
        }
        off += H[k].lv[depth - 1].A.n;
    }
    CoarseLU& lu = U.lu;
    lu.blk_off.assign(1, 0);
    lu.mat_off.assign(1, 0);
    off = 0;
    for (int k = 0; k < ns; k++) {
        const CoarseLU& c = H[k].lu;
        for (int i : c.dense_idx) lu.dense_idx.push_back(i + off);
        for (size_t t = 0; t < c.trivial_idx.size(); t++) {
            lu.trivial_idx.push_back(c.trivial_idx[t] + off);
            lu.trivial_diag.push_back(c.trivial_diag[t]);
        }
        lu.dense.insert(lu.dense.end(), c.dense.begin(), c.dense.end());
        lu.blk_off.push_back((int)lu.dense_idx.size());
        lu.mat_off.push_back(lu.dense.size());
        off += c.n;
    }
    lu.n = off;
    return U;
}

}  // namespace

class BatchSolver {
  public:
    BatchSolver(int ns, const int* row_off, const Csr& A, const lina_amg_params& p, int chunk_rows)
        : ns_(ns), row_off_(row_off, row_off + ns + 1) {
        if (ns < 1) throw std::invalid_argument("batch: no systems");
        if (row_off_[0] != 0 || row_off_[ns] != A.n) throw std::invalid_argument("batch: row offsets");
        for (int k = 0; k < ns; k++)
            if (row_off_[k + 1] <= row_off_[k]) throw std::invalid_argument("batch: empty system");
        // split into systems with local column indices
        std::vector<Csr> sys(ns);
        for (int k = 0; k < ns; k++) {
            Csr& s = sys[k];
            const int r0 = row_off_[k], r1 = row_off_[k + 1];
            s.n = s.ncols = r1 - r0;
            s.rp.assign(s.n + 1, 0);
            for (int i = r0; i < r1; i++) {
                for (int q = A.rp[i]; q < A.rp[i + 1]; q++) {
                    const int c = A.ci[q];
                    if (c < r0 || c >= r1) throw std::invalid_argument("batch: entry couples two systems");
                    s.ci.push_back(c - r0);
                    s.v.push_back(A.v[q]);
                }
                s.rp[i - r0 + 1] = (int)s.ci.size();
            }
        }
        SetupParams sp;
        sp.dof = p.dof;
        sp.depth = p.depth;
        sp.smoothing_iterations = p.smoothing_iterations;
        sp.aggr_smoothing_iterations = p.aggr_smoothing_iterations;
        sp.coarsening = p.coarsening;
        sp.eps_minus = p.eps_minus;
        sp.eps_plus = p.eps_plus;
        sp.matrix_tol = p.matrix_tol;
        std::vector<HostHierarchy> H(ns);
        std::vector<std::string> err(ns);
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < ns; k++) {
            try {
                amg_setup(sys[k], sp, H[k]);
            } catch (const std::exception& e) {
                err[k] = e.what();
            }
        }
        for (int k = 0; k < ns; k++)
            if (!err[k].empty()) throw std::runtime_error("batch: setup of system " + std::to_string(k) + ": " + err[k]);
        HostHierarchy U = join(H);
        for (double sg : U.signs)
            if (sg < 0) throw std::invalid_argument("batch: systems need positive diagonals");
        // the systems' row offsets on every level (the union is block-diagonal)
        std::vector<std::vector<int>> seg(U.lv.size(), std::vector<int>(ns + 1, 0));
        for (size_t l = 0; l < U.lv.size(); l++)
            for (int k = 0; k < ns; k++) seg[l][k + 1] = seg[l][k] + H[k].lv[l].A.n;
        S_.set_hierarchy(std::move(U), p);
        S_.set_segments(std::move(seg));
        S_.upload();
        // row chunks inside systems
        std::vector<int> crow{0}, cinst, ichunk{0};
        for (int k = 0; k < ns; k++) {
            for (int r = row_off_[k]; r < row_off_[k + 1]; r += chunk_rows) {
                crow.push_back(std::min(r + chunk_rows, row_off_[k + 1]));
                cinst.push_back(k);
            }
            ichunk.push_back((int)cinst.size());
        }
        nchunk_ = (int)cinst.size();
        st_ = S_.stream();
        B_.reset(new DevBuffers());
        crow_ = up(crow);
        cinst_ = up(cinst);
        ichunk_ = up(ichunk);
        const int n = A.n;
        for (double** v : {&R_, &P_, &Q_, &Z_}) *v = alloc(n);
        part_ = alloc(std::max(nchunk_, 4096));
        for (double** v : {&s_, &sold_, &pq_, &rr_, &hb_, &active_, &iters_}) *v = alloc(ns_);
        S_.set_active(active_);  // the V-cycle skips the row blocks of converged systems
        count_ = alloc(1);
        hip_check(hipHostMalloc((void**)&host_, 2 * sizeof(double)), "hipHostMalloc");
        hip_check(hipStreamSynchronize(st_), "batch upload");
    }

    ~BatchSolver() {
        if (st_) (void)hipStreamSynchronize(st_);
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        if (host_) (void)hipHostFree(host_);
    }

    int n() const { return row_off_[ns_]; }
    Solver& solver() { return S_; }

    // X (in: guess, out: solution) and B device vectors of the union
    int solve_dev(double* X, const double* Bv, double tol, int max_iter, int graph_iters, lina_batch_stats* stats) {
        hipStream_t st = st_;
        const int n = this->n();
        hip_check(hipEventCreate(&e0_), "event");
        hip_check(hipEventCreate(&e1_), "event");
        hip_check(hipEventRecord(e0_, st), "event");
        // |B_k|, R = B - A X, rr, active, s = 1, P = 0
        dev::seg_dot(st, nchunk_, crow_, Bv, Bv, part_);
        dev::seg_sum(st, ns_, ichunk_, part_, hb_, nullptr);
        sqrt_inplace(hb_);
        DevLevel& top = S_.finest();
        dev::fill<double>(st, ns_, active_, 1.0);  // before any row kernel (they skip inactive systems)
        dev::fill<double>(st, ns_, iters_, 0.0);
        dev::spmv<double, double>(st, top.A, X, Bv, R_, dev::MODE_RES, nullptr, nullptr);
        dev::seg_dot(st, nchunk_, crow_, R_, R_, part_);
        dev::seg_sum(st, ns_, ichunk_, part_, rr_, nullptr);
        dev::bpcg_check(st, ns_, rr_, hb_, tol, max_iter, active_, iters_, 0);
        dev::fill<double>(st, ns_, s_, 1.0);
        dev::fill<double>(st, n, P_, 0.0);
        int launched = 0;
        int active = count_active();
        if (graph_iters > 0) {
            if (!exec_ || graph_iters != graph_len_ || X != gX_ || tol != gtol_ || max_iter != gmax_) {
                if (exec_) (void)hipGraphExecDestroy(exec_);
                if (graph_) (void)hipGraphDestroy(graph_);
                exec_ = nullptr;
                graph_ = nullptr;
                hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
                for (int k = 0; k < graph_iters; k++) body(X, tol, max_iter);
                hip_check(hipStreamEndCapture(st, &graph_), "end capture");
                hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "instantiate");
                graph_len_ = graph_iters;
                gX_ = X;
                gtol_ = tol;
                gmax_ = max_iter;
            }
            while (active > 0 && launched < max_iter) {
                hip_check(hipGraphLaunch(exec_, st), "graph launch");
                launched += graph_iters;
                active = count_active();
            }
        } else {
            while (active > 0 && launched < max_iter) {
                body(X, tol, max_iter);
                launched++;
                active = count_active();
            }
        }
        hip_check(hipEventRecord(e1_, st), "event");
        hip_check(hipEventSynchronize(e1_), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0_, e1_), "elapsed");
        (void)hipEventDestroy(e0_);
        (void)hipEventDestroy(e1_);
        if (stats) {
            stats->iterations_launched = launched;
            stats->active = active;
            stats->t_solve_ms = ms;
        }
        return launched;
    }

    void results(double* iters, double* relres) {
        std::vector<double> it(ns_), rr(ns_), hb(ns_);
        hip_check(hipMemcpyAsync(it.data(), iters_, ns_ * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipMemcpyAsync(rr.data(), rr_, ns_ * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipMemcpyAsync(hb.data(), hb_, ns_ * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        for (int k = 0; k < ns_; k++) {
            if (iters) iters[k] = it[k];
            if (relres) relres[k] = std::sqrt(rr[k]) / (hb[k] == 0 ? 1.0 : hb[k]);
        }
    }

  private:
    // one pcg iteration of every system, no host synchronisation
    void body(double* X, double tol, int max_iter) {
        hipStream_t st = st_;
        S_.vcycle_once(R_, Z_);
        dev::seg_dot(st, nchunk_, crow_, R_, Z_, part_);
        dev::seg_sum(st, ns_, ichunk_, part_, s_, sold_);
        dev::bpcg_p(st, nchunk_, crow_, cinst_, P_, Z_, s_, sold_, active_);
        DevLevel& top = S_.finest();
        dev::spmv<double, double>(st, top.A, P_, nullptr, Q_, dev::MODE_OUTER, nullptr, nullptr);
        dev::seg_dot(st, nchunk_, crow_, P_, Q_, part_);
        dev::seg_sum(st, ns_, ichunk_, part_, pq_, nullptr);
        dev::bpcg_xr(st, nchunk_, crow_, cinst_, X, R_, P_, Q_, s_, pq_, active_);
        dev::seg_dot(st, nchunk_, crow_, R_, R_, part_);
        dev::seg_sum(st, ns_, ichunk_, part_, rr_, nullptr);
        dev::bpcg_check(st, ns_, rr_, hb_, tol, max_iter, active_, iters_, 1);
    }

    int count_active() {
        int np = dev::dot_part<double>(st_, ns_, active_, active_, part_);
        dev::finalize(st_, np, part_, count_, 0, 0);
        hip_check(hipMemcpyAsync(host_, count_, sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        return (int)std::lround(host_[0]);
    }

    void sqrt_inplace(double* v) {
        // |B_k| from the sum of squares: sqrt through the finalize op on a
        // one-element reduction per system is avoided by a host round trip
        // (once per solve)
        std::vector<double> h(ns_);
        hip_check(hipMemcpyAsync(h.data(), v, ns_ * sizeof(double), hipMemcpyDeviceToHost, st_), "d2h");
        hip_check(hipStreamSynchronize(st_), "sync");
        for (double& x : h) x = std::sqrt(x);
        hip_check(hipMemcpyAsync(v, h.data(), ns_ * sizeof(double), hipMemcpyHostToDevice, st_), "h2d");
        hip_check(hipStreamSynchronize(st_), "sync");
    }

    template <typename T>
    T* up(const std::vector<T>& h) {
        void* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(h.size(), 1) * sizeof(T)), "hipMalloc");
        B_->ptrs.push_back(d);
        if (!h.empty()) hip_check(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st_), "h2d");
        return (T*)d;
    }
    double* alloc(size_t m) {
        void* d = nullptr;
        hip_check(hipMalloc(&d, std::max<size_t>(m, 1) * sizeof(double)), "hipMalloc");
        B_->ptrs.push_back(d);
        return (double*)d;
    }

    Solver S_;
    int ns_;
    std::vector<int> row_off_;
    int nchunk_ = 0;
    hipStream_t st_ = nullptr;
    std::unique_ptr<DevBuffers> B_;
    int *crow_ = nullptr, *cinst_ = nullptr, *ichunk_ = nullptr;
    double *R_ = nullptr, *P_ = nullptr, *Q_ = nullptr, *Z_ = nullptr, *part_ = nullptr;
    double *s_ = nullptr, *sold_ = nullptr, *pq_ = nullptr, *rr_ = nullptr, *hb_ = nullptr, *active_ = nullptr,
           *iters_ = nullptr, *count_ = nullptr;
    double* host_ = nullptr;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    int graph_len_ = 0, gmax_ = 0;
    double gtol_ = 0;
    double* gX_ = nullptr;
};

}  // namespace lina

// ------------------------------------------------------------------ C ABI

struct lina_batch {
    std::unique_ptr<lina::BatchSolver> b;
    double *dx = nullptr, *db = nullptr;
    ~lina_batch() {
        if (dx) (void)hipFree(dx);
        if (db) (void)hipFree(db);
    }
};

namespace {
template <typename F>
int batch_guard(F&& f) {
    try {
        return f();
    } catch (const lina::HipError& e) {
        lina::set_error("%s", e.what());
        return LINA_EHIP;
    } catch (const lina::NumericError& e) {
        lina::set_error("%s", e.what());
        return LINA_ENUMERIC;
    } catch (const std::invalid_argument& e) {
        lina::set_error("%s", e.what());
        return LINA_EARG;
    } catch (const std::exception& e) {
        lina::set_error("%s", e.what());
        return LINA_ESETUP;
    }
}
}  // namespace

extern "C" {

lina_batch* lina_batch_create(int nsys, const int* row_off, const int* rowptr, const int* col, const double* val,
                              const lina_amg_params* p) {
    lina_batch* h = nullptr;
    int rc = batch_guard([&]() {
        if (nsys < 1 || !row_off || !rowptr || !col || !val || !p) throw std::invalid_argument("null argument");
        lina::check_params(p);
        // the batch preconditions every CG iteration with one V-cycle from a
        // zero guess; other cycle settings would be silently ignored
        if (p->cycle != LINA_CYCLE_V || p->amg_max_iter != 1 || p->zero_init != 1)
            throw std::invalid_argument("a batch preconditions with one V-cycle from a zero guess: cycle must be "
                                        "LINA_CYCLE_V, amg_max_iter 1 and zero_init 1");
        const int n = row_off[nsys];
        lina::Csr A;
        A.n = A.ncols = n;
        A.rp.assign(rowptr, rowptr + n + 1);
        if (A.rp[0] != 0) throw std::invalid_argument("rowptr[0] != 0");
        for (int i = 0; i < n; i++)
            if (A.rp[i + 1] <= A.rp[i]) throw std::invalid_argument("every row needs at least one entry");
        A.ci.assign(col, col + A.rp[n]);
        A.v.assign(val, val + A.rp[n]);
        for (int i = 0; i < n; i++)
            for (int q = A.rp[i]; q < A.rp[i + 1]; q++)
                if (A.ci[q] < 0 || A.ci[q] >= n || (q > A.rp[i] && A.ci[q] <= A.ci[q - 1]))
                    throw std::invalid_argument("columns out of range or not sorted");
        h = new lina_batch();
        h->b.reset(new lina::BatchSolver(nsys, row_off, A, *p, 2048));
        lina::hip_check(hipMalloc((void**)&h->dx, (size_t)n * sizeof(double)), "hipMalloc");
        lina::hip_check(hipMalloc((void**)&h->db, (size_t)n * sizeof(double)), "hipMalloc");
        return 0;
    });
    if (rc < 0) {
        delete h;
        return nullptr;
    }
    return h;
}

void lina_batch_destroy(lina_batch* h) { delete h; }

int lina_batch_pcg_device(lina_batch* h, double* x, const double* b, double tol, int max_iter, int graph_iters,
                          double* iters, double* relres, lina_batch_stats* st) {
    return batch_guard([&]() {
        if (!h || !x || !b || max_iter < 0 || graph_iters < 0) throw std::invalid_argument("bad arguments");
        int it = h->b->solve_dev(x, b, tol, max_iter, graph_iters, st);
        h->b->results(iters, relres);
        return it;
    });
}

int lina_batch_pcg(lina_batch* h, double* x, const double* b, double tol, int max_iter, int graph_iters,
                   double* iters, double* relres, lina_batch_stats* st) {
    return batch_guard([&]() {
        if (!h || !x || !b || max_iter < 0 || graph_iters < 0) throw std::invalid_argument("bad arguments");
        const size_t bytes = (size_t)h->b->n() * sizeof(double);
        lina::hip_check(hipMemcpy(h->dx, x, bytes, hipMemcpyHostToDevice), "h2d");
        lina::hip_check(hipMemcpy(h->db, b, bytes, hipMemcpyHostToDevice), "h2d");
        int it = h->b->solve_dev(h->dx, h->db, tol, max_iter, graph_iters, st);
        h->b->results(iters, relres);
        lina::hip_check(hipMemcpy(x, h->dx, bytes, hipMemcpyDeviceToHost), "d2h");
        return it;
    });
}

int lina_batch_tuned(lina_batch* h, int* level, char* which, int* lpr64, int* lpr32, int cap) {
    return batch_guard([&]() {
        if (!h) throw std::invalid_argument("null handle");
        const auto& t = h->b->solver().tuned();
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

}  // extern "C"
